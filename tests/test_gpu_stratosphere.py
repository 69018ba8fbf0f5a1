"""GPU: the forced panel step (StratosphereRK4 with the Juckes forcing)
against the reference's own StratosphereRK4 output and the oracle.
Tolerance as for the unforced step: 1e-12 relative to the step's
displacement for positions, normwise for vorticity and KE."""
import numpy as np
import pytest

import oracle
from conftest import forced_manifest, load_golden, normwise_rel
from thesiscode_amd import _lib
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic

pytestmark = pytest.mark.gpu
TOL = 1e-12
FORCED = sorted(forced_manifest())


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", FORCED)
@pytest.mark.parametrize("precision", [_lib.BVE_PRECISION_FP64, _lib.BVE_PRECISION_FP64_FAST,
                                       _lib.BVE_PRECISION_MIXED_BOUNDED])
def test_forced_steps_match_reference(ctx, name, precision):
    """The float64 modes to 1e-12, the bounded mode to its error target."""
    bar = 1e-10 if precision == _lib.BVE_PRECISION_MIXED_BOUNDED else TOL
    c = forced_manifest()[name]
    g = load_golden(name)
    ctx.reset()
    ctx.set_omega(c["omega"])
    ctx.set_smooth(c["smooth"])
    ctx.set_precision(precision)
    try:
        ctx.load_panels(g["s0_active_x"], g["s0_active_relvort"], g["s0_active_area"],
                        g["s0_passive_x"])
        ctx.load_panel_absvort(g["active_absvort"])
        qx, qa = g["s0_particles_x"], g["s0_particles_absvort"]
        t0 = float(g["t0"][0])
        kes = []
        for n in range(c["nsteps"]):
            qx, qrel, tot, ke = ctx.stratosphere_rk4_step(t0 + n * c["dt"], c["dt"], qx, qa)
            kes.append(tot)
            if n == 0:
                assert normwise_rel(ke, g["ke_step1"]) < bar
        ax, av, px = ctx.store_panels()
    finally:
        ctx.set_precision(_lib.BVE_PRECISION_FP64)
        ctx.set_smooth(0.05)
    disp = np.abs(g["sn_active_x"] - g["s0_active_x"]).max()
    assert np.abs(ax - g["sn_active_x"]).max() < bar * disp
    assert np.abs(px - g["sn_passive_x"]).max() < bar * disp
    assert np.abs(qx - g["sn_particles_x"]).max() < bar * disp
    assert normwise_rel(av, g["sn_active_relvort"]) < bar
    assert normwise_rel(qrel, g["sn_particles_relvort"]) < bar
    assert np.allclose(kes, g["total_ke"], rtol=bar, atol=0)


def test_unforced_stratosphere_step_is_the_bve_step(ctx):
    """Past the forcing ramp (A = 0) the forced step equals the plain RK4 step."""
    g = gridmod.cubed_sphere(4)
    omega = 2 * np.pi
    z = ic.rh4_wave(g.active_x)
    aabs = z + 2 * omega * g.active_x[:, 2]
    qa = ic.rh4_wave(g.particles_x) + 2 * omega * g.particles_x[:, 2]
    ctx.reset()
    ctx.set_omega(omega)
    ctx.load_panels(g.active_x, z, g.active_area, g.passive_x)
    ctx.load_panel_absvort(aabs)
    fq, frel, ftot, _ = ctx.stratosphere_rk4_step(20.0, 0.01, g.particles_x, qa)
    fax, fav, fpx = ctx.store_panels()
    ctx.reset()
    ctx.load_panels(g.active_x, z, g.active_area, g.passive_x)
    rq, rrel, rtot, _ = ctx.rk4_step(0.01, g.particles_x, qa)
    rax, rav, rpx = ctx.store_panels()
    disp = np.abs(rax - g.active_x).max()
    for a, b in ((fax, rax), (fpx, rpx), (fq, rq)):
        assert np.abs(a - b).max() < 1e-13 * disp
    assert normwise_rel(fav, rav) < 1e-13 and normwise_rel(frel, rrel) < 1e-13
    assert abs(ftot - rtot) < 1e-13 * rtot


def test_forced_step_against_oracle_at_level5(ctx):
    """A larger grid than the golden cases, full forcing (t = 6)."""
    g = gridmod.cubed_sphere(5)
    omega = 2 * np.pi
    z = ic.gauss_vortex(g, lat0=np.pi / 3, lon0=0.5, width=0.2)[0]
    aabs = z + 2 * omega * g.active_x[:, 2]
    qa = np.zeros(g.n_particles) + 2 * omega * g.particles_x[:, 2]
    r = oracle.stratosphere_rk4_step(g.active_x, aabs, g.active_area, g.passive_x, g.particles_x,
                                     qa, 6.0, 0.02, omega, 0.05)
    ctx.reset()
    ctx.set_omega(omega)
    ctx.load_panels(g.active_x, z, g.active_area, g.passive_x)
    ctx.load_panel_absvort(aabs)
    qx, qrel, tot, _ = ctx.stratosphere_rk4_step(6.0, 0.02, g.particles_x, qa)
    ax, av, px = ctx.store_panels()
    disp = np.abs(r["active_x"] - g.active_x).max()
    assert np.abs(ax - r["active_x"]).max() < TOL * disp
    assert np.abs(px - r["passive_x"]).max() < TOL * disp
    assert np.abs(qx - r["particles_x"]).max() < TOL * disp
    assert normwise_rel(av, r["active_relvort"]) < TOL
    assert normwise_rel(qrel, r["particles_relvort"]) < TOL
    assert abs(tot - r["total_ke"]) < TOL * r["total_ke"]


def test_forced_step_needs_panel_absvort(ctx):
    g = gridmod.cubed_sphere(2)
    ctx.reset()
    ctx.load_panels(g.active_x, np.zeros(g.n_active), g.active_area, g.passive_x)
    with pytest.raises(_lib.BVEError, match="no panel absolute vorticity"):
        ctx.stratosphere_rk4_step(0.0, 0.01, g.particles_x)
