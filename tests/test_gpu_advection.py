"""GPU: the analytic-wind steps (AdvectionRK4, DivergentAdvectionRK4)
against the reference's own output and the oracle.  Positions to 1e-12
relative to the step's displacement (pointwise trigonometry in another
order and another libm)."""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_golden
from thesiscode_amd import _lib
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic

pytestmark = pytest.mark.gpu
TOL = 1e-12


def advection_manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["advection_cases"]


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", sorted(advection_manifest()))
def test_advection_steps_match_reference(ctx, name):
    c = advection_manifest()[name]
    g = load_golden(name)
    ctx.reset()
    ctx.load_panels(g["s0_active_x"], g["s0_active_relvort"], g["s0_active_area"],
                    g["s0_passive_x"])
    qx = g["s0_particles_x"]
    t0 = float(g["t0"][0])
    for n in range(c["nsteps"]):
        qx = ctx.advection_rk4_step(t0 + n * c["dt"], c["dt"], qx, c["divergent"])
    ax, av, px = ctx.store_panels()
    disp = np.abs(g["sn_active_x"] - g["s0_active_x"]).max()
    assert np.abs(ax - g["sn_active_x"]).max() < TOL * disp
    assert np.abs(px - g["sn_passive_x"]).max() < TOL * disp
    assert np.abs(qx - g["sn_particles_x"]).max() < TOL * disp
    assert np.array_equal(av, g["s0_active_relvort"])     # vorticity untouched


@pytest.mark.parametrize("divergent", [False, True])
def test_advection_against_oracle_then_bve_step(ctx, divergent):
    """A level-5 grid: advection steps, then a BVE step on the moved panels
    (the source records follow the advected positions)."""
    g = gridmod.cubed_sphere(5)
    z = ic.rh4_wave(g.active_x)
    ctx.reset()
    ctx.load_panels(g.active_x, z, g.active_area, g.passive_x)
    qx = g.particles_x
    ax_ref, px_ref, qx_ref = g.active_x, g.passive_x, g.particles_x
    for n in range(3):
        t = 0.4 + n * 0.03
        qx = ctx.advection_rk4_step(t, 0.03, qx, divergent)
        ax_ref = oracle.advection_rk4_step(ax_ref, t, 0.03, divergent)
        px_ref = oracle.advection_rk4_step(px_ref, t, 0.03, divergent)
        qx_ref = oracle.advection_rk4_step(qx_ref, t, 0.03, divergent)
    ax, av, px = ctx.store_panels()
    disp = np.abs(ax_ref - g.active_x).max()
    assert np.abs(ax - ax_ref).max() < TOL * disp
    assert np.abs(px - px_ref).max() < TOL * disp
    assert np.abs(qx - qx_ref).max() < TOL * disp
    qx2, _, _, _ = ctx.rk4_step(0.01, qx)
    ax2, _, _ = ctx.store_panels()
    r = oracle.rk4_step(ax, z, g.active_area, px, qx, np.zeros(len(qx)), 0.01, 2 * np.pi, 0.05)
    d2 = np.abs(r["active_x"] - ax).max()
    assert np.abs(ax2 - r["active_x"]).max() < TOL * d2
    assert np.abs(qx2 - r["particles_x"]).max() < TOL * d2


def test_advection_argument_errors(ctx):
    g = gridmod.cubed_sphere(2)
    ctx.reset()
    with pytest.raises(_lib.BVEError, match="no panels loaded"):
        ctx.advection_rk4_step(0.0, 0.01, g.particles_x)
    ctx.load_panels(g.active_x, np.zeros(g.n_active), g.active_area, g.passive_x)
    with pytest.raises(_lib.BVEError, match="bad arguments"):
        ctx._check(ctx._L.bve_advection_rk4_step(ctx._h, 0.0, 0.01, 2, 0, None),
                   "bve_advection_rk4_step")
