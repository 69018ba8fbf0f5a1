"""GPU ensembles: each member of a batched (persistent-launch, graph-captured)
ensemble equals the same problem run alone, bit for bit, in every
arithmetic mode; graph replays equal eager launches."""
import numpy as np
import pytest

import oracle
from conftest import normwise_rel
from thesiscode_amd import _lib
from thesiscode_amd.ensemble import autotune, load_members, make_members

pytestmark = pytest.mark.gpu
OMEGA = 2 * np.pi


def alone(d, precision, dt, n_steps, chunk):
    """The member's problem in its own context, with the member's launch
    geometry (256-source tiles, default targets per lane, `chunk`)."""
    g = d["grid"]
    with _lib.Context(0) as c:
        c.set_omega(OMEGA)
        c.set_precision(precision, 0.02)
        c.set_plan(0, 256, 0, 0)
        c.set_chunk(chunk)
        c.load_panels(g.active_x, d["zeta_active"], g.active_area, g.passive_x)
        c.load_particles(g.particles_x, d["absvort"])
        c.advance(dt, n_steps)
        qx, _ = c.store_particles()
        ax, av, px = c.store_panels()
    return dict(active_x=ax, active_relvort=av, passive_x=px, particles_x=qx)


@pytest.mark.parametrize("precision", [_lib.BVE_PRECISION_FP64, _lib.BVE_PRECISION_FP64_FAST,
                                       _lib.BVE_PRECISION_MIXED, _lib.BVE_PRECISION_MIXED_KAHAN])
@pytest.mark.parametrize("use_graph", [False, True])
def test_members_equal_unbatched_runs(precision, use_graph):
    members = make_members(5, levels=(2, 3, 4), seed=3)
    dt, n = 0.01, 2
    chunks = [512, 1024, 512, 768, 1536]   # several source tiles per chunk
    with _lib.Ensemble(len(members)) as ens:
        ens.set_params(OMEGA, 0.05, precision, 0.02)
        load_members(ens, members)
        for m, ch in enumerate(chunks):
            ens.set_member_plan(m, 0, ch)
        ens.advance(dt, n, use_graph)
        got = [ens.store_member(m) for m in range(len(members))]
    for m, d in enumerate(members):
        ref = alone(d, precision, dt, n, chunks[m])
        for k in ("active_x", "active_relvort", "passive_x", "particles_x"):
            assert np.array_equal(got[m][k], ref[k]), (m, k)


def test_graph_replays_match_eager():
    """The captured graph replayed three times, with eager work and host
    synchronisation in between, equals three eager advances each time."""
    members = make_members(4, levels=(3, 4), seed=1)
    dt = 0.01
    with _lib.Ensemble(len(members)) as eg, _lib.Ensemble(len(members)) as ee:
        for e in (eg, ee):
            e.set_params(OMEGA, 0.05, _lib.BVE_PRECISION_FP64, 0.02)
            load_members(e, members)
        for rep in range(3):
            eg.advance(dt, 2, use_graph=True)
            ee.advance(dt, 2, use_graph=False)
            a = [eg.store_member(m) for m in range(len(members))]
            b = [ee.store_member(m) for m in range(len(members))]
            for m in range(len(members)):
                for k in ("active_x", "particles_x", "passive_x", "active_relvort"):
                    assert np.array_equal(a[m][k], b[m][k]), (rep, m, k)


@pytest.mark.parametrize("precision,bar", [(_lib.BVE_PRECISION_FP64, 1e-12),
                                           (_lib.BVE_PRECISION_FP64_FAST, 1e-12),
                                           (_lib.BVE_PRECISION_MIXED, 1e-5),
                                           (_lib.BVE_PRECISION_MIXED_KAHAN, 1e-5)])
def test_member_plans_keep_parity(precision, bar):
    """Explicit plans with a different targets-per-lane value per member (so
    one pair launch per value, each a kernel compiled for it) give the
    oracle's step to the mode's bar."""
    members = make_members(3, levels=(3,), seed=2)
    plans = [(1, 512), (2, 1024), (4, 256)]
    dt = 0.01
    with _lib.Ensemble(len(members)) as ens:
        ens.set_params(OMEGA, 0.05, precision, 0.02)
        load_members(ens, members)
        for m, (T, ch) in enumerate(plans):
            ens.set_member_plan(m, T, ch)
        ens.advance(dt, 1, use_graph=True)
        got = [ens.store_member(m) for m in range(len(members))]
    for m, d in enumerate(members):
        g = d["grid"]
        r = oracle.rk4_step(g.active_x, d["zeta_active"], g.active_area, g.passive_x,
                            g.particles_x, d["absvort"], dt, OMEGA, 0.05)
        disp = np.abs(r["active_x"] - g.active_x).max()
        assert np.abs(got[m]["active_x"] - r["active_x"]).max() < bar * disp
        assert np.abs(got[m]["particles_x"] - r["particles_x"]).max() < bar * disp


def test_autotuner_returns_a_supported_plan():
    """Supported plans, one targets-per-lane value for every level (the
    members then share one pair launch), and per level the fastest chunk
    tried at that value."""
    table = autotune([2, 3], members_per_level=2)
    for L, t in table.items():
        assert t["targets_per_lane"] in (1, 2, 4) and t["chunk"] % 256 == 0
        assert len(t["tried"]) >= 3
        same_t = [r for r in t["tried"] if r[0] == t["targets_per_lane"]]
        assert t["ms_per_stage"] == min(r[2] for r in same_t)
    assert len({t["targets_per_lane"] for t in table.values()}) == 1
