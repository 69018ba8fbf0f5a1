"""Host logic of the ensemble autotuner (CPU only): the choice of one
targets-per-lane value shared by every level, and the fastest chunk per
level at that value."""
import pytest

from thesiscode_amd.ensemble import candidate_plans, common_plans


def test_common_value_minimises_the_weighted_total():
    # T = 4 is fastest on level 6 alone, T = 1 on levels 4 and 5; with level
    # 6 weighted by its cost the common choice is the one with the smallest
    # weighted sum of per-level bests, not any level's own best
    tried = {
        4: [(1, 256, 0.10), (2, 256, 0.12), (4, 256, 0.20)],
        5: [(1, 256, 0.30), (1, 512, 0.28), (2, 256, 0.33), (4, 256, 0.50)],
        6: [(1, 512, 2.00), (1, 1024, 1.90), (2, 512, 1.80), (4, 512, 1.75)],
    }
    got = common_plans(tried)
    # sums: T=1 0.10 + 0.28 + 1.90 = 2.28; T=2 0.12 + 0.33 + 1.80 = 2.25;
    # T=4 0.20 + 0.50 + 1.75 = 2.45
    assert {L: p[0] for L, p in got.items()} == {4: 2, 5: 2, 6: 2}
    assert got[6] == (2, 512, 1.80)
    # weights shift the choice: level 4 counted ten times favours T = 1
    got_w = common_plans(tried, weights={4: 10.0, 5: 1.0, 6: 1.0})
    assert {p[0] for p in got_w.values()} == {1}
    assert got_w[5] == (1, 512, 0.28) and got_w[6] == (1, 1024, 1.90)


def test_values_not_tried_on_every_level_are_skipped():
    tried = {4: [(1, 256, 1.0), (4, 256, 0.1)], 5: [(1, 256, 1.0)]}
    assert common_plans(tried) == {4: (1, 256, 1.0), 5: (1, 256, 1.0)}
    with pytest.raises(ValueError):
        common_plans({4: [(4, 256, 0.1)], 5: [(1, 256, 1.0)]})


def test_candidate_plans_cover_every_value():
    for level in (2, 4, 6):
        plans = candidate_plans(level)
        assert {T for T, _ in plans} == {1, 2, 4}
        assert all(c % 256 == 0 for _, c in plans)
