"""The bound behind the bounded pass's routing (host only).

A bounded pass (`pair_bounded_range`, `bounded_route64` in
thesiscode_amd/csrc/bve_device.h) splits each pair's denominator at a tile
centre c, D = E - delta with E = c0 - x.c and delta = x.(x_j - c), and sends
a (target group, tile) pair to float64 when t = |delta| / E may exceed 0.9
or when its remainder's rounding bound, which grows as t^2, exceeds the
error target.  It bounds t by 2q + q^2 with q = R / rho (R the tile's
radius about c, rho the target's distance from c): for unit vectors
E >= |x - c|^2 / 2 and |x.(x_j - c)| <= rho R + R^2 / 2, because
1 - |c|^2 is the mean of |x_j - c|^2 when c is the mean of the x_j.

These tests check that bound by brute force on random tiles of unit
vectors, check that it is needed in full (dropping the q^2 term fails), and
check the identity it rests on.
"""
import numpy as np


def _cap(rng, centre, angle, n):
    """n unit vectors within `angle` of the unit vector `centre`."""
    centre = centre / np.linalg.norm(centre)
    helper = np.array([1.0, 0.0, 0.0]) if abs(centre[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(centre, helper)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(centre, e1)
    cos_t = rng.uniform(np.cos(angle), 1.0, n)
    sin_t = np.sqrt(1.0 - cos_t ** 2)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    return (cos_t[:, None] * centre + sin_t[:, None] * np.cos(phi)[:, None] * e1
            + sin_t[:, None] * np.sin(phi)[:, None] * e2)


def _tiles(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        centre = rng.normal(size=3)
        angle = rng.uniform(0.01, 0.6)
        src = _cap(rng, centre, angle, int(rng.integers(8, 257)))
        c = src.mean(axis=0)
        o = src - c
        R = np.linalg.norm(o, axis=1).max()
        # targets anywhere on the sphere; keep those outside the tile sphere
        tgt = rng.normal(size=(512, 3))
        tgt /= np.linalg.norm(tgt, axis=1)[:, None]
        rho = np.linalg.norm(tgt - c, axis=1)
        keep = rho > R
        yield src, c, o, R, tgt[keep], rho[keep]


def _t_and_q(c, o, R, tgt, rho, c0=1.0):
    E = c0 - tgt @ c
    t = np.abs(tgt @ o.T).max(axis=1) / E
    return t, R / rho


def test_mean_of_unit_vectors_identity():
    """1 - |c|^2 = mean |x_j - c|^2 <= R^2 for c the mean of unit vectors."""
    for src, c, o, R, _, _ in _tiles(1, 200):
        lhs = 1.0 - c @ c
        rhs = np.mean(np.sum(o * o, axis=1))
        assert abs(lhs - rhs) <= 1e-14
        assert rhs <= R * R * (1 + 1e-12)


def test_t_within_2q_plus_q2():
    worst = 0.0
    for src, c, o, R, tgt, rho in _tiles(2, 300):
        if len(tgt) == 0:
            continue
        for c0 in (1.0, 1.0 + 1e-4):
            t, q = _t_and_q(c, o, R, tgt, rho, c0)
            bound = 2 * q + q * q
            assert np.all(t <= bound * (1 + 1e-12)), (t / bound).max()
            worst = max(worst, (t / bound).max())
    # the bound is close to tight on these tiles
    assert worst > 0.8


def test_q2_term_is_needed():
    """Tiles near the overlap limit exceed 2q alone: the q^2 term matters."""
    over = 0
    for src, c, o, R, tgt, rho in _tiles(3, 300):
        if len(tgt) == 0:
            continue
        t, q = _t_and_q(c, o, R, tgt, rho)
        over += int(np.sum(t > 2 * q))
    assert over > 0


def test_group_sphere_bound():
    """Per group, q = R / (|g - c| - r_g) bounds every target of the group."""
    rng = np.random.default_rng(4)
    for src, c, o, R, _, _ in _tiles(5, 100):
        g_dir = rng.normal(size=3)
        grp = _cap(rng, g_dir, rng.uniform(0.005, 0.2), 64)
        g = grp.mean(axis=0)
        r_g = np.linalg.norm(grp - g, axis=1).max()
        rho_c = np.linalg.norm(g - c) - r_g
        if not rho_c > R:
            continue
        q = R / rho_c
        E = 1.0 - grp @ c
        t = np.abs(grp @ o.T).max(axis=1) / E
        assert np.all(t <= (2 * q + q * q) * (1 + 1e-12))
