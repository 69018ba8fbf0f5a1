"""Host-side inputs: the cubed-sphere generator against the reference's own
grid (golden vectors), and the host bookkeeping of the drop-in."""
import numpy as np
import pytest

from conftest import load_golden
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic
from thesiscode_amd.panel_velocity import gather_panels, load_balance, panels_from_grid


@pytest.mark.parametrize("level", range(0, 6))
def test_counts_and_total_area(level):
    g = gridmod.cubed_sphere(level)
    assert (g.n_active, g.n_particles, g.n_passive) == gridmod.counts(level)
    assert abs(g.active_area.sum() - 4 * np.pi) < 1e-12
    assert np.abs(np.linalg.norm(g.active_x, axis=1) - 1).max() < 3e-16
    # Euler: V - E + F = 2 with E = 2F for quads
    assert g.n_particles - 2 * g.n_active + g.n_active == 2


@pytest.mark.parametrize("name,level", [("gauss_L3_2steps_smooth010", 3), ("rh4ref_L4_1step", 4),
                                        ("rh4_L5_1step", 5)])
def test_grid_matches_reference_grid(name, level):
    """Active and passive panel centres equal the reference's bit for bit in
    GatherPanels order, areas to 1e-16, particles as a set."""
    g = load_golden(name)
    cs = gridmod.cubed_sphere(level)
    assert np.array_equal(cs.active_x, g["s0_active_x"])
    assert np.array_equal(cs.passive_x, g["s0_passive_x"])
    assert np.abs(cs.active_area - g["s0_active_area"]).max() < 1e-16
    a = cs.particles_x[np.lexsort(cs.particles_x.T)]
    b = g["s0_particles_x"][np.lexsort(g["s0_particles_x"].T)]
    assert np.array_equal(a, b)


def test_reference_rh4_matches_reference_vorticity():
    g = load_golden("rh4ref_L4_1step")
    z = ic.rh4_reference(g["s0_active_x"])
    assert np.abs(z - g["s0_active_relvort"]).max() < 1e-13


def test_benchmark_rh4_matches_harness_vorticity():
    g = load_golden("rh4_L4_12steps")
    z = ic.rh4_wave(g["s0_active_x"])
    assert np.abs(z - g["s0_active_relvort"]).max() < 1e-13
    assert np.abs(g["s0_particles_absvort"] - (ic.rh4_wave(g["s0_particles_x"]) + 2 * 2 * np.pi * g["s0_particles_x"][:, 2])).max() < 1e-12


def test_gauss_has_zero_mean():
    g = gridmod.cubed_sphere(4)
    za, zp, zq = ic.gauss_vortex(g)
    assert abs(np.sum(za * g.active_area)) < 1e-12
    za2, _, _ = ic.two_vortices(g)
    assert abs(np.sum(za2 * g.active_area)) < 1e-12


def test_load_balance_is_the_reference_partition():
    parts = load_balance(10, 3)
    assert parts == [(0, 3), (3, 6), (6, 10)]
    assert load_balance(5, 1) == [(0, 5)]


def test_gather_order():
    g = gridmod.cubed_sphere(2)
    p = panels_from_grid(g, ic.rh4_wave(g.active_x))
    a, q = gather_panels(p)
    assert np.array_equal(p.x[a], g.active_x) and np.array_equal(p.x[q], g.passive_x)
