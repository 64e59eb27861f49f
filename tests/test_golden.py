"""The committed golden fixtures (tests/golden/make_golden.py) still hold: the
scene generator reproduces the fixture's input and the oracle its outputs."""
import os

import numpy as np

import oracle
from libclsph_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_dam4096_fixture_reproduced():
    g = np.load(os.path.join(GOLDEN, "dam4096_step0.npz"))
    sc = scenes.dam_break(4096, seed=1234)
    np.testing.assert_array_equal(sc.particles["position"][:, :3], g["position0"])
    r = oracle.stages(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices, sc.indices)
    np.testing.assert_array_equal(r.params["min_point"][:3], g["min_point"])
    np.testing.assert_array_equal(r.params["max_point"][:3], g["max_point"])
    assert int(r.params["grid_cell_count"]) == int(g["grid_cell_count"])
    np.testing.assert_array_equal(r.hashed["grid_index"], g["keys"])
    np.testing.assert_array_equal(np.argsort(r.hashed["grid_index"], kind="stable"), g["order"])
    np.testing.assert_array_equal(r.table, g["table"])
    np.testing.assert_array_equal(r.density["density"], g["density"])
    np.testing.assert_array_equal(r.forces["acceleration"][:, :3], g["acceleration"])
    np.testing.assert_array_equal(r.advected["position"][:, :3], g["position1"])


def test_series_fixture_is_physical():
    g = np.load(os.path.join(GOLDEN, "dam16k_series.npz"))
    rho, ke = g["mean_density"], g["kinetic_energy"]
    assert len(rho) == 1000 and np.all(np.isfinite(rho)) and np.all(np.isfinite(ke))
    # early on the block free-falls: KE ~ 0.5 M (g t)^2 while nothing touches it
    sc = scenes.dam_break(16384)
    M = float(sc.params["particle_mass"]) * 16384
    dt = float(sc.params["time_delta"])
    t = 20 * dt
    assert abs(ke[19] / (0.5 * M * (9.81 * t) ** 2) - 1) < 0.1
    assert 900 < rho.mean() < 1100
