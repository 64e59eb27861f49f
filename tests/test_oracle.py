"""Pins of the CPU oracle, each against something independent of its code:
bit-loop Morton codes, numpy's stable argsort / searchsorted, the textbook
smoothing kernels (Mueller et al. 2003) through finite differences and their
normalisation, momentum conservation, lattice symmetry, and closed-form
free fall / plane collisions. The reference ships no tests or golden vectors
of its own, so these are the pins."""
import math

import numpy as np
import pytest

import oracle
from libclsph_amd import layout, native, scenes


def bit_morton(x, y, z):
    k = 0
    for b in range(10):
        k |= ((x >> b) & 1) << (3 * b)
        k |= ((y >> b) & 1) << (3 * b + 1)
        k |= ((z >> b) & 1) << (3 * b + 2)
    return k


def test_morton_bit_placement_and_roundtrip():
    for v in range(1024):
        assert oracle.morton(v, 0, 0) == bit_morton(v, 0, 0)
        assert oracle.morton(0, v, 0) == bit_morton(0, v, 0)
        assert oracle.morton(0, 0, v) == bit_morton(0, 0, v)
        assert oracle.morton_decode(bit_morton(v, 0, 0)) == (v, 0, 0)
        assert oracle.morton_decode(bit_morton(0, 0, v)) == (0, 0, v)
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        k = oracle.morton(int(x), int(y), int(z))
        assert k == bit_morton(int(x), int(y), int(z))
        assert oracle.morton_decode(k) == (x, y, z)
    # monotone per axis (what the sparse cell table relies on)
    assert all(oracle.morton(v, 5, 9) < oracle.morton(v + 1, 5, 9) for v in range(1023))


def random_cloud(n, seed, lo=0.1, hi=0.9, h=0.03):
    rng = np.random.default_rng(seed)
    parts = layout.particles(n)
    parts["position"][:, :3] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    p = scenes.fluid_params(n, h, 1e-3)
    return parts, p


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 1), (500, 2), (3000, 3)])
def test_grid_params_invariants(n, seed):
    parts, p = random_cloud(n, seed)
    assert oracle.grid_params(parts, p) == 0
    h = float(p["h"])
    L = 2 * h
    pos = parts["position"][:, :3].astype(np.float64)
    # padding of two cells (4h) on every side
    np.testing.assert_allclose(p["min_point"][:3], pos.min(axis=0) - 2 * L, rtol=0, atol=1e-6)
    np.testing.assert_allclose(p["max_point"][:3], pos.max(axis=0) + 2 * L, rtol=0, atol=1e-6)
    gs = np.array([p["grid_size_x"], p["grid_size_y"], p["grid_size_z"]])
    ext = (p["max_point"][:3].astype(np.float64) - p["min_point"][:3])
    np.testing.assert_array_equal(gs, np.floor(ext / L + 1e-9).astype(int))
    assert int(p["grid_cell_count"]) == bit_morton(*[int(v) for v in gs])
    hashed = oracle.locate_in_grid(parts, p)
    for i in range(n):
        c = oracle.morton_decode(int(hashed["grid_index"][i]))
        for a in range(3):
            # each particle lies in its cell and the 27-cell stencil stays inside the grid
            lo = float(p["min_point"][a]) + c[a] * L
            assert lo - 1e-6 <= pos[i, a] < lo + L + 1e-6
            assert 1 <= c[a] <= gs[a] - 2


def np32_grid(pos, h):
    """The grid of sph_simulation.cpp:201-252 and the cell ids of
    grid.cl:43-67 restated in numpy float32 arithmetic (every operation
    rounded to float32 as the reference's float variables are): bounding box
    from float(INT_MAX) / float(INT_MIN) start values, two cells of padding,
    truncated dimensions, the Morton count and each particle's truncated
    cell coordinates."""
    f = np.float32
    L = f(h) * f(2)
    mn = np.full(3, f(2147483647), dtype=f)
    mx = np.full(3, f(-2147483648), dtype=f)
    mn = np.minimum(mn, pos.min(axis=0))
    mx = np.maximum(mx, pos.max(axis=0))
    mn = mn - L * f(2)
    mx = mx + L * f(2)
    gs = ((mx - mn) / L).astype(np.uint32)
    cell = ((pos - mn) / L).astype(np.uint32)
    keys = np.array([bit_morton(int(x), int(y), int(z)) for x, y, z in cell], dtype=np.uint32)
    return mn, mx, gs, bit_morton(*[int(v) for v in gs]), keys


def boundary_cloud(seed, h=0.03):
    """Particles exactly on cell faces of the padded grid, one float32 ulp
    either side of them, and random ones, all inside the box an anchor
    particle and a far corner particle fix."""
    rng = np.random.default_rng(seed)
    f = np.float32
    L = f(h) * f(2)
    anchor = rng.uniform(0.1, 0.4, 3).astype(f)
    mn = anchor - L * f(2)
    pts = [anchor, anchor + f(0.3)]
    for _ in range(300):
        k = rng.integers(3, 12, 3)
        b = (mn + k.astype(f) * L).astype(f)
        side = rng.integers(-1, 2, 3)
        b = np.where(side < 0, np.nextafter(b, f(-1)), np.where(side > 0, np.nextafter(b, f(2)), b))
        pts.append(b.astype(f))
    pts += list(rng.uniform(anchor, anchor + f(0.3), (200, 3)).astype(f))
    parts = layout.particles(len(pts))
    parts["position"][:, :3] = np.array(pts, dtype=f)
    p = scenes.fluid_params(len(pts), h, 1e-3)
    return parts, p


@pytest.mark.parametrize("kind,seed", [("random", 11), ("random", 12), ("boundary", 13),
                                       ("boundary", 14), ("boundary", 15)])
def test_grid_and_cell_ids_bit_exact_against_float32_restatement(kind, seed):
    if kind == "random":
        parts, p = random_cloud(2000, seed)
    else:
        parts, p = boundary_cloud(seed)
    assert oracle.grid_params(parts, p) == 0
    mn, mx, gs, count, keys = np32_grid(parts["position"][:, :3].copy(), p["h"])
    np.testing.assert_array_equal(p["min_point"][:3].view(np.uint32), mn.view(np.uint32))
    np.testing.assert_array_equal(p["max_point"][:3].view(np.uint32), mx.view(np.uint32))
    np.testing.assert_array_equal([p["grid_size_x"], p["grid_size_y"], p["grid_size_z"]], gs)
    assert int(p["grid_cell_count"]) == count
    hashed = oracle.locate_in_grid(parts, p)
    np.testing.assert_array_equal(hashed["grid_index"], keys)


def test_grid_dimension_assert():
    parts, p = random_cloud(2, 0)
    parts["position"][1, 0] = 100.0  # 100 m away with h = 3 cm: > 1024 cells
    assert oracle.grid_params(parts, p) == -1


@pytest.mark.parametrize("n", [1, 2, 127, 128, 129, 1000, 5003])
def test_sort_is_stable_sort(n):
    rng = np.random.default_rng(n)
    parts = layout.particles(n)
    parts["grid_index"] = rng.integers(0, 50, n).astype(np.uint32) * 7919
    parts["density"] = np.arange(n, dtype=np.float32)  # tags the input slot
    out = oracle.sort_particles(parts)
    order = np.argsort(parts["grid_index"], kind="stable")
    np.testing.assert_array_equal(out["grid_index"], parts["grid_index"][order])
    np.testing.assert_array_equal(out["density"], parts["density"][order])


def test_sort_all_digit_positions():
    # keys spanning all four 8-bit digits
    rng = np.random.default_rng(9)
    parts = layout.particles(4000)
    parts["grid_index"] = rng.integers(0, 1 << 30, 4000).astype(np.uint32)
    parts["grid_index"][1000:1800] = parts["grid_index"][::5]  # duplicates
    parts["pressure"] = np.arange(4000, dtype=np.float32)
    out = oracle.sort_particles(parts)
    order = np.argsort(parts["grid_index"], kind="stable")
    np.testing.assert_array_equal(out["pressure"], parts["pressure"][order])


def test_cell_table_is_lower_bound():
    rng = np.random.default_rng(4)
    parts = layout.particles(3000)
    parts["grid_index"] = np.sort(rng.integers(3, 900, 3000)).astype(np.uint32)
    G = 1000
    t = oracle.cell_table(parts, G)
    np.testing.assert_array_equal(t, np.searchsorted(parts["grid_index"], np.arange(G), "left"))


# ---------------------------------------------------------------- physics

def one_step_inputs(parts, p):
    assert oracle.grid_params(parts, p) == 0
    s = oracle.sort_particles(oracle.locate_in_grid(parts, p))
    return s, oracle.cell_table(s, int(p["grid_cell_count"]))


def lattice(nside, spacing, origin=0.3):
    g = np.arange(nside) * spacing + origin
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    parts = layout.particles(nside ** 3)
    parts["position"][:, 0] = X.ravel()
    parts["position"][:, 1] = Y.ravel()
    parts["position"][:, 2] = Z.ravel()
    return parts


def poly6(r, h):
    return 315.0 / (64.0 * math.pi * h ** 9) * np.where(r < h, (h * h - r * r) ** 3, 0.0)


def test_poly6_normalisation_fine_lattice():
    # spacing h/8: the interior density approaches rho0 = m / s^3 (int W dV = 1)
    h = 0.04
    s = h / 8
    parts = lattice(21, s)
    n = len(parts)
    p = scenes.fluid_params(n, h, 1000.0 * s ** 3)
    t = native.kernel_terms(h)
    srt, table = one_step_inputs(parts, p)
    d = oracle.density_pressure(srt, p, t, table)
    centre = np.argmin(np.linalg.norm(d["position"][:, :3] - (0.3 + 10 * s), axis=1))
    assert abs(d["density"][centre] / 1000.0 - 1.0) < 3e-3


def test_density_matches_textbook_sum_on_lattice():
    # spacing 0.5 h: interior value of sum_j m W(r_j) from the textbook poly6
    h = 0.04
    s = 0.5 * h
    parts = lattice(9, s)
    n = len(parts)
    m = 1000.0 * s ** 3
    p = scenes.fluid_params(n, h, m)
    t = native.kernel_terms(h)
    srt, table = one_step_inputs(parts, p)
    d = oracle.density_pressure(srt, p, t, table)
    X = srt["position"][:, :3].astype(np.float64)
    for i in range(0, n, 37):
        r = np.linalg.norm(X - X[i], axis=1)
        want = m * poly6(r, h).sum()
        assert abs(d["density"][i] - want) <= 2e-6 * want
    # Tait pressure: K ((rho/rho0)^7 - 1)
    K = float(p["K"])
    np.testing.assert_allclose(d["pressure"], K * ((d["density"] / 1000.0) ** 7 - 1), rtol=2e-5,
                               atol=1e-6)
    # the well-known interior value at spacing 0.5 h: 1.00978 rho0
    centre = np.argmin(np.linalg.norm(X - (0.3 + 4 * s), axis=1))
    assert abs(d["density"][centre] / 1000.0 - 1.00978) < 5e-5


def two_particles(r_vec, h, dens=(1000.0, 1000.0), pres=(0.0, 0.0), vel=((0, 0, 0), (0, 0, 0)),
                  mu=0.0, sigma=0.0, tau=1e30):
    parts = layout.particles(2)
    parts["position"][0, :3] = (0.5, 0.5, 0.5)
    parts["position"][1, :3] = np.array((0.5, 0.5, 0.5)) - np.asarray(r_vec)
    m = 1e-3
    p = scenes.fluid_params(2, h, m, viscosity=mu, surface_tension=sigma, threshold=tau,
                            gravity=0.0)
    srt, table = one_step_inputs(parts, p)
    # put the chosen fields on the sorted particles (forces reads them as given)
    idx0 = int(np.argmin(np.linalg.norm(srt["position"][:, :3] - 0.5, axis=1)))
    for k, who in ((idx0, 0), (1 - idx0, 1)):
        srt["density"][k] = dens[who]
        srt["pressure"][k] = pres[who]
        srt["velocity"][k, :3] = vel[who]
    t = native.kernel_terms(h)
    out = oracle.forces(srt, p, t, table)
    return out["acceleration"][idx0, :3].astype(np.float64), out["acceleration"][1 - idx0, :3], m


def fd_grad(f, x, eps=1e-7):
    g = np.zeros(3)
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        g[a] = (f(x + e) - f(x - e)) / (2 * eps)
    return g


def fd_lap(f, x, eps=1e-5):
    return sum((f(x + e) - 2 * f(x) + f(x - e)) / eps ** 2 for e in np.eye(3) * eps)


def test_pressure_is_spiky_gradient():
    h = 0.04
    rv = np.array([0.011, -0.007, 0.013])
    rho, pr = (1010.0, 990.0), (2.0, 0.5)
    a_i, a_j, m = two_particles(rv, h, dens=rho, pres=pr)
    w_spiky = lambda x: 15.0 / (math.pi * h ** 6) * (h - np.linalg.norm(x)) ** 3
    want = -(pr[1] / rho[1] ** 2 + pr[0] / rho[0] ** 2) * m * fd_grad(w_spiky, rv)
    np.testing.assert_allclose(a_i, want, rtol=1e-4)
    # equal and opposite momentum change
    np.testing.assert_allclose(a_i * m, -a_j * m, rtol=1e-5)


def test_viscosity_is_laplacian_of_viscosity_kernel():
    h = 0.04
    rv = np.array([0.012, 0.005, -0.009])
    vel = ((0.0, 0.0, 0.0), (0.3, -0.2, 0.1))
    mu = 0.5
    rho = (1000.0, 1020.0)
    a_i, a_j, m = two_particles(rv, h, dens=rho, vel=vel, mu=mu)

    def w_visc(x):
        r = np.linalg.norm(x)
        return 15.0 / (2 * math.pi * h ** 3) * (-r ** 3 / (2 * h ** 3) + r * r / h ** 2
                                                 + h / (2 * r) - 1)

    lap = fd_lap(w_visc, rv)
    want = mu * (np.array(vel[1]) - np.array(vel[0])) * (m / rho[1]) * lap / rho[0]
    np.testing.assert_allclose(a_i, want, rtol=2e-4)
    # a_i = mu (v_j - v_i) m lapW / (rho_i rho_j) = -a_j: equal masses, equal and opposite
    np.testing.assert_allclose(a_i, -a_j, rtol=1e-5)


def test_surface_tension_is_colour_field():
    h = 0.04
    rv = np.array([0.010, 0.014, -0.006])
    rho = (1000.0, 980.0)
    sigma = 0.7
    a_i, _, m = two_particles(rv, h, dens=rho, sigma=sigma, tau=0.0)
    w = lambda x: 315.0 / (64 * math.pi * h ** 9) * (h * h - x @ x) ** 3
    normal = (m / rho[1]) * fd_grad(w, rv)  # self gradient is zero
    kappa = (m / rho[0]) * fd_lap(w, np.zeros(3) + 1e-9) + (m / rho[1]) * fd_lap(w, rv)
    want = -sigma * kappa * normal / np.linalg.norm(normal) / rho[0]
    np.testing.assert_allclose(a_i, want, rtol=2e-3)


def test_momentum_conservation_random_cloud():
    sc = scenes.dam_break(4096, seed=3)
    parts = sc.particles.copy()
    rng = np.random.default_rng(1)
    parts["velocity"][:, :3] = rng.normal(0, 0.1, size=(len(parts), 3)).astype(np.float32)
    p = scenes.fluid_params(sc.n, float(sc.params["h"]), float(sc.params["particle_mass"]),
                            surface_tension=0.0, gravity=0.0, viscosity=0.1)
    srt, table = one_step_inputs(parts, p)
    d = oracle.density_pressure(srt, p, sc.terms, table)
    f = oracle.forces(d, p, sc.terms, table)
    # pressure and viscosity accelerations are antisymmetric pair by pair, so
    # with equal masses the total momentum change vanishes
    a = f["acceleration"][:, :3].astype(np.float64)
    tot = np.abs(a).sum(axis=0)
    assert np.all(np.abs(a.sum(axis=0)) <= 2e-5 * tot)


def test_interior_lattice_forces_cancel():
    h = 0.04
    s = 0.5 * h
    parts = lattice(11, s)
    p = scenes.fluid_params(len(parts), h, 1000.0 * s ** 3, gravity=-9.81)
    t = native.kernel_terms(h)
    srt, table = one_step_inputs(parts, p)
    d = oracle.density_pressure(srt, p, t, table)
    f = oracle.forces(d, p, t, table)
    X = d["position"][:, :3]
    inner = np.all(np.abs(X - (0.3 + 5 * s)) < 2.5 * s + 1e-6, axis=1)
    a = f["acceleration"][inner, :3]
    np.testing.assert_allclose(a[:, 1], -9.81, atol=1e-3)
    np.testing.assert_allclose(a[:, [0, 2]], 0.0, atol=1e-3)


def test_forces_copies_particle_and_advection_zeroes_acceleration():
    sc = scenes.dam_break(2048, seed=5)
    srt, table = one_step_inputs(sc.particles, sc.params.copy())
    p = sc.params.copy()
    oracle.grid_params(sc.particles, p)
    d = oracle.density_pressure(srt, p, sc.terms, table)
    f = oracle.forces(d, p, sc.terms, table)
    for fld in ("position", "velocity", "intermediate_velocity"):
        np.testing.assert_array_equal(f[fld][:, :3], d[fld][:, :3])
    for fld in ("density", "pressure", "grid_index"):
        np.testing.assert_array_equal(f[fld], d[fld])
    adv, capped = oracle.advection_collision(f, p, sc.normals, sc.vertices, sc.indices)
    assert capped == 0
    assert np.all(adv["acceleration"] == 0)
    np.testing.assert_array_equal(adv["density"], f["density"])


# ---------------------------------------------------------- advection

def single(pos, vh, acc, dt=1e-3, vmax=1e9, e=0.0):
    parts = layout.particles(1)
    parts["position"][0, :3] = pos
    parts["intermediate_velocity"][0, :3] = vh
    parts["acceleration"][0, :3] = acc
    p = scenes.fluid_params(1, 0.04, 1e-3, restitution=e)
    p["time_delta"] = np.float32(dt)
    p["max_velocity"] = np.float32(vmax)
    return parts, p


PLANE_V = np.array([1, -1, 1, -1, -1, 1, 1, -1, -1, -1, -1, -1], dtype=np.float32)
PLANE_I = np.array([1, 0, 2, 3, 1, 2], dtype=np.uint32)  # two triangles covering y = -1


def plane_normals():
    v = PLANE_V.reshape(-1, 3).astype(np.float64)
    out = []
    for t in PLANE_I.reshape(-1, 3):
        n = np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])
        out.extend(n / np.linalg.norm(n))
    return np.array(out, dtype=np.float32)


def test_free_fall_closed_form():
    g = -9.81
    dt = 1e-3
    parts, p = single((0.1, 0.5, 0.2), (0, 0, 0), (0, g, 0), dt=dt)
    out, capped = oracle.advection_collision(parts, p, np.zeros(0), np.zeros(0), np.zeros(0))
    assert capped == 0
    np.testing.assert_allclose(out["intermediate_velocity"][0, :3], (0, g * dt, 0), rtol=1e-6)
    np.testing.assert_allclose(out["velocity"][0, :3], (0, g * dt / 2, 0), rtol=1e-6)
    np.testing.assert_allclose(out["position"][0, :3], (0.1, 0.5 + g * dt * dt, 0.2), rtol=1e-6)


def test_speed_clamp():
    parts, p = single((0.0, 0.0, 0.0), (3.0, 4.0, 0.0), (0, 0, 0), dt=1e-3, vmax=2.5)
    out, _ = oracle.advection_collision(parts, p, np.zeros(0), np.zeros(0), np.zeros(0))
    np.testing.assert_allclose(out["intermediate_velocity"][0, :3], (1.5, 2.0, 0.0), rtol=1e-6)
    np.testing.assert_allclose(out["position"][0, :3], (1.5e-3, 2.0e-3, 0.0), rtol=1e-6)


@pytest.mark.parametrize("e", [0.0, 0.5])
def test_plane_collision_known_answer(e):
    # start 1 mm above the plane y = -1, moving (0.5, -2) m/s for 1 ms:
    # the segment hits the plane after half the step
    dt = 1e-3
    v0 = np.array([0.5, -2.0, 0.0])
    x0 = np.array([0.0, -0.999, 0.0])
    parts, p = single(x0, v0, (0, 0, 0), dt=dt, e=e)
    # the closed form uses the values the oracle sees (FP32 inputs, 0.001f offset)
    x0 = parts["position"][0, :3].astype(np.float64)
    dt = float(p["time_delta"])
    out, capped = oracle.advection_collision(parts, p, plane_normals(), PLANE_V, PLANE_I)
    assert capped == 0
    hit = x0 + ((x0[1] + 1.0) / 2.0 / dt) * dt * v0  # reaches y = -1 (mid-step)
    depth = np.linalg.norm(x0 + dt * v0 - hit)
    vn = -2.0  # velocity along the motion-oriented normal (0, -1, 0) is +2
    k = (1 + e * depth / (dt * np.linalg.norm(v0))) * 2.0
    v1 = v0 - k * np.array([0.0, -1.0, 0.0])
    x1 = hit + np.array([0.0, float(np.float32(0.001)), 0.0])  # pulled back 1 mm
    frac = np.linalg.norm(x1 - x0) / np.linalg.norm(dt * v0)
    rem = dt - dt * frac
    x2 = x1 + v1 * rem  # second pass: no further hit (moving away / along)
    # FP64 instantiation against the closed form; FP32 within its cancellation
    # error (the penetration depth is a difference of numbers near 1)
    out64, _ = oracle.advection_collision_f64(oracle.to_f64(parts), p, plane_normals(),
                                              PLANE_V, PLANE_I)
    np.testing.assert_allclose(out64["intermediate_velocity"][0, :3], v1, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(out64["position"][0, :3], x2, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(out64["velocity"][0, :3], (v0 + v1) / 2, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(out["intermediate_velocity"][0, :3], v1, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(out["position"][0, :3], x2, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(out["velocity"][0, :3], (v0 + v1) / 2, rtol=1e-4, atol=1e-6)
    assert vn < 0


def test_collision_tie_goes_to_later_triangle():
    # the same triangle twice; the stored normals differ only in length, which
    # leaves the hit point unchanged but scales the 1 mm pull-back: on an
    # equal-distance tie the later triangle's normal is the one applied
    v = np.array([-1, 0, -1, 1, 0, -1, 0, 0, 1], dtype=np.float32)
    idx = np.array([0, 1, 2, 0, 1, 2], dtype=np.uint32)
    parts, p = single((0.0, 0.0005, 0.0), (0.0, -1.0, 0.0), (0, 0, 0), dt=1e-3)
    later_long = np.array([0, 1, 0, 0, 2, 0], dtype=np.float32)
    later_unit = np.array([0, 2, 0, 0, 1, 0], dtype=np.float32)
    out_a, _ = oracle.advection_collision(parts, p, later_long, v, idx)
    out_b, _ = oracle.advection_collision(parts, p, later_unit, v, idx)
    # Both hit y = 0 half-way through the 1 ms step. Worked by hand:
    #  later normal (0,2,0): n = (0,-2,0) along the motion, x -> 0.002,
    #    v -> (0,-1,0) - 2 (0,-2,0) = (0,3,0), consumed time 1.5 ms (> dt), so
    #    the second pass runs for -0.5 ms and ends at y = 0.002 - 0.0015.
    #  later normal (0,1,0): x -> 0.001, v -> 0, nothing moves afterwards.
    np.testing.assert_allclose(out_a["intermediate_velocity"][0, :3], (0, 3, 0), atol=1e-5)
    assert abs(out_a["position"][0, 1] - 0.0005) < 1e-7
    np.testing.assert_allclose(out_b["intermediate_velocity"][0, :3], (0, 0, 0), atol=1e-6)
    assert abs(out_b["position"][0, 1] - 0.001) < 1e-7


def test_bounce_cap_counts():
    # planes y = 0 and y = 0.01; at 100 m/s with restitution 1 the particle
    # ricochets 5-8 times within one 1 ms step (each hit keeps the fraction
    # depth / (dt |v|) of the normal speed), so a cap of 3 is reached
    v = np.array([-1, 0, -1, 1, 0, -1, 0, 0, 1, -1, 0.01, -1, 1, 0.01, -1, 0, 0.01, 1],
                 dtype=np.float32)
    idx = np.array([0, 1, 2, 3, 4, 5], dtype=np.uint32)
    normals = np.array([0, 1, 0, 0, 1, 0], dtype=np.float32)
    parts, p = single((0.0, 0.005, 0.0), (0.0, 100.0, 0.0), (0, 0, 0), dt=1e-3, e=1.0)
    out, capped = oracle.advection_collision(parts, p, normals, v, idx, max_bounces=3)
    assert capped == 1
    assert 0.0 < out["position"][0, 1] < 0.01
    out, capped = oracle.advection_collision(parts, p, normals, v, idx, max_bounces=100)
    assert capped == 0
    assert 0.0 < out["position"][0, 1] < 0.01


# ------------------------------------------------------- FP32 vs FP64

def test_fp32_error_against_fp64():
    sc = scenes.dam_break(4096, seed=2)
    r = oracle.stages(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices, sc.indices)
    s64 = oracle.to_f64(r.sorted)
    d64 = oracle.density_pressure_f64(s64, r.params, sc.terms, r.table)
    f64 = oracle.forces_f64(d64, r.params, sc.terms, r.table)
    rel = np.abs(r.density["density"] - d64["density"]) / d64["density"]
    assert rel.max() < 2e-6
    a64 = f64["acceleration"][:, :3]
    err = np.linalg.norm(r.forces["acceleration"][:, :3] - a64, axis=1).max()
    assert err < 1e-5 * np.linalg.norm(a64, axis=1).max()
