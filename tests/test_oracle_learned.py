"""Pins of the learned surrogate's CPU oracle (oracle/learned.py).

The network is this project's own definition, so there is no reference
output to match. The oracle is pinned by:
  - central finite differences of its FP64 forward (its backward);
  - a literal particle-by-particle evaluation with the pair features written
    out from the forces kernel's formulas (the vectorised gather/scatter);
  - the zero network returning the forces-stage acceleration exactly;
  - invariance to the order of each particle's neighbours and to a
    translation of all positions;
  - bfloat16 round-to-nearest-even on hand-worked bit patterns;
  - Adam's first step, which moves each parameter by lr * sign(grad).
"""
import numpy as np
import pytest

from oracle import learned as L


def tiny_graph(n=14, seed=0, h=0.01):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 0.02, (n, 3))
    v = rng.normal(0, 0.3, (n, 3))
    rho = rng.uniform(900, 1300, n)
    prs = rng.uniform(-2, 50, n)
    lists = []
    for i in range(n):
        d = np.linalg.norm(x - x[i], axis=1)
        lists.append([j for j in range(n) if j != i and d[j] < h])
    off, src, dst, deg = L.graph_from_lists(lists)
    # (particle mass, viscosity, spiky, viscosity kernel, poly6 gradient, poly6 laplacian)
    # of spacing h/2, as the dam breaks use them
    m = 1000.0 * (h / 2) ** 3
    consts = (m, 3.5e-3, -45 / (np.pi * h ** 6), 45 / (np.pi * h ** 6),
              -945 / (32 * np.pi * h ** 9), -945 / (32 * np.pi * h ** 9))
    acc0 = rng.normal(0, 20, (n, 3))
    return dict(x=x, v=v, rho=rho, prs=prs, acc0=acc0, src=src, dst=dst, deg=deg, h=h,
                rho0=1000.0, K=3.0, g=np.array([0.0, -9.81, 0.0]), consts=consts)


def test_param_layout():
    assert L.param_count() == 11456
    flat = L.init_params(0)
    P = L.unpack(flat)
    assert np.all(P["e.w1"][27:] == 0) and np.all(P["n.w1"][38:] == 0)
    assert np.all(P["n.w3"][:, 3:] == 0)
    assert np.array_equal(L.pack(P), flat)


def test_backward_matches_finite_differences():
    g = tiny_graph()
    assert len(g["src"]) > 20
    flat = L.random_params(1)
    target = np.random.default_rng(2).normal(0, 1, (len(g["rho"]), 3))
    out, c = L.forward(flat, keep=True, **g)
    grad = L.backward(flat, c, target, g["src"], g["dst"], g["deg"])
    f = lambda w: L.loss(L.forward(w, **g), target)
    rng = np.random.default_rng(3)
    o, idx = 0, []
    for name, shape in L.LAYOUT:
        k = int(np.prod(shape))
        idx += list(o + rng.choice(k, min(k, 4), replace=False))
        o += k
    gmax = np.abs(grad).max()
    eps = 1e-6
    for i in idx:
        wp, wm = flat.copy(), flat.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (f(wp) - f(wm)) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-4 * gmax + 1e-9, (i, fd, grad[i])


def test_padded_parameters_get_zero_gradient():
    g = tiny_graph(seed=4)
    flat = L.random_params(5)
    target = np.ones((len(g["rho"]), 3))
    out, c = L.forward(flat, keep=True, **g)
    G = L.unpack(L.backward(flat, c, target, g["src"], g["dst"], g["deg"]))
    assert np.all(G["e.w1"][27:] == 0) and np.all(G["n.w1"][38:] == 0)
    assert np.all(G["n.w3"][:, 3:] == 0) and np.all(G["n.b3"][3:] == 0)
    assert np.abs(G["e.w1"][:27]).max() > 0 and np.abs(G["n.w1"][:38]).max() > 0


def test_literal_per_particle_evaluation():
    g = tiny_graph(seed=6)
    flat = L.random_params(7)
    P = L.unpack(flat)
    out = L.forward(flat, **g)
    n = len(g["rho"])
    h = g["h"]
    m, mu, c_spiky, c_visc, c_grad, c_lap = g["consts"]
    silu = lambda z: z / (1.0 + np.exp(-z))
    for i in range(n):
        S = np.zeros(L.EDGE_W)
        deg = 0
        nu_i = [g["rho"][i] / g["rho0"] - 1.0, np.arcsinh(g["prs"][i] / g["K"]) / 4.0]
        for j in range(n):
            if j == i or np.linalg.norm(g["x"][i] - g["x"][j]) >= h:
                continue
            deg += 1
            d = g["x"][i] - g["x"][j]
            r = np.sqrt(d @ d)
            q = r / h
            nu_j = [g["rho"][j] / g["rho0"] - 1.0, np.arcsinh(g["prs"][j] / g["K"]) / 4.0]
            ps = (g["prs"][i] / g["rho"][i] ** 2 + g["prs"][j] / g["rho"][j] ** 2) * m
            pres = -ps * c_spiky * (h - r) ** 2 / r * d / L.ACC_SCALE
            visc = mu / g["rho"][i] * (m / g["rho"][j]) * c_visc * (h - r) * \
                (g["v"][j] - g["v"][i]) / L.ACC_SCALE
            t = h * h - r * r
            grad = (m / g["rho"][j]) * c_grad * t * t * h * d
            lap = (m / g["rho"][j]) * c_lap * t * (3 * h * h - 7 * r * r) * h * h
            dv = g["v"][i] - g["v"][j]
            xv = d @ dv
            rate = -ps * c_spiky * ((h - r) ** 2 / r * dv - xv * (h - r) * (h + r) / r ** 3 * d) \
                * h / (L.VEL_SCALE * L.ACC_SCALE)
            rho_rate = m * c_grad * t * t * xv * 1000.0 * h / (g["rho0"] * L.VEL_SCALE)
            e = np.concatenate([d / h, dv, [q, (1 - q * q) ** 3], nu_i, nu_j, pres, visc, grad,
                                [lap], rate, [rho_rate, xv * 10.0 / (h * L.VEL_SCALE)],
                                np.zeros(5)])
            S += np.maximum(e @ P["e.w1"] + P["e.b1"], 0.0) / L.AGG_C
        u = np.concatenate([S, nu_i, (g["acc0"][i] - g["g"]) / L.ACC_SCALE, [deg / L.AGG_C],
                            np.zeros(26)])
        y1 = silu(u @ P["n.w1"] + P["n.b1"])
        y2 = silu(y1 @ P["n.w2"] + P["n.b2"])
        want = y2 @ P["n.w3"] + P["n.b3"]
        assert np.allclose(out[i], want, rtol=1e-11, atol=1e-12), i


def test_motion_rate_inputs_are_time_derivatives():
    """Inputs 22..25 are rates along the pair's motion: moving both particles
    along their velocities, central differences in time of the pressure pair
    acceleration (at fixed pressures) and of the pair's poly6 density term
    give them."""
    g = tiny_graph(seed=13)
    h = g["h"]
    m, mu, c_spiky, c_visc, c_grad, c_lap = g["consts"]
    poly6 = 315.0 / (64.0 * np.pi * h ** 9)
    nu = L.node_inputs(g["rho"], g["prs"], g["rho0"], g["K"])
    e = L.edge_inputs(g["x"], g["v"], nu, g["src"], g["dst"], h, g["rho"], g["prs"],
                      g["consts"], g["rho0"])
    eps = 1e-7
    for k in range(len(g["src"])):
        i, j = g["dst"][k], g["src"][k]
        ps = (g["prs"][i] / g["rho"][i] ** 2 + g["prs"][j] / g["rho"][j] ** 2) * m

        def pres(tt):
            d = (g["x"][i] + tt * g["v"][i]) - (g["x"][j] + tt * g["v"][j])
            r = np.sqrt(d @ d)
            return -ps * c_spiky * (h - r) ** 2 / r * d

        def dens(tt):
            d = (g["x"][i] + tt * g["v"][i]) - (g["x"][j] + tt * g["v"][j])
            return m * poly6 * (h * h - d @ d) ** 3

        rate = (pres(eps) - pres(-eps)) / (2 * eps) * h / (L.VEL_SCALE * L.ACC_SCALE)
        rho_rate = (dens(eps) - dens(-eps)) / (2 * eps) * 1000.0 * h / (g["rho0"] * L.VEL_SCALE)
        scale = np.abs(rate).max() + 1e-12
        assert np.allclose(e[k, 22:25], rate, rtol=1e-5, atol=1e-6 * scale), k
        assert np.isclose(e[k, 25], rho_rate, rtol=1e-5, atol=1e-9), k


def test_zero_network_returns_the_forces_stage_acceleration():
    g = tiny_graph(seed=11)
    out = L.forward(np.zeros(L.param_count()), **g)
    assert np.all(out == 0.0)
    assert np.array_equal(L.acceleration(out, g["acc0"]), g["acc0"])


def test_invariance_to_neighbour_order_and_translation():
    g = tiny_graph(seed=8)
    flat = L.random_params(9)
    ref = L.forward(flat, **g)
    rng = np.random.default_rng(10)
    off = np.concatenate([[0], np.cumsum(g["deg"])])
    perm = np.concatenate([off[i] + rng.permutation(g["deg"][i]) for i in range(len(g["deg"]))])
    g2 = dict(g, src=g["src"][perm], dst=g["dst"][perm])
    assert np.allclose(L.forward(flat, **g2), ref, rtol=1e-12, atol=1e-12)
    g3 = dict(g, x=g["x"] + np.array([0.25, -0.5, 0.125]))
    assert np.allclose(L.forward(flat, **g3), ref, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("value,bits", [
    (1.0, 0x3F80),
    (1.0 + 2.0 ** -8, 0x3F80),          # halfway, ties to the even pattern 0x3F80
    (1.0 + 3 * 2.0 ** -8, 0x3F82),      # halfway, ties to the even pattern 0x3F82
    (1.0 + 2.0 ** -8 + 2.0 ** -20, 0x3F81),
    (-2.5, 0xC020),
    (0.0, 0x0000),
])
def test_bf16_rounding(value, bits):
    assert int(np.ravel(L.bf16_bits(np.float32(value)))[0]) == bits
    assert L.bf16(np.float32(value)) == L.bits_to_float(bits)


def test_adam_first_step_is_sign_step():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    m, v = np.zeros(3), np.zeros(3)
    L.adam(p, m, v, g, step=1, lr=0.01, eps=0.0)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01, 0.5 - 0.01], rtol=0, atol=1e-15)


def test_product_layout_and_init_match_the_oracle():
    from libclsph_amd import learned as product
    assert [(n, s) for n, s, _ in product.LAYOUT] == L.LAYOUT
    assert np.array_equal(product.init_params(3), L.init_params(3).astype(np.float32))


def test_pair_features_sum_to_the_forces_stage():
    """Summed over each particle's neighbours, the pair-term inputs rebuild
    the oracle's FP64 forces-stage acceleration (pressure and viscosity
    directly; surface tension from the colour-field gradient and laplacian
    with the self term and the threshold)."""
    import oracle
    from libclsph_amd import scenes
    sc = scenes.dam_break(512, seed=3)
    st32, _, _, p = oracle.run(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices,
                               sc.indices, 30)
    r = oracle.stages(st32, p, sc.terms, sc.normals, sc.vertices, sc.indices)
    s64 = oracle.to_f64(r.sorted)
    d64 = oracle.density_pressure_f64(s64, r.params, sc.terms, r.table)
    f64 = oracle.forces_f64(d64, r.params, sc.terms, r.table)
    x = d64["position"][:, :3]
    v = d64["velocity"][:, :3]
    rho, prs = d64["density"], d64["pressure"]
    h = float(p["h"])
    t = sc.terms
    consts = (float(p["particle_mass"]), float(p["dynamic_viscosity"]), float(t["spiky"]),
              float(t["viscosity"]), float(t["poly_6_gradient"]), float(t["poly_6_laplacian"]))
    n = len(rho)
    lists = []
    for i in range(n):
        d = np.linalg.norm(x - x[i], axis=1)
        lists.append([j for j in np.nonzero(d < h)[0] if j != i])
    off, src, dst, deg = L.graph_from_lists(lists)
    assert deg.max() < L.MAX_EDGES
    nu = L.node_inputs(rho, prs, float(p["fluid_density"]), float(p["K"]))
    e = L.edge_inputs(x, v, nu, src, dst, h, rho, prs, consts)
    m, c_lap = consts[0], consts[5]
    a = L.ACC_SCALE * L.segsum(e[:, 12:18], dst, n)
    a = a[:, 0:3] + a[:, 3:6]
    normal = L.segsum(e[:, 18:21], dst, n) / h
    lap = m / rho * c_lap * h * h * 3 * h * h + L.segsum(e[:, 21:22], dst, n)[:, 0] / (h * h)
    nl = np.linalg.norm(normal, axis=1)
    on = nl > float(p["surface_tension_threshold"])
    st = np.zeros((n, 3))
    st[on] = (-float(p["surface_tension"]) * lap[on] / nl[on])[:, None] * normal[on]
    a += st / rho[:, None] + p["constant_acceleration"][:3].astype(np.float64)
    ref = f64["acceleration"][:, :3]
    assert on.any() and (~on).any()
    assert np.abs(a - ref).max() <= 1e-9 * np.abs(ref).max()
