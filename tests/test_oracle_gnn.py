"""Pins of the two-round network's CPU oracle (oracle/gnn.py).

The network is this project's own definition, so there is no reference
output to match. The oracle is pinned by:
  - central finite differences of its FP64 forward (its backward);
  - a literal per-edge evaluation of round 1 with the third edge layer
    applied on every edge (the sum-then-project restructuring);
  - invariance to the order of each particle's neighbours and to a
    translation of all positions;
  - bfloat16 round-to-nearest-even on hand-worked bit patterns;
  - Adam's first step, which moves each parameter by lr * sign(grad).
"""
import numpy as np
import pytest

from oracle import gnn as L


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
    return dict(x=x, v=v, rho=rho, prs=prs, src=src, dst=dst, deg=deg, h=h, rho0=1000.0, K=3.0)


def test_param_layout():
    assert L.param_count() == 208288
    flat = L.init_params(0)
    P = L.unpack(flat)
    assert np.all(P["e1.w1"][12:] == 0) and np.all(P["e2.w1"][12:] == 0)
    assert np.all(P["n1.w1"][130:] == 0) and np.all(P["n2.w3"][:, 3:] == 0)
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
    assert np.all(G["e1.w1"][12:] == 0) and np.all(G["e2.w1"][12:] == 0)
    assert np.all(G["n1.w1"][130:] == 0)
    assert np.all(G["n2.w3"][:, 3:] == 0) and np.all(G["n2.b3"][3:] == 0)


def test_sum_then_project_equals_per_edge_layer():
    g = tiny_graph(seed=6)
    flat = L.random_params(7)
    P = L.unpack(flat)
    out, c = L.forward(flat, keep=True, **g)
    n = len(g["rho"])
    nu = L.node_inputs(g["rho"], g["prs"], g["rho0"], g["K"])
    M1 = np.zeros((n, L.WIDTH))
    for i in range(n):
        for j in range(n):
            if j == i or np.linalg.norm(g["x"][i] - g["x"][j]) >= g["h"]:
                continue
            d = g["x"][i] - g["x"][j]
            q = np.sqrt(d @ d) / g["h"]
            e = np.concatenate([d / g["h"], g["v"][i] - g["v"][j], [q, (1 - q * q) ** 3],
                                nu[i], nu[j], np.zeros(4)])
            s1 = L.silu(e @ P["e1.w1"] + P["e1.b1"])
            s2 = L.silu(s1 @ P["e1.w2"] + P["e1.b2"])
            M1[i] += (s2 @ P["e1.w3"] + P["e1.b3"]) / L.AGG_C
    assert np.allclose(c["n1_u"][:, :L.WIDTH], M1, rtol=1e-12, atol=1e-12)


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


def test_param_count_matches_the_library():
    from libclsph_amd import native
    assert int(native.lib.clsph_gnn_param_count()) == L.param_count()
