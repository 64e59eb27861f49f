"""CPU oracle of the two-round message-passing network (test infrastructure
only), the full-size network of libclsph_amd/csrc/gnn.h.

Only tests, the smoke entry and the benchmark's CPU baseline may import this
module; the product path never calls it.

The network has no counterpart in the reference: the reference predicts
accelerations with the `forces` kernel only (kernels/sph.cl:42-62). This file
is the definition the GPU kernels are checked against; tests/test_oracle_gnn.py
pins it (finite differences of the backward, the per-edge form of the third
edge layer, invariances, bf16 bit patterns, Adam's first step). It is a
message-passing network over the particles' neighbour graph (r < h, self
excluded, at most 128 neighbours per particle) with two rounds:

  node inputs   nu_i = [rho_i / rho0 - 1, asinh(p_i / K) / 4]
  edge inputs   e_ij = [(x_i - x_j) / h, (v_i - v_j) / V, q, (1 - q^2)^3,
                        nu_i, nu_j, 0 0 0 0]            q = |x_i - x_j| / h
  round 1       s1 = silu(e W1 + b1); s2 = silu(s1 W2 + b2)
                M1_i = (sum_j s2_ij / C) W3 + (deg_i / C) b3
                H_i  = MLP_n1([M1_i, nu_i, 0...])        (3 layers, 144 -> 128)
  round 2       z1 = e W1' + b1' + (H_i Wa) + (H_j Wb)
                s1 = silu(z1); s2 = silu(s1 W2' + b2')
                M2_i = (sum_j s2_ij / C) W3' + (deg_i / C) b3'
                out_i = MLP_n2([H_i, M2_i])              (3 layers, 256 -> 32)
  acceleration  a_i = g + A * out_i[0:3]

The third edge layer is applied after the sum over neighbours. That is exact,
because the layer is linear: sum_j (s2_ij W3 + b3) = (sum_j s2_ij) W3 + deg_i b3.
The identity is pinned in the tests against the literal per-edge form.

Rounding points. The GPU rounds every matrix-product operand to bfloat16
(round to nearest even) and accumulates in float32. `forward(..., bf16=True)`
rounds at the same points, so the GPU forward can be checked tightly;
`bf16=False` is the exact FP64 network. `backward` is the analytic gradient
of the FP64 network (rounding treated as the identity). It is pinned by
central finite differences of `forward`.
"""
import numpy as np

WIDTH = 128
FEAT = 16          # edge input features (12 used, 4 zero)
NODE_IN1 = 144     # node MLP 1 input: M1 (128) + nu (2) + zeros
NODE_IN2 = 256     # node MLP 2 input: H (128) + M2 (128)
OUT = 32           # node MLP 2 output (3 used)
AGG_C = 32.0       # aggregation normaliser
VEL_SCALE = 1.0    # V, m/s
ACC_SCALE = 10.0   # A, m/s^2
MAX_EDGES = 128

# Flat parameter vector: (name, shape) in storage order (row-major, in x out).
LAYOUT = [
    ("e1.w1", (FEAT, WIDTH)), ("e1.b1", (WIDTH,)),
    ("e1.w2", (WIDTH, WIDTH)), ("e1.b2", (WIDTH,)),
    ("e1.w3", (WIDTH, WIDTH)), ("e1.b3", (WIDTH,)),
    ("n1.w1", (NODE_IN1, WIDTH)), ("n1.b1", (WIDTH,)),
    ("n1.w2", (WIDTH, WIDTH)), ("n1.b2", (WIDTH,)),
    ("n1.w3", (WIDTH, WIDTH)), ("n1.b3", (WIDTH,)),
    ("e2.w1", (FEAT, WIDTH)), ("e2.b1", (WIDTH,)),
    ("e2.wa", (WIDTH, WIDTH)), ("e2.wb", (WIDTH, WIDTH)),
    ("e2.w2", (WIDTH, WIDTH)), ("e2.b2", (WIDTH,)),
    ("e2.w3", (WIDTH, WIDTH)), ("e2.b3", (WIDTH,)),
    ("n2.w1", (NODE_IN2, WIDTH)), ("n2.b1", (WIDTH,)),
    ("n2.w2", (WIDTH, WIDTH)), ("n2.b2", (WIDTH,)),
    ("n2.w3", (WIDTH, OUT)), ("n2.b3", (OUT,)),
]


def param_count():
    return sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(flat):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + k]).reshape(shape)
        o += k
    return out


def pack(params):
    return np.concatenate([np.asarray(params[n], dtype=np.float64).ravel() for n, _ in LAYOUT])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def graph_from_lists(nbr_lists):
    """CSR (dst-major) edge arrays from per-particle neighbour lists, each
    truncated to MAX_EDGES entries."""
    deg = np.array([min(len(l), MAX_EDGES) for l in nbr_lists], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(deg)])
    src = np.concatenate([np.asarray(l[:MAX_EDGES], dtype=np.int64) for l in nbr_lists]
                         + [np.zeros(0, np.int64)])
    dst = np.repeat(np.arange(len(nbr_lists)), deg)
    return off, src, dst, deg


def node_inputs(rho, prs, rho0, K):
    nu = np.zeros((len(rho), 2))
    nu[:, 0] = np.asarray(rho, np.float64) / rho0 - 1.0
    nu[:, 1] = np.arcsinh(np.asarray(prs, np.float64) / K) * 0.25
    return nu


def edge_inputs(x, v, nu, src, dst, h):
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    d = x[dst] - x[src]
    q = np.sqrt((d * d).sum(1)) / h
    e = np.zeros((len(src), FEAT))
    e[:, 0:3] = d / h
    e[:, 3:6] = (v[dst] - v[src]) / VEL_SCALE
    e[:, 6] = q
    e[:, 7] = (1.0 - q * q) ** 3
    e[:, 8:10] = nu[dst]
    e[:, 10:12] = nu[src]
    return e


def segsum(vals, idx, n):
    out = np.zeros((n,) + vals.shape[1:])
    np.add.at(out, idx, vals)
    return out


def forward(flat, x, v, rho, prs, src, dst, deg, h, rho0, K, bf16_round=False, keep=False):
    """Network output `out` (N x 32; columns 0..2 are the acceleration
    residual in units of A). With keep=True also returns the intermediates
    the backward pass uses."""
    P = unpack(flat)
    R = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    n = len(rho)
    degc = np.asarray(deg, np.float64)[:, None] / AGG_C
    nu = node_inputs(rho, prs, rho0, K)
    e = R(edge_inputs(x, v, nu, src, dst, h))
    c = {"e": e, "nu": nu}
    # round 1 edges
    z1 = e @ R(P["e1.w1"]) + P["e1.b1"]
    s1 = R(silu(z1))
    z2 = s1 @ R(P["e1.w2"]) + P["e1.b2"]
    S = R(segsum(silu(z2), dst, n) / AGG_C)
    M1 = S @ R(P["e1.w3"]) + degc * P["e1.b3"]
    c.update(r1_z1=z1, r1_s1=s1, r1_z2=z2, r1_S=S)
    # node MLP 1
    u1 = np.zeros((n, NODE_IN1))
    u1[:, :WIDTH] = M1
    u1[:, WIDTH:WIDTH + 2] = nu
    u1 = R(u1)
    p1 = u1 @ R(P["n1.w1"]) + P["n1.b1"]
    y1 = R(silu(p1))
    p2 = y1 @ R(P["n1.w2"]) + P["n1.b2"]
    y2 = R(silu(p2))
    H = R(y2 @ R(P["n1.w3"]) + P["n1.b3"])
    c.update(n1_u=u1, n1_p1=p1, n1_y1=y1, n1_p2=p2, n1_y2=y2, H=H)
    # round 2 edges
    PA = R(H @ R(P["e2.wa"]))
    PB = R(H @ R(P["e2.wb"]))
    z1 = e @ R(P["e2.w1"]) + P["e2.b1"] + PA[dst] + PB[src]
    s1 = R(silu(z1))
    z2 = s1 @ R(P["e2.w2"]) + P["e2.b2"]
    S = R(segsum(silu(z2), dst, n) / AGG_C)
    M2 = S @ R(P["e2.w3"]) + degc * P["e2.b3"]
    c.update(r2_z1=z1, r2_s1=s1, r2_z2=z2, r2_S=S)
    # node MLP 2
    u2 = R(np.concatenate([H, M2], axis=1))
    q1 = u2 @ R(P["n2.w1"]) + P["n2.b1"]
    w1 = R(silu(q1))
    q2 = w1 @ R(P["n2.w2"]) + P["n2.b2"]
    w2 = R(silu(q2))
    out = w2 @ R(P["n2.w3"]) + P["n2.b3"]
    c.update(n2_u=u2, n2_p1=q1, n2_y1=w1, n2_p2=q2, n2_y2=w2, out=out)
    return (out, c) if keep else out


def loss(out, target):
    """Mean over particles of the squared error of the three used outputs."""
    d = out[:, :3] - target
    return float((d * d).sum() / len(target))


def backward(flat, c, target, src, dst, deg):
    """Gradient of loss(forward(...), target) with respect to the flat
    parameter vector, from the intermediates of forward(..., keep=True)."""
    P = unpack(flat)
    n = len(target)
    degc = np.asarray(deg, np.float64)[:, None] / AGG_C
    G = {name: np.zeros(shape) for name, shape in LAYOUT}
    dout = np.zeros((n, OUT))
    dout[:, :3] = 2.0 * (c["out"][:, :3] - target) / n

    def mlp3_back(prefix, u, p1, y1, p2, y2, dy):
        G[prefix + ".w3"] += y2.T @ dy
        G[prefix + ".b3"] += dy.sum(0)
        dp2 = (dy @ P[prefix + ".w3"].T) * dsilu(p2)
        G[prefix + ".w2"] += y1.T @ dp2
        G[prefix + ".b2"] += dp2.sum(0)
        dp1 = (dp2 @ P[prefix + ".w2"].T) * dsilu(p1)
        G[prefix + ".w1"] += u.T @ dp1
        G[prefix + ".b1"] += dp1.sum(0)
        return dp1 @ P[prefix + ".w1"].T

    def edges_back(prefix, r, dM):
        S = c[r + "_S"]
        G[prefix + ".w3"] += S.T @ dM
        G[prefix + ".b3"] += (degc * dM).sum(0)
        dS = (dM @ P[prefix + ".w3"].T) / AGG_C
        dz2 = dS[dst] * dsilu(c[r + "_z2"])
        G[prefix + ".w2"] += c[r + "_s1"].T @ dz2
        G[prefix + ".b2"] += dz2.sum(0)
        dz1 = (dz2 @ P[prefix + ".w2"].T) * dsilu(c[r + "_z1"])
        G[prefix + ".w1"] += c["e"].T @ dz1
        G[prefix + ".b1"] += dz1.sum(0)
        return dz1

    du2 = mlp3_back("n2", c["n2_u"], c["n2_p1"], c["n2_y1"], c["n2_p2"], c["n2_y2"], dout)
    dH = du2[:, :WIDTH].copy()
    dz1 = edges_back("e2", "r2", du2[:, WIDTH:])
    dPA = segsum(dz1, dst, n)
    dPB = segsum(dz1, src, n)
    G["e2.wa"] += c["H"].T @ dPA
    G["e2.wb"] += c["H"].T @ dPB
    dH += dPA @ P["e2.wa"].T + dPB @ P["e2.wb"].T
    du1 = mlp3_back("n1", c["n1_u"], c["n1_p1"], c["n1_y1"], c["n1_p2"], c["n1_y2"], dH)
    edges_back("e1", "r1", du1[:, :WIDTH])
    return pack(G)


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights over the inputs that
    are used, zero biases, zero rows for the padded inputs and a last layer
    scaled by 0.1."""
    rng = np.random.default_rng(seed)
    used_in = {"e1.w1": 12, "e2.w1": 12, "n1.w1": WIDTH + 2}
    P = {}
    for name, shape in LAYOUT:
        if len(shape) == 1:
            P[name] = np.zeros(shape)
            continue
        fan = used_in.get(name, shape[0])
        w = rng.uniform(-1.0, 1.0, size=shape) / np.sqrt(fan)
        w[fan:] = 0.0
        if name == "n2.w3":
            w[:, 3:] = 0.0
            w *= 0.1
        P[name] = w
    return pack(P)


def adam(flat, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected), in place."""
    m *= beta1
    m += (1.0 - beta1) * grad
    v *= beta2
    v += (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    flat -= lr * mh / (np.sqrt(vh) + eps)
    return flat


def bf16_bits(x):
    """Raw bfloat16 bit patterns (uint16) of float32 values rounded to
    nearest even."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bits_to_float(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def random_params(seed, bias_scale=0.1):
    """init_params with nonzero random biases, keeping every padded row and
    column at zero."""
    P = unpack(init_params(seed))
    rng = np.random.default_rng(seed + 1000)
    for name, shape in LAYOUT:
        if len(shape) == 1:
            b = rng.normal(0.0, bias_scale, size=shape)
            if name == "n2.b3":
                b[3:] = 0.0
            P[name] = b
    return pack(P)
