"""CPU oracle of the learned force surrogate (test infrastructure only).

Only tests, the smoke entry and the benchmark's CPU baseline may import this
module; the product path never calls it.

The surrogate has no counterpart in the reference: the reference predicts
accelerations with the `forces` kernel only (kernels/sph.cl:42-62). This file
is the definition the GPU kernels are checked against. The network corrects
the forces stage's acceleration a0_i so that it stands for the mean
acceleration over a macro step of 4 substeps. It is one round of message
passing over the particles' neighbour graph (r < h, self excluded, at most
128 neighbours per particle):

  node inputs   nu_i = [rho_i / rho0 - 1, asinh(p_i / K) / 4]
  edge inputs   e_ij = [(x_i - x_j) / h, (v_i - v_j) / V, q, (1 - q^2)^3,
                        nu_i, nu_j,                     q = |x_i - x_j| / h
                        pressure and viscosity pair accelerations / A,
                        colour-gradient pair term * h, colour-laplacian term * h^2,
                        d/dt of the pressure pair acceleration at fixed
                        pressures * h / (V A), the density-rate pair term
                        * 1000 h / (rho0 V), (x_i - x_j).(v_i - v_j) * 10 / (h V),
                        0 x 5]
                the pair terms are those the forces kernel sums
                (kernels/forces.cl:60-101), from positions, velocities,
                densities and pressures only
  edges         S_i = bf16(sum_j relu(e_ij W1 + b1) / C)              (32 wide)
  node          u_i = [S_i, nu_i, (a0_i - g) / A, deg_i / C, 0 x 26]  (64 wide)
                out_i = MLP(u_i): 64 -> 64 -> 64 -> 32, SiLU between layers
  acceleration  a_i = a0_i + R * out_i[0:3]

With every parameter zero the network returns a0 exactly, so the method
without it is the plain substep at the macro step.

Rounding points. The GPU rounds every matrix-product operand to bfloat16
(round to nearest even) and accumulates in float32. `forward(...,
bf16_round=True)` rounds at the same points, so the GPU forward can be
checked tightly; `bf16_round=False` is the exact FP64 network. `backward`
is the analytic gradient of the FP64 network (rounding treated as the
identity). It is pinned by central finite differences of `forward`.
"""
import numpy as np

EDGE_W = 32        # edge layer width
NODE_W = 64        # node MLP width
FEAT = 32          # edge input features (27 used, 5 zero)
FEAT_USED = 27
NODE_IN = 64       # node MLP input: S (32) + nu (2) + a0 (3) + deg (1) + zeros
NODE_IN_USED = 38
OUT = 32           # network output (3 used)
AGG_C = 32.0       # aggregation normaliser
VEL_SCALE = 1.0    # V, m/s
ACC_SCALE = 10.0   # A, m/s^2
RES_SCALE = 2.0    # R, m/s^2
MAX_EDGES = 128

# Flat parameter vector: (name, shape) in storage order (row-major, in x out).
LAYOUT = [
    ("e.w1", (FEAT, EDGE_W)), ("e.b1", (EDGE_W,)),
    ("n.w1", (NODE_IN, NODE_W)), ("n.b1", (NODE_W,)),
    ("n.w2", (NODE_W, NODE_W)), ("n.b2", (NODE_W,)),
    ("n.w3", (NODE_W, OUT)), ("n.b3", (OUT,)),
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


def relu(x):
    return np.maximum(x, 0.0)


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


def edge_inputs(x, v, nu, src, dst, h, rho=None, prs=None, consts=None, rho0=1000.0):
    """Edge inputs; the pair terms (columns 12..26) need rho, prs and
    consts = (particle_mass, dynamic_viscosity, spiky, viscosity,
    poly_6_gradient, poly_6_laplacian) and are zero without them."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    d = x[dst] - x[src]
    r = np.sqrt((d * d).sum(1))
    q = r / h
    e = np.zeros((len(src), FEAT))
    e[:, 0:3] = d / h
    e[:, 3:6] = (v[dst] - v[src]) / VEL_SCALE
    e[:, 6] = q
    e[:, 7] = (1.0 - q * q) ** 3
    e[:, 8:10] = nu[dst]
    e[:, 10:12] = nu[src]
    if consts is not None:
        m, mu, c_spiky, c_visc, c_grad, c_lap = (float(c) for c in consts)
        rho = np.asarray(rho, np.float64)
        prs = np.asarray(prs, np.float64)
        hr = h - r
        t = h * h - r * r
        # pressure: -(p_i/rho_i^2 + p_j/rho_j^2) m c_spiky (h - r)^2 (x_i - x_j) / r
        ps = (prs[dst] / rho[dst] ** 2 + prs[src] / rho[src] ** 2) * m
        g = np.where(r >= 1e-7, ps * c_spiky * hr * hr / np.maximum(r, 1e-30), 0.0)
        e[:, 12:15] = -g[:, None] * d / ACC_SCALE
        # viscosity: mu / rho_i (v_j - v_i) (m / rho_j) c_visc (h - r)
        vw = (m / rho[src]) * c_visc * hr
        e[:, 15:18] = (mu / rho[dst] * vw)[:, None] * (v[src] - v[dst]) / ACC_SCALE
        # colour field: gradient (m / rho_j) c_grad t^2 (x_i - x_j) and laplacian
        # (m / rho_j) c_lap t (3 h^2 - 7 r^2), scaled by h and h^2
        w = m / rho[src]
        e[:, 18:21] = (w * c_grad * t * t * h)[:, None] * d
        e[:, 21] = w * c_lap * t * (3 * h * h - 7 * r * r) * h * h
        # rates along the pair's motion (x = x_i - x_j, v = v_i - v_j): the
        # time derivative of the pressure pair acceleration at fixed
        # pressures, -ps c_spiky [(h - r)^2 / r v - (x.v)(h - r)(h + r) / r^3 x]
        dv = v[dst] - v[src]
        xv = (d * dv).sum(1)
        rs = np.maximum(r, 1e-30)
        ra = np.where(r >= 1e-7, hr * hr / rs, 0.0)
        rb = np.where(r >= 1e-7, xv * hr * (h + r) / rs ** 3, 0.0)
        cr = -ps * c_spiky * h / (VEL_SCALE * ACC_SCALE)
        e[:, 22:25] = cr[:, None] * (ra[:, None] * dv - rb[:, None] * d)
        # the density-rate pair term m c_grad (h^2 - r^2)^2 (x.v), x 1000 h / (rho0 V)
        e[:, 25] = m * c_grad * t * t * xv * 1000.0 * h / (rho0 * VEL_SCALE)
        e[:, 26] = xv * 10.0 / (h * VEL_SCALE)
    return e


def segsum(vals, idx, n):
    out = np.zeros((n,) + vals.shape[1:])
    np.add.at(out, idx, vals)
    return out


def node_vector(S, nu, acc0, deg, g):
    """u_i = [S_i, nu_i, (a0_i - g) / A, deg_i / C, 0...] (unrounded)."""
    n = len(nu)
    u = np.zeros((n, NODE_IN))
    u[:, :EDGE_W] = S
    u[:, EDGE_W:EDGE_W + 2] = nu
    u[:, EDGE_W + 2:EDGE_W + 5] = (np.asarray(acc0, np.float64)[:, :3]
                                   - np.asarray(g, np.float64)[:3]) / ACC_SCALE
    u[:, EDGE_W + 5] = np.asarray(deg, np.float64) / AGG_C
    return u


def forward(flat, x, v, rho, prs, acc0, src, dst, deg, h, rho0, K, g, consts=None,
            bf16_round=False, keep=False):
    """Network output `out` (N x 32; columns 0..2 are the correction in units
    of R). acc0 (N x 3 or N x 4) is the forces stage's acceleration and g
    the constant acceleration. With keep=True also returns the
    intermediates the backward pass uses."""
    P = unpack(flat)
    R = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    n = len(rho)
    nu = node_inputs(rho, prs, rho0, K)
    e = R(edge_inputs(x, v, nu, src, dst, h, rho, prs, consts, rho0))
    z = e @ R(P["e.w1"]) + P["e.b1"]
    S = R(segsum(relu(z), dst, n) / AGG_C)
    u = R(node_vector(S, nu, acc0, deg, g))
    p1 = u @ R(P["n.w1"]) + P["n.b1"]
    y1 = R(silu(p1))
    p2 = y1 @ R(P["n.w2"]) + P["n.b2"]
    y2 = R(silu(p2))
    out = y2 @ R(P["n.w3"]) + P["n.b3"]
    c = dict(e=e, z=z, S=S, u=u, p1=p1, y1=y1, p2=p2, y2=y2, out=out)
    return (out, c) if keep else out


def acceleration(out, acc0):
    """a_i = a0_i + R out_i[0:3]."""
    return np.asarray(acc0, np.float64)[:, :3] + RES_SCALE * np.asarray(out)[:, :3]


def loss(out, target):
    """Mean over particles of the squared error of the three used outputs."""
    d = out[:, :3] - target
    return float((d * d).sum() / len(target))


def backward(flat, c, target, src, dst, deg):
    """Gradient of loss(forward(...), target) with respect to the flat
    parameter vector, from the intermediates of forward(..., keep=True)."""
    P = unpack(flat)
    n = len(target)
    G = {name: np.zeros(shape) for name, shape in LAYOUT}
    dout = np.zeros((n, OUT))
    dout[:, :3] = 2.0 * (c["out"][:, :3] - target) / n
    G["n.w3"] += c["y2"].T @ dout
    G["n.b3"] += dout.sum(0)
    dp2 = (dout @ P["n.w3"].T) * dsilu(c["p2"])
    G["n.w2"] += c["y1"].T @ dp2
    G["n.b2"] += dp2.sum(0)
    dp1 = (dp2 @ P["n.w2"].T) * dsilu(c["p1"])
    G["n.w1"] += c["u"].T @ dp1
    G["n.b1"] += dp1.sum(0)
    du = dp1 @ P["n.w1"].T
    dS = du[:, :EDGE_W] / AGG_C
    dz = dS[dst] * (c["z"] > 0.0)
    G["e.w1"] += c["e"].T @ dz
    G["e.b1"] += dz.sum(0)
    return pack(G)


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights over the inputs that
    are used, zero biases, zero rows for the padded inputs and an output
    layer scaled by 0.1 with only its first 3 columns nonzero."""
    rng = np.random.default_rng(seed)
    used_in = {"e.w1": FEAT_USED, "n.w1": NODE_IN_USED}
    P = {}
    for name, shape in LAYOUT:
        if len(shape) == 1:
            P[name] = np.zeros(shape)
            continue
        fan = used_in.get(name, shape[0])
        w = rng.uniform(-1.0, 1.0, size=shape) / np.sqrt(fan)
        w[fan:] = 0.0
        if name == "n.w3":
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
            if name == "n.b3":
                b[3:] = 0.0
            P[name] = b
    return pack(P)
