"""Learned force surrogate on the GPU against its CPU oracle (oracle/learned.py).

  GEMM (all operand layouts, split K)        : 1e-5 of sum |a||b| per entry
  neighbour graph                             : the density kernel's r < h set
  forward vs the oracle with bf16 rounding    : output within 1 % of its scale
  forward vs the exact FP64 network           : bf16 error reported, <= 5 %
  gradients vs the FP64 oracle's backward     : per tensor, relative L2 <= 5 %
  fused inference vs the training forward     : within 1 % of scale, rms 1e-3
  training targets, identities                : closed form, 1e-5
  zero network                                : the forces stage, bit for bit;
                                                the plain substep within 1e-4 max|a| dt
"""
import numpy as np
import pytest

from libclsph_amd import native, scenes
from oracle import learned as L
from sorted_view import internal_sorted

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    print("device:", c.device_name(), flush=True)
    yield c
    c.close()


def bits(x):
    return L.bf16_bits(x)


@pytest.mark.parametrize("ta,tb,acc,m,n,k", [
    (False, False, False, 200, 136, 72),
    (False, True, False, 200, 136, 72),
    (True, False, False, 200, 136, 72),
    (True, True, False, 200, 136, 72),
    (False, False, True, 64, 32, 16),
    (True, False, True, 128, 128, 9000),   # split over K, like a weight gradient
    (True, False, True, 16, 128, 5000),
])
def test_gemm(ctx, ta, tb, acc, m, n, k):
    print(f"gemm ta={ta} tb={tb} acc={acc} {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m * 7 + n + k)
    a = rng.normal(size=(k, m) if ta else (m, k)).astype(np.float32)
    b = rng.normal(size=(n, k) if tb else (k, n)).astype(np.float32)
    c0 = rng.normal(size=(m, n)).astype(np.float32) if acc else np.zeros((m, n), np.float32)
    bias = None if acc else rng.normal(size=n).astype(np.float32)
    ab, bb = bits(a), bits(b)
    got = native.nn_gemm(ctx, ab, bb, c0.copy(), m, n, k, ta, tb, acc, bias)
    A = L.bits_to_float(ab).reshape(a.shape)
    B = L.bits_to_float(bb).reshape(b.shape)
    A = A.T if ta else A
    B = B.T if tb else B
    ref = A @ B + (c0 if acc else bias[None, :])
    bound = np.abs(A) @ np.abs(B) + 1.0
    assert np.all(np.abs(got - ref) <= 1e-5 * bound), np.abs(got - ref).max()


def settled_state(ctx, n=3000, steps=30, seed=5):
    sc = scenes.dam_break(n, seed=seed)
    st = native.State(ctx, sc.n)
    st.set_params(sc.params, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    st.upload(sc.particles)
    st.step(steps)
    for name in ("grid", "hash", "sort", "cell_table", "density", "forces"):
        st.stage(name)
    return sc, st


def oracle_inputs(sc, st, net):
    s, _ = internal_sorted(st)
    off, src = net.graph()
    deg = np.diff(off.astype(np.int64))
    dst = np.repeat(np.arange(st.n), deg)
    p, t = sc.params, sc.terms
    consts = (p["particle_mass"], p["dynamic_viscosity"], t["spiky"], t["viscosity"],
              t["poly_6_gradient"], t["poly_6_laplacian"])
    return dict(x=s["position"][:, :3], v=s["velocity"][:, :3], rho=s["density"],
                prs=s["pressure"], acc0=s["acceleration"][:, :3].astype(np.float64),
                src=src.astype(np.int64), dst=dst, deg=deg,
                h=float(p["h"]), rho0=float(p["fluid_density"]), K=float(p["K"]),
                g=p["constant_acceleration"][:3].astype(np.float32).astype(np.float64),
                consts=tuple(float(np.float32(c)) for c in consts))


def test_graph_is_the_density_neighbourhood(ctx):
    sc, st = settled_state(ctx)
    net = native.Network(ctx, st.n)
    net.set_params(np.zeros(net.count, np.float32))
    e = net.forward(st)
    off, src = net.graph()
    s, _ = internal_sorted(st)
    x = s["position"][:, :3].astype(np.float32)
    h2 = np.float32(sc.params["h"]) ** 2
    assert off[-1] == e and e > 0
    for i in range(0, st.n, 7):
        d = x[i] - x
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        want = set(np.nonzero(d2 < h2)[0]) - {i}
        got = set(src[off[i]:off[i + 1]].tolist())
        fuzzy = set(np.nonzero(np.abs(d2 - h2) <= 1e-5 * h2)[0])
        assert (want ^ got) <= fuzzy, i
        assert len(got) == off[i + 1] - off[i]


def test_forward_matches_oracle(ctx):
    sc, st = settled_state(ctx)
    net = native.Network(ctx, st.n)
    flat = L.random_params(3)
    net.set_params(flat.astype(np.float32))
    e = net.forward(st)
    print(f"n={st.n} edges={e}", flush=True)
    gpu = net.output().astype(np.float64)
    args = oracle_inputs(sc, st, net)
    # the GPU holds float32 parameters and rounds matrix operands to bf16
    ref = L.forward(flat.astype(np.float32).astype(np.float64), bf16_round=True, **args)
    exact = L.forward(flat.astype(np.float32).astype(np.float64), bf16_round=False, **args)
    scale = np.abs(ref[:, :3]).max()
    err = np.abs(gpu[:, :3] - ref[:, :3])
    rms = np.sqrt((err ** 2).mean()) / np.sqrt((ref[:, :3] ** 2).mean())
    ex = np.sqrt(((gpu[:, :3] - exact[:, :3]) ** 2).mean()) / np.sqrt((exact[:, :3] ** 2).mean())
    print(f"max|gpu-oracle_bf16|/scale={err.max() / scale:.2e} rel.rms={rms:.2e} "
          f"rel.rms vs fp64={ex:.2e}", flush=True)
    assert err.max() <= 1e-2 * scale
    assert rms <= 2e-3
    assert ex <= 5e-2
    assert np.all(gpu[:, 3:] == 0.0)


def test_gradients_match_oracle(ctx):
    sc, st = settled_state(ctx, n=2000)
    net = native.Network(ctx, st.n)
    flat = L.random_params(4).astype(np.float32).astype(np.float64)
    net.set_params(flat.astype(np.float32))
    net.forward(st)
    rng = np.random.default_rng(0)
    target = rng.normal(0, 0.5, size=(st.n, 3)).astype(np.float32)
    net.set_target(target)
    loss = net.backward()
    g = net.grad().astype(np.float64)
    args = oracle_inputs(sc, st, net)
    out, c = L.forward(flat, bf16_round=True, keep=True, **args)
    ref_loss = L.loss(out, target.astype(np.float64))
    gref = L.backward(flat, c, target.astype(np.float64), args["src"], args["dst"], args["deg"])
    print(f"loss gpu={loss:.6e} oracle={ref_loss:.6e}", flush=True)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    G, R = L.unpack(g), L.unpack(gref)
    worst = 0.0
    for name, _ in L.LAYOUT:
        gn, rn = G[name].ravel(), R[name].ravel()
        nr = np.linalg.norm(rn)
        if nr == 0.0:
            assert np.all(gn == 0.0), name
            continue
        rel = np.linalg.norm(gn - rn) / nr
        worst = max(worst, rel)
        print(f"  {name:6s} |g|={nr:.3e} rel.err={rel:.2e}", flush=True)
        assert rel <= 5e-2, name
    print(f"worst relative gradient error {worst:.2e}", flush=True)


def test_target_and_identities(ctx):
    sc, st = settled_state(ctx, n=2500, steps=1500)
    teacher = native.State(ctx, st.n)
    teacher.set_params(sc.params, sc.terms)
    teacher.scene(sc.normals, sc.vertices, sc.indices)
    teacher.copy_from(st)
    teacher.track_ids(True)
    teacher.step(4)
    ids = teacher.ids()
    assert np.array_equal(np.sort(ids), np.arange(st.n))
    net = native.Network(ctx, st.n)
    dt = 4.0 * float(sc.params["time_delta"])
    net.n = st.n
    net.target_from(st, teacher, dt)
    got = net.target(st.n).astype(np.float64)
    # host restatement: learner's sorted i <- canonical perm[i]; the teacher
    # holds that particle at loc[perm[i]]
    _, perm = st.order()
    perm = perm.astype(np.int64)
    canon = st.download()
    t_parts = teacher.download()
    loc = np.empty(st.n, np.int64)
    loc[ids] = np.arange(st.n)
    a0 = internal_sorted(st)[0]["acceleration"][:, :3].astype(np.float32)
    vs = canon["intermediate_velocity"][:, :3].astype(np.float32)
    vt = t_parts["intermediate_velocity"][:, :3].astype(np.float32)
    want = ((vt[loc[perm]] - vs[perm]).astype(np.float32) * np.float32(1.0 / dt)
            - a0) / np.float32(2.0)
    # the target is a difference of terms of size |a0|: float32 rounding of
    # those terms sets the tolerance
    scale = max(np.abs(a0).max(), np.abs(want).max())
    assert np.abs(got - want).max() <= 1e-5 * scale
    assert np.abs(want).max() > 1e-3 * np.abs(a0).max()  # the fluid interacts
    # a particle moves continuously: over 4 substeps it stays within 4 dt vmax
    x0 = canon["position"][:, :3]
    x1 = t_parts["position"][:, :3][loc]
    assert np.abs(x1 - x0).max() < 4 * float(sc.params["time_delta"]) * 5.0


def test_zero_network_leaves_the_forces_stage_unchanged(ctx):
    """With every parameter zero the correction is exactly zero: inference
    after the forces stage leaves its accelerations bit for bit."""
    sc, st = settled_state(ctx, n=3000, steps=200, seed=21)
    a0 = internal_sorted(st)[0]["acceleration"].copy()
    net = native.Network(ctx, st.n)
    net.set_params(np.zeros(net.count, np.float32))
    net.infer(st)
    assert np.array_equal(internal_sorted(st)[0]["acceleration"], a0)


def mixed_cloud(seed=6):
    """300 particles inside one smoothing radius (their neighbour lists
    overflow) beside 700 spread ones."""
    rng = np.random.default_rng(seed)
    sc = scenes.dam_break(1000)
    parts = sc.particles.copy()
    pos = np.concatenate([rng.uniform(0.49, 0.51, size=(300, 3)),
                          rng.uniform(0.2, 0.8, size=(700, 3))])
    parts["position"][:, :3] = pos.astype(np.float32)
    parts["velocity"][:, :3] = rng.normal(0, 0.05, size=(1000, 3)).astype(np.float32)
    parts["intermediate_velocity"][:, :3] = parts["velocity"][:, :3]
    p = sc.params.copy()
    h = np.float32(0.04)
    p["h"] = h
    p["max_velocity"] = np.float32(0.8) * h / p["time_delta"]
    return scenes.Scene(parts, p, native.kernel_terms(float(h)), sc.normals, sc.vertices,
                        sc.indices)


@pytest.mark.parametrize("case", ["dam_break", "overflowing_lists"])
def test_learned_step_sums_the_forces_stage(ctx, case):
    """The learned step sums the forces stage's pair terms inside the network
    kernel (the forces kernel runs only for particles whose neighbour list
    overflowed). With a zero network one learned step is the plain substep
    up to the order of that summation: velocities agree within 1e-4 of
    max |a| dt, the forces parity bound."""
    sc = scenes.dam_break(3000, seed=2) if case == "dam_break" else mixed_cloud()
    warm = 1500 if case == "dam_break" else 0
    a = native.State(ctx, sc.n)
    b = native.State(ctx, sc.n)
    for s in (a, b):
        s.set_params(sc.params, sc.terms)
        s.scene(sc.normals, sc.vertices, sc.indices)
        s.upload(sc.particles)
        s.step(warm)
    for name in ("grid", "hash", "sort", "cell_table", "density", "forces"):
        b.stage(name)
    amag = np.linalg.norm(internal_sorted(b)[0]["acceleration"][:, :3].astype(np.float64), axis=1)
    # the bound follows the particles summed in the network kernel; those of
    # the dense ball (|a| ~ 1e16, lists overflowing) take the forces kernel's
    # own cell walk and must agree to rounding
    amax = float(amag[amag < 1e6].max())
    if case == "overflowing_lists":
        assert b.neighbour_pairs() > 300 * 128  # the ball's lists overflow
        assert (amag >= 1e6).sum() >= 300
    b.stage("advect")
    net = native.Network(ctx, sc.n)
    net.set_params(np.zeros(net.count, np.float32))
    a.learned_step(net, 1)
    pa, pb = a.download(), b.download()
    dt = float(sc.params["time_delta"])
    vb = pb["intermediate_velocity"][:, :3]
    dvp = np.abs(pa["intermediate_velocity"][:, :3] - vb)
    dxp = np.abs(pa["position"][:, :3] - pb["position"][:, :3])
    # per particle: the acceleration bound plus the rounding of v + a dt
    # (2 ulp of the particle's |v|)
    ulp_v = 2.0 * np.spacing(np.abs(vb).max(axis=1).astype(np.float32)).astype(np.float64)
    bound = 1e-4 * amax * dt + ulp_v
    ulp_x = 2.0 * np.spacing(np.abs(pb["position"][:, :3]).max(axis=1).astype(np.float32))
    print(f"max |a| {amax:.3e}, velocity difference {dvp.max():.3e} (bound {bound.min():.3e}"
          f"..{bound.max():.3e}), position difference {dxp.max():.3e}", flush=True)
    assert np.all(dvp.max(axis=1) <= bound)
    assert np.all(dxp.max(axis=1) <= bound * dt + ulp_x)


def test_compare_density_of_a_copy_is_zero(ctx):
    sc, st = settled_state(ctx, n=2000, steps=5)
    other = native.State(ctx, st.n)
    other.set_params(sc.params, sc.terms)
    other.scene(sc.normals, sc.vertices, sc.indices)
    other.copy_from(st)
    st.track_ids(True)
    other.track_ids(True)
    for s in (st, other):
        s.step(2)
        for name in ("grid", "hash", "sort", "cell_table", "density"):
            s.stage(name)
    net = native.Network(ctx, st.n)
    mx, mean = net.compare_density(st, other)
    assert mx == 0.0 and mean == 0.0


def test_training_reduces_loss(ctx):
    sc, st = settled_state(ctx, n=3000, steps=20)
    teacher = native.State(ctx, st.n)
    teacher.set_params(sc.params, sc.terms)
    teacher.scene(sc.normals, sc.vertices, sc.indices)
    teacher.copy_from(st)
    teacher.track_ids(True)
    teacher.step(4)
    net = native.Network(ctx, st.n)
    net.set_params(L.init_params(0).astype(np.float32))
    net.forward(st)
    net.target_from(st, teacher, 4.0 * float(sc.params["time_delta"]))
    losses = []
    for _ in range(60):
        net.forward(st)
        losses.append(net.backward())
        net.adam(1e-3)
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]


def test_fused_inference_matches_training_forward(ctx):
    sc, st = settled_state(ctx, n=4000, steps=40, seed=9)
    net = native.Network(ctx, st.n)
    net.set_params(L.random_params(11).astype(np.float32))
    net.forward(st)
    ref = net.output().astype(np.float64)
    a0 = internal_sorted(st)[0]["acceleration"][:, :3].astype(np.float64)
    net.infer(st)
    got = net.output().astype(np.float64)
    a1 = internal_sorted(st)[0]["acceleration"][:, :3].astype(np.float64)
    # the inference added R out to the forces stage's accelerations, once
    assert np.allclose(a1, a0 + 2.0 * got[:, :3], rtol=1e-6, atol=1e-5)
    with pytest.raises(native.ClsphError):
        net.infer(st)
    scale = np.abs(ref[:, :3]).max()
    err = np.abs(got[:, :3] - ref[:, :3])
    rms = np.sqrt((err ** 2).mean()) / np.sqrt((ref[:, :3] ** 2).mean())
    print(f"fused vs training forward: max {err.max() / scale:.2e} of scale, rel.rms {rms:.2e}",
          flush=True)
    assert err.max() <= 1e-2 * scale
    assert rms <= 1e-3


def test_mean_acceleration_target(ctx):
    sc, st = settled_state(ctx, n=2500, steps=1500, seed=13)
    teacher = native.State(ctx, st.n)
    twin = native.State(ctx, st.n)
    for s in (teacher, twin):
        s.set_params(sc.params, sc.terms)
        s.scene(sc.normals, sc.vertices, sc.indices)
        s.copy_from(st)
        s.track_ids(True)
    net = native.Network(ctx, st.n)
    net.target_mean_acc(st, teacher, 4)
    got = net.target(st.n).astype(np.float64)
    # host restatement: the twin runs the same 4 substeps stage by stage and
    # its forces-stage accelerations are summed by identity
    total = np.zeros((st.n, 3))
    for _ in range(4):
        for name in ("grid", "hash", "sort", "cell_table", "density", "forces"):
            twin.stage(name)
        parts, perm = internal_sorted(twin)
        ids = twin.ids()
        total[ids[perm]] += parts["acceleration"][:, :3]
        twin.stage("advect")
    _, perm_l = st.order()
    a0 = internal_sorted(st)[0]["acceleration"][:, :3].astype(np.float64)
    want = (total[perm_l.astype(np.int64)] / 4.0 - a0) / 2.0
    scale = max(np.abs(a0).max(), np.abs(want).max())
    assert np.abs(got - want).max() <= 1e-5 * scale
    assert np.abs(want).max() > 1e-3 * np.abs(a0).max()
    # the teacher ended where the twin did
    assert np.array_equal(teacher.download()["position"], twin.download()["position"])


def test_judged_run_checks_its_states(ctx):
    """The main run needs identities in both states and a judge at a
    quarter of the learner's time step; it runs without a judge when
    judge_every is 0."""
    sc = scenes.dam_break(2048, seed=4)
    dt = float(sc.params["time_delta"])
    states = {}
    for name, step in (("learner", 4 * dt), ("judge", dt), ("wrong", 2 * dt)):
        s = native.State(ctx, sc.n)
        p = sc.params.copy()
        p["time_delta"] = np.float32(step)
        s.set_params(p, sc.terms)
        s.scene(sc.normals, sc.vertices, sc.indices)
        s.upload(sc.particles)
        states[name] = s
    net = native.Network(ctx, sc.n)
    net.set_params(np.zeros(net.count, np.float32))
    with pytest.raises(native.ClsphError):  # no identities yet
        net.learned_run(states["learner"], states["judge"], 2, 1)
    for s in states.values():
        s.track_ids(True)
    with pytest.raises(native.ClsphError):
        net.learned_run(states["learner"], states["wrong"], 2, 1)
    judged, fallbacks, errs, means = net.learned_run(states["learner"], states["judge"], 2, 1)
    assert judged == 2 and len(errs) == 2 and len(means) == 2
    assert all(0.0 <= m <= e for m, e in zip(means, errs))
    assert net.learned_run(states["learner"], None, 3, 0)[:2] == (0, 0)
