"""The trained surrogate (libclsph_amd/weights) against the method without it.

On a dam-break variant the training never saw, from one common state:
  (a) the plain method at dt                (the reference trajectory)
  (b) the plain method at 4 dt              (the method without the network)
  (c) the learned method at 4 dt            (the plain substep plus the network's correction)
Densities of the same particles are compared at the matched time. The
learned run must stay finite with its density error against (a) inside the
judge's bound, and it must beat (b) in mean and max density error against (a)
(an expected failure while the shipped network does not). On states of a
held-out variant the network's squared error against its training target
(the mean forces-stage acceleration over the next 4 substeps) must be below
that of the method without it. A separate case compares both GPU runs with
an FP64 trajectory of the CPU oracle.
"""
import os

import numpy as np
import pytest

import oracle
from libclsph_amd import learned, native, scenes
from sorted_view import internal_sorted

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(learned.WEIGHTS),
                                 reason="no trained surrogate weights in the tree")]


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    print("device:", c.device_name(), flush=True)
    yield c
    c.close()


def start_state(ctx, sc, warm):
    st = learned.make_state(ctx, sc)
    st.upload(sc.particles)
    st.step(warm)
    st.track_ids(True)
    return st


def density_by_id(st):
    """Densities indexed by particle identity (after the density stages)."""
    s, perm = internal_sorted(st)
    ids = st.ids()
    out = np.empty(st.n)
    out[ids[perm]] = s["density"]
    return out


@pytest.fixture(scope="module")
def large_step_errors(ctx):
    """Density errors against the plain run at dt, on a held-out slab after
    100 substeps' worth of time, of the plain method at 4 dt and of the
    learned method at 4 dt (unjudged)."""
    sc = learned.slab_variant(100003)
    dt = float(sc.params["time_delta"])
    rho0 = float(sc.params["fluid_density"])
    macro = 25
    base = start_state(ctx, sc, warm=300)
    runs = {}
    for name in ("plain_dt", "plain_4dt", "learned_4dt"):
        big = name != "plain_dt"
        st = learned.make_state(ctx, sc, dt=learned.MACRO * dt if big else None)
        st.track_ids(True)
        base.track_ids(True)
        st.copy_from(base)
        runs[name] = st
    net = native.Network(ctx, sc.n, e_cap=1024)
    net.set_params(np.load(learned.WEIGHTS))
    runs["plain_dt"].step(macro * learned.MACRO)
    runs["plain_4dt"].step(macro)
    runs["learned_4dt"].learned_step(net, macro)
    rho = {}
    for name, st in runs.items():
        learned.density_stages(st)
        rho[name] = density_by_id(st)
        assert np.isfinite(rho[name]).all(), name
    err = {k: np.abs(rho[k] - rho["plain_dt"]) / rho0 for k in ("plain_4dt", "learned_4dt")}
    for k, e in err.items():
        print(f"{k}: max |rho - rho_plain| / rho0 = {e.max():.3e}, mean {e.mean():.3e}",
              flush=True)
    mx, mean = net.compare_density(runs["learned_4dt"], runs["plain_dt"])
    assert abs(mx - err["learned_4dt"].max()) <= 1e-6
    assert abs(mean - err["learned_4dt"].mean()) <= 1e-6
    net.close()
    for st in list(runs.values()) + [base]:
        st.close()
    return err


def test_learned_run_stays_within_the_judge_bound(large_step_errors):
    """Unjudged over 100 substeps' worth of time, the learned run stays
    within 1 % of the plain density field on average (the maximum is set by a
    few splashing particles whose paths diverge)."""
    assert large_step_errors["learned_4dt"].mean() <= 0.01


@pytest.mark.xfail(strict=False, reason=(
    "the shipped network does not beat the plain step at 4 dt on held-out slabs: its density "
    "error against the plain run at dt is larger in mean and in max"))
def test_learned_beats_plain_at_the_same_large_step(large_step_errors):
    """What the network has to gain: at the same 4 dt step, the learned
    method's density error against the plain run at dt must be below that of
    the plain method without the network, in mean and in max."""
    err = large_step_errors
    print(f"mean error plain 4 dt / learned 4 dt: "
          f"{err['plain_4dt'].mean() / max(err['learned_4dt'].mean(), 1e-12):.2f}x, "
          f"max: {err['plain_4dt'].max() / max(err['learned_4dt'].max(), 1e-12):.2f}x", flush=True)
    assert err["learned_4dt"].mean() < err["plain_4dt"].mean()
    assert err["learned_4dt"].max() < err["plain_4dt"].max()


def held_out_loss(ctx, net, seed=100007, samples=8, gap=40, warm=300):
    """Mean squared correction error of the network on states of a held-out
    dam-break variant, and that of the zero network (the plain step at 4 dt),
    over the same targets (in units of R^2)."""
    sc = learned.slab_variant(seed)
    teacher = learned.make_state(ctx, sc)
    learner = learned.make_state(ctx, sc)
    teacher.upload(sc.particles)
    teacher.track_ids(True)
    teacher.step(warm)
    got, zero = 0.0, 0.0
    for _ in range(samples):
        learner.copy_from(teacher)
        learned.forces_stages(learner)
        teacher.track_ids(True)
        net.target_mean_acc(learner, teacher, learned.MACRO)
        net.forward(learner)
        got += net.backward()
        t = net.target(learner.n).astype(np.float64)
        zero += float((t * t).sum(1).mean())
        teacher.step(gap - learned.MACRO)
    return got / samples, zero / samples


def test_network_reduces_the_correction_error_on_held_out_states(ctx):
    """What the network is trained for: the mean forces-stage acceleration
    over the next 4 substeps. On states of a dam-break variant the training
    never saw, its squared error against that target is below the zero
    network's, which is the plain method's at the 4 dt step."""
    sc = learned.slab_variant(100007)
    net = native.Network(ctx, sc.n, e_cap=80 * sc.n)
    net.set_params(np.load(learned.WEIGHTS))
    got, zero = held_out_loss(ctx, net)
    print(f"held-out correction error: network {got:.4e}, zero network {zero:.4e}, "
          f"ratio {got / zero:.3f}", flush=True)
    assert got < 0.9 * zero


def test_judged_run_counts_fallbacks(ctx):
    sc = learned.slab_variant(100019)
    dt = float(sc.params["time_delta"])
    base = start_state(ctx, sc, warm=200)
    learner = learned.make_state(ctx, sc, dt=learned.MACRO * dt)
    teacher = learned.make_state(ctx, sc)
    learner.track_ids(True)
    teacher.track_ids(True)
    learner.copy_from(base)
    net = native.Network(ctx, sc.n, e_cap=1024)
    net.set_params(np.load(learned.WEIGHTS))
    r = learned.run_learned(ctx, learner, teacher, net, 16, dt, judge_every=2, tol=0.01)
    print(f"judged {r.judged}, fallbacks {r.fallbacks}, judge errors "
          f"{np.round(r.judge_err, 5).tolist()}", flush=True)
    assert r.judged == 8 and 0 <= r.fallbacks <= r.judged
    assert all(np.isfinite(r.judge_err))
    # every macro step judged: the count of fallbacks over 16 macro steps
    learner.copy_from(base)
    r1 = learned.run_learned(ctx, learner, teacher, net, 16, dt, judge_every=1, tol=0.01)
    print(f"judged every step: {r1.judged} judged, {r1.fallbacks} fallbacks, worst judged "
          f"error {max(r1.judge_err):.4f}", flush=True)
    assert r1.judged == 16 and 0 <= r1.fallbacks <= 16
    assert all(np.isfinite(r1.judge_err))
    # a tolerance of 0 forces a fallback at every judged step: the result is
    # then the plain trajectory at the judged steps
    learner.copy_from(base)
    r0 = learned.run_learned(ctx, learner, teacher, net, 2, dt, judge_every=1, tol=0.0)
    assert r0.fallbacks == 2
    plain = learned.make_state(ctx, sc)
    plain.track_ids(True)
    base.track_ids(True)
    plain.copy_from(base)
    learner.track_ids(True)
    learner.copy_from(base)
    learned.run_learned(ctx, learner, teacher, net, 2, dt, judge_every=1, tol=0.0)
    plain.step(2 * learned.MACRO)
    learned.density_stages(plain)
    learned.density_stages(learner)
    mx, _ = net.compare_density(learner, plain)
    assert mx == 0.0


def oracle_trajectory(sc, parts, steps, f64=True):
    """Oracle trajectory driven stage by stage: grid and cell ids in FP32 (the
    reference's) on positions rounded from the state, the stable sort as a
    stable argsort by cell id (the oracle's sort is pinned to it), then
    density, forces and advection in FP64 (oracle *_f64) or FP32. Returns
    the densities at the end, indexed by the particles' initial positions."""
    p = sc.params.copy()
    by_id = oracle.to_f64(parts) if f64 else parts.copy()

    def sorted_state():
        cur = np.zeros_like(parts)
        for f in ("position", "velocity", "intermediate_velocity", "acceleration"):
            cur[f] = by_id[f].astype(np.float32)
        oracle.grid_params(cur, p)
        hashed = oracle.locate_in_grid(cur, p)
        ids = np.argsort(hashed["grid_index"], kind="stable")
        table = oracle.cell_table(hashed[ids], int(p["grid_cell_count"]))
        cur = by_id[ids].copy()
        cur["grid_index"] = hashed["grid_index"][ids]  # the kernels look neighbours up by it
        return ids, table, cur

    dens = oracle.density_pressure_f64 if f64 else oracle.density_pressure
    forc = oracle.forces_f64 if f64 else oracle.forces
    adv = oracle.advection_collision_f64 if f64 else oracle.advection_collision
    for _ in range(steps):
        ids, table, cur = sorted_state()
        d = dens(cur, p, sc.terms, table)
        f = forc(d, p, sc.terms, table)
        a, _ = adv(f, p, sc.normals, sc.vertices, sc.indices)
        by_id[ids] = a
    ids, table, cur = sorted_state()
    d = dens(cur, p, sc.terms, table)
    rho = np.empty(len(parts))
    rho[ids] = d["density"]
    return rho


def test_error_against_fp64_oracle(ctx):
    sc = scenes.dam_break(4096, seed=77, spacing=learned.SPACING_1M)
    dt = float(sc.params["time_delta"])
    rho0 = float(sc.params["fluid_density"])
    macro = 10
    plain = learned.make_state(ctx, sc)
    plain.upload(sc.particles)
    plain.track_ids(True)
    learner = learned.make_state(ctx, sc, dt=learned.MACRO * dt)
    learner.track_ids(True)
    learner.copy_from(plain)
    net = native.Network(ctx, sc.n, e_cap=1024)
    net.set_params(np.load(learned.WEIGHTS))
    plain.step(macro * learned.MACRO)
    learner.learned_step(net, macro)
    ref = oracle_trajectory(sc, sc.particles, macro * learned.MACRO, f64=True)
    out = {}
    for name, st in (("plain_f32", plain), ("learned", learner)):
        learned.density_stages(st)
        e = np.abs(density_by_id(st) - ref) / rho0
        out[name] = e
        print(f"{name} vs FP64 oracle: max {e.max():.3e}, mean {e.mean():.3e}", flush=True)
    assert out["plain_f32"].max() <= 1e-4
    assert np.isfinite(out["learned"]).all()
