"""Training and running the learned force surrogate.

The surrogate (the network of libclsph_amd/csrc/nn.h) corrects each
particle's forces-stage acceleration a0 so that it stands for the mean
acceleration over one macro step of 4 substeps: the learned method advances
at dt_macro = 4 dt with one plain substep per macro step plus one network
evaluation, a = a0 + R out.

Training data comes from the plain GPU solver, generated online: each
trajectory is a seeded dam-break variant, advanced by the plain substep;
every `gap` substeps the current state is copied into a learner state,
which runs the plain substep's stages up to forces (a0), and the teacher's
next 4 plain substeps (with particle identities tracked) give each
particle's target

    out*_i = (mean over the 4 substeps of the forces stage's a_i - a0_i) / R

(gravity plus pressure, viscosity and surface tension along the plain
trajectory; collision impulses stay with advection, which applies them in
the learned step too). Each sample is one forward, one backward and one
Adam step, all on the GPU.

Running (run_learned): macro steps of the plain substep at 4 dt with the
network's correction. Every `judge_every` macro steps the plain method is the judge: the
same start is advanced by 4 plain substeps, both states get their densities,
and when max_i |rho_i - rho_i^plain| / rho0 exceeds `tol` (1 %), the macro
step is replaced by the plain result (a fallback).

    python -m libclsph_amd.learned --out weights.npy --trajectories 200 --minutes 8
    python -m libclsph_amd.learned --profile     (times the shipped network at 1M)
"""
import argparse
import math
import os
import sys
import time

import numpy as np

from . import native, scenes

MACRO = 4
SPACING_1M = (0.128 / 1048576) ** (1.0 / 3.0)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                       "surrogate_mi355x.npy")
DENSITY_STAGES = ("grid", "hash", "sort", "cell_table", "density")
FORCES_STAGES = DENSITY_STAGES + ("forces",)

# Parameter vector layout of the network (libclsph_amd/csrc/nn.h, enum Param):
# (name, shape, used input rows) in storage order, row-major inputs x outputs.
EW, NW, FEAT, IN, OUT = 32, 64, 32, 64, 32
LAYOUT = [
    ("e.w1", (FEAT, EW), 27), ("e.b1", (EW,), 0),
    ("n.w1", (IN, NW), 38), ("n.b1", (NW,), 0),
    ("n.w2", (NW, NW), NW), ("n.b2", (NW,), 0),
    ("n.w3", (NW, OUT), NW), ("n.b3", (OUT,), 0),
]


def init_params(seed=0):
    """Initial parameters: weights U(-1, 1) / sqrt(fan_in) over the used
    input rows (padded rows zero), biases zero, and the output layer scaled
    by 0.1 with only its first 3 columns nonzero."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape, used in LAYOUT:
        if len(shape) == 1:
            parts.append(np.zeros(shape))
            continue
        w = rng.uniform(-1.0, 1.0, size=shape) / np.sqrt(used)
        w[used:] = 0.0
        if name == "n.w3":
            w[:, 3:] = 0.0
            w *= 0.1
        parts.append(w)
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def slab_variant(seed, spacing=SPACING_1M, depth_sites=12):
    """A dam-break variant at the spacing of the 1,048,576-particle dam
    break: block b x 2b x d (d = depth_sites spacings) with b ~ U(0.3, 0.5) m
    in the corner of a 2.5 b x 2.5 b x 2.5 d box, as the 1M block fills 40 %
    of its box along x and z (walls behind it, free faces in front); jitter
    ~ U(0, 0.02) spacings and dynamic viscosity ~ U(1e-3, 1e-2). Twelve sites
    in depth leave the middle layers with complete neighbourhoods (radius
    h = 2 spacings) under the full column height."""
    rng = np.random.default_rng(seed)
    b = float(rng.uniform(0.3, 0.5))
    s = np.float32(spacing)
    nx = int(round(b / float(s)))
    ny = int(round(2.0 * b / float(s)))
    nz = int(depth_sites)
    n = nx * ny * nz
    depth = nz * float(s)
    jitter = float(rng.uniform(0.0, 0.02))
    visc = float(rng.uniform(1e-3, 1e-2))
    h = np.float32(2.0) * s
    mass = np.float32(1000.0 * float(s) ** 3)
    parts = native.make_block(n, float(s), (0.0, 0.0, 0.0), (nx * float(s), 0.0, depth),
                              jitter, int(seed) * 7919 + 11)
    params = scenes.fluid_params(n, h, mass, viscosity=visc)
    terms = native.kernel_terms(float(h))
    normals, verts, idx = native.make_box((0.0, 0.0, 0.0), (2.5 * b, 2.5 * b, 2.5 * depth))
    return scenes.Scene(parts, params, terms, normals, verts, idx)


def make_state(ctx, sc, n_cap=None, dt=None):
    st = native.State(ctx, n_cap or sc.n)
    p = sc.params.copy()
    if dt is not None:
        p["time_delta"] = np.float32(dt)
    st.set_params(p, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    return st


def density_stages(st):
    for name in DENSITY_STAGES:
        st.stage(name)


def forces_stages(st):
    for name in FORCES_STAGES:
        st.stage(name)


def train(net, ctx, trajectories=200, steps=2000, gap=20, minutes=None, lr=1e-3, lr_end=1e-5,
          seed=0, log=print, warm=50):
    """Online training over plain-solver trajectories; returns the loss
    history. Stops early after `minutes` of wall-clock time."""
    t0 = time.perf_counter()
    samples_per_traj = max(1, (steps - warm) // gap)
    total = trajectories * samples_per_traj
    losses = []
    k = 0
    skipped = 0
    for tr in range(trajectories):
        sc = slab_variant(seed + tr)
        teacher = make_state(ctx, sc)
        learner = make_state(ctx, sc)
        teacher.upload(sc.particles)
        teacher.track_ids(True)
        teacher.step(warm)
        for smp in range(samples_per_traj):
            learner.copy_from(teacher)
            forces_stages(learner)
            teacher.track_ids(True)
            net.target_mean_acc(learner, teacher, MACRO)  # advances the teacher by MACRO
            try:
                net.forward(learner)
            except native.ClsphError as err:  # a graph larger than the workspace
                if err.code != -3:
                    raise
                skipped += 1
                teacher.step(gap - MACRO)
                continue
            frac = k / max(1, total - 1)
            if minutes is not None:
                frac = max(frac, (time.perf_counter() - t0) / (60.0 * minutes))
            frac = min(1.0, frac)
            rate = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * frac))
            loss = net.backward(read_loss=(smp % 10 == 0))
            net.adam(rate)
            if loss is not None:
                losses.append((k, loss))
            k += 1
            teacher.step(gap - MACRO)
        teacher.close()
        learner.close()
        last = losses[-1][1] if losses else float("nan")
        log(f"trajectory {tr + 1}/{trajectories} (n={sc.n}): {k} samples ({skipped} skipped), "
            f"loss {last:.4e}, {time.perf_counter() - t0:.0f} s")
        if minutes is not None and time.perf_counter() - t0 > 60.0 * minutes:
            break
    return losses


class Run:
    """Outcome of run_learned."""

    def __init__(self):
        self.macro_steps = 0
        self.judged = 0
        self.fallbacks = 0
        self.judge_err = []
        self.judge_mean = []
        self.wall_s = 0.0
        self.sim_s = 0.0


def run_learned(ctx, learner, teacher, net, macro_steps, dt, judge_every=8, tol=0.01):
    """Advances `learner` (parameters at dt_macro = 4 dt) by `macro_steps`
    macro steps with the network, judged every `judge_every` macro steps by
    4 plain substeps of `teacher` (parameters at dt) from the same start
    (clsph_learned_run). Both states must track identities."""
    r = Run()
    ctx.sync()
    t0 = time.perf_counter()
    r.judged, r.fallbacks, r.judge_err, r.judge_mean = net.learned_run(
        learner, teacher, macro_steps, judge_every, tol)
    ctx.sync()
    r.wall_s = time.perf_counter() - t0
    r.macro_steps = macro_steps
    r.sim_s = macro_steps * MACRO * dt
    return r


def profile_network(substeps=20, reps=5):
    """Times the fused network on the 1,048,576-particle dam break after
    `substeps` plain substeps, every kernel launched directly (no graphs), for
    kernel counters; returns (ms per evaluation, edges)."""
    ctx = native.Context(0)
    sc = scenes.dam_break(1048576)
    st = native.State(ctx, sc.n)
    st.set_params(sc.params, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    st.upload(sc.particles)
    if substeps:
        st.profile_steps(substeps)
    forces_stages(st)
    net = native.Network(ctx, sc.n, e_cap=1024)
    net.set_params(np.load(WEIGHTS))
    return net.profile(st, reps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile", action="store_true",
                    help="time the shipped network on the 1M dam break instead of training")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--trajectories", type=int, default=200)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--gap", type=int, default=20)
    ap.add_argument("--minutes", type=float, default=None)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these weights")
    a = ap.parse_args()
    if a.profile:
        ms, edges = profile_network()
        print(f"network {ms:.4f} ms per evaluation, {edges} edges", flush=True)
        return
    ctx = native.Context(0)
    n_cap = 101 * 202 * 12  # largest slab variant
    net = native.Network(ctx, n_cap, e_cap=80 * n_cap)
    net.set_params(np.load(a.init) if a.init else init_params(a.seed))
    losses = train(net, ctx, trajectories=a.trajectories, steps=a.steps, gap=a.gap,
                   minutes=a.minutes, lr=a.lr, seed=a.seed,
                   log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, net.params())
    np.save(os.path.splitext(a.out)[0] + "_loss.npy", np.asarray(losses, dtype=np.float64))
    print(f"wrote {a.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
