"""Throughput of the plain SPH substep on one MI355X.

Default workload: the 1,048,576-particle dam break (0.4 x 0.8 x 0.4 m block
at spacing 0.5 h in a 1 m box, dt 1e-4 s). W untimed warm-up substeps, then
K timed substeps replayed from one captured HIP graph, with state resident
in HBM. Prints one JSON line with particle-steps/s, the roofline of the
dominant kernel (timed live with HIP events on the library stream) and a CPU
baseline (the oracle on the host cores over a bounded sample).

--workload heightfield: 4,194,304 particles poured from a random-velocity
column onto the 128 x 128 heightfield (32,768 triangles + 10 wall
triangles), dt 5e-5 s, 50 + 450 substeps.

--workload batched: 64 independent dam breaks (counts log-uniform in
[4,096, 262,144], viscosity uniform in [1e-3, 5e-2], spacing of the 1M case)
advanced together, one launch per stage, work-group sizes from the autotuner
table; reports aggregate particle-steps/s, the per-scene latency (every scene
completes with the batched substep) and the time of the same substep run
scene by scene.

    python bench.py [--gpus 1] [--steps 900] [--warmup 100] [--workload dam_break]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per particle for each stage (FP32 SoA, each logical
# array read or written once; positions/velocities count 12 B).
#   hash     : x 12 read, key 4 + index 4 written
#   sort     : 3 LSD passes x (key+index 8 read + 8 written), then the gather
#              of x, v, v_half (index 4 + 36 read, 36 written)
#   cell tab : sorted keys 4 read + 4 * G / N table written
#   density  : x 12 read, rho and p 8 written
#   forces   : x, v 24 + rho, p 8 read, a 12 written
#   advect   : x, v_half, a 36 read, x, v, v_half 36 written
SORT_PASSES = 3


def stage_bytes(n, cells):
    return {
        "grid": 0.0,
        "hash": 20.0 * n,
        "sort": (16.0 * SORT_PASSES + 76.0) * n,
        "cell_table": 4.0 * n + 4.0 * cells,
        "density": 20.0 * n,
        "forces": 44.0 * n,
        "advect": 72.0 * n,
    }


STAGE_KERNEL = {"hash": "clsph::k_hash", "density": "clsph::k_density",
                "forces": "clsph::k_forces", "advect": "clsph::k_advect"}

# VALU issue: a wave64 VALU instruction occupies one of the 1,024 SIMDs (256
# CUs x 4) for 4 cycles at up to 2.4 GHz.
SIMDS = 1024
CLOCK_HZ = 2.4e9


def measured_counters(stage, workload, n):
    """HBM bytes and VALU instructions per launch of the stage's kernel from
    profiles/counters.json (rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a
    copy kernel of known size, and SQ_INSTS_VALU), when those counters were
    taken on the sources being run (libclsph_amd.counters.source_stamp) and
    this workload; else {}."""
    from libclsph_amd import counters
    data = counters.load()
    if data is None or workload != "dam_break" or data.get("workload") != f"dam_break {n}":
        return {}
    return data["kernels"].get(STAGE_KERNEL.get(stage, ""), {})


def roofline(dom, achieved, ms, workload, n):
    """The dominant kernel against the HBM peak (algorithmic bytes over its
    measured duration) and, from the committed counters of the same sources,
    its measured HBM traffic and its VALU issue fraction (VALU instructions x
    4 cycles over the SIMDs' cycles in its duration): the bound it meets.
    The counter fractions are taken over the kernel's duration in the
    counter run itself (its sample of substeps), not over this run's."""
    c = measured_counters(dom, workload, n)
    out = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": c.get("traffic")}
    cms = c["us"] * 1e-3 if c.get("us") else ms
    if c.get("traffic"):
        out["traffic_frac"] = c["traffic"] / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS
    if c.get("valu"):
        out["valu_issue_frac"] = c["valu"] * 4.0 / (SIMDS * CLOCK_HZ * cms * 1e-3)
        out["limiter"] = "valu" if out["valu_issue_frac"] > out.get("traffic_frac", 0.0) else "hbm"
    if c:
        from libclsph_amd import counters
        out["counter_sample"] = counters.load().get("sample")
        out["counter_kernel_ms"] = cms
    return out


def cpu_baseline(scene, label, budget_s=20.0, max_steps=3):
    """The oracle (plain C port of the reference substep, OpenMP over the
    host cores, collisions tested against every triangle as the reference
    does) on a bounded number of substeps of `scene`."""
    import oracle
    parts = scene.particles.copy()
    params = scene.params.copy()
    params["particles_count"] = scene.n
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        parts, _, _, _ = oracle.run(parts, params, scene.terms, scene.normals,
                                    scene.vertices, scene.indices, 1)
        steps += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    dt = time.perf_counter() - t0
    return {
        "value": scene.n * steps / dt,
        "unit": "particle-steps/s",
        "cores": oracle.threads(),
        "kind": "port",
        "sample": f"{steps} substep(s) of {label} from its initial state, {dt:.1f} s on the host",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["dam_break", "heightfield", "batched"],
                    default="dam_break")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="substeps of the per-stage timing (0: the timed substeps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--direct", action="store_true",
                    help="launch every substep stage by stage instead of replaying its captured "
                         "graph, timing the stages on the way (for kernel tracing; implies "
                         "--no-learned)")
    ap.add_argument("--track-ids", action="store_true",
                    help="also track particle identities each substep; its copy kernel "
                         "(k_copy_u32, 4 n bytes read and written) calibrates HBM counters")
    ap.add_argument("--learned-macro-steps", type=int, default=100)
    ap.add_argument("--judge-every", type=int, default=16,
                    help="learned macro steps between checks against the plain method")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the plain SPH substep runs on a single "
              "MI355X in this build (multi-GPU sharding is a design, not a measured path)",
              file=sys.stderr)
        sys.exit(2)

    from libclsph_amd import native, scenes

    if args.workload == "batched":
        return bench_batched(args)

    defaults = {"dam_break": (1048576, 900, 100), "heightfield": (4194304, 450, 50)}
    n0, k0, w0 = defaults[args.workload]
    n = args.particles or n0
    args.steps = k0 if args.steps is None else args.steps
    args.warmup = w0 if args.warmup is None else args.warmup
    if args.workload == "dam_break":
        scene = scenes.dam_break(n)
        metric = "particle-steps/s @1M particles (plain SPH substep)"
        cfg = {"workload": "dam_break", "particles": n}
    else:
        scene = scenes.heightfield_pour(n)
        metric = "particle-steps/s @4M particles poured on a 32,768-triangle heightfield"
        cfg = {"workload": "heightfield_pour", "particles": n,
               "triangles": int(len(scene.indices) // 3)}
    ctx = native.Context(0)
    st = native.State(ctx, n)
    st.set_params(scene.params, scene.terms)
    st.scene(scene.normals, scene.vertices, scene.indices)
    st.upload(scene.particles)
    print(f"# device {ctx.device_name()}; {n} particles", file=sys.stderr, flush=True)
    if args.track_ids:
        st.track_ids(True)

    if args.direct:
        # every substep launched stage by stage, HIP events between stages
        if args.warmup:
            st.profile_steps(args.warmup)
        ctx.sync()
        t0 = time.perf_counter()
        prof = st.profile_steps(args.steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        ms = wall * 1e3
        pairs = st.neighbour_pairs()
    else:
        if args.warmup:
            st.step(args.warmup)
        # the state at the start of the timed region, to time its stages below
        start = native.State(ctx, n)
        start.set_params(scene.params, scene.terms)
        start.scene(scene.normals, scene.vertices, scene.indices)
        start.copy_from(st)
        ctx.sync()
        t0 = time.perf_counter()
        ms = st.time_steps(args.steps)  # HIP events around K graph replays
        ctx.sync()
        wall = time.perf_counter() - t0
    ms_per_step = ms / args.steps
    value = n * args.steps / (ms / 1e3)
    print(f"# timed {args.steps} substeps: {ms:.2f} ms (events), {wall * 1e3:.2f} ms (wall)",
          file=sys.stderr, flush=True)

    if not args.direct:
        # Per-stage durations over the same substeps: the timed region
        # replayed stage by stage from its starting state, HIP events between
        # stages.
        prof = start.profile_steps(args.steps if args.profile_steps <= 0 else args.profile_steps)
        pairs = start.neighbour_pairs()  # found by the last replayed substep
        start.close()
    grid = st.grid()
    cells = int(grid["grid_cell_count"])
    nbytes = stage_bytes(n, cells)
    dom = max((k for k in prof if k != "grid"), key=lambda k: prof[k])
    achieved = nbytes[dom] / (prof[dom] * 1e-3) / 1e9
    stages = {k: {"ms": round(v, 5),
                  "GB/s": round(nbytes[k] / (v * 1e-3) / 1e9, 1) if v > 0 else None}
              for k, v in prof.items()}
    print("# stages " + json.dumps(stages), file=sys.stderr, flush=True)

    final = st.download()
    finite = bool(np.isfinite(final["position"][:, :3]).all())
    capped = st.bounce_capped()

    learned = None
    if args.workload == "dam_break" and not (args.no_learned or args.direct):
        learned = bench_learned(ctx, st, scene, ms_per_step, args.learned_macro_steps,
                                judge_every=args.judge_every)

    cfg.update({
        "dt": float(scene.params["time_delta"]),
        "timed_substeps": f"{args.warmup + 1}-{args.warmup + args.steps}",
        "launch": "stage by stage" if args.direct else "graph replay",
        "grid_cell_count_end": cells,
        "state_finite": finite,
        "bounce_capped_last": capped,
    })
    out = {
        "metric": metric,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "stages_ms": {k: v["ms"] for k, v in stages.items()},
        # density and forces are gather work: their pair-interaction rates,
        # with the neighbour pairs (r < h, ordered) of the last timed substep
        "pair_interactions": {
            "pairs_per_substep": pairs,
            "density_pairs_per_s": pairs / (prof["density"] * 1e-3),
            "forces_pairs_per_s": pairs / (prof["forces"] * 1e-3),
        },
        "roofline": roofline(dom, achieved, prof[dom], args.workload, n),
    }
    if learned is not None:
        out["learned"] = learned
    if not args.no_cpu_baseline:
        if args.workload == "dam_break":
            out["cpu_baseline"] = cpu_baseline(scene, f"the {n}-particle dam break")
        else:
            sub = scenes.Scene(scene.particles[:65536].copy(), scene.params.copy(), scene.terms,
                               scene.normals, scene.vertices, scene.indices)
            out["cpu_baseline"] = cpu_baseline(
                sub, f"the bottom 65,536 particles of the {n}-particle column (every particle "
                     f"tested against all {len(scene.indices) // 3} triangles)")
    print(json.dumps(out), flush=True)


# Learned surrogate: multiply-adds per evaluated edge and round in the fused
# edge kernel (layer 1: 16 inputs -> 128; layer 2: 128 -> 128).
# Matrix work of one surrogate evaluation over the used inputs: the edge
# layer (27 inputs -> 32) per edge and the node MLP (38 -> 64 -> 64 -> 3)
# per particle, 2 flops per multiply-add.
NN_SHAPE = ("1 round of message passing over the neighbours within h (at most 128): edge "
            "layer 27 -> 32 (ReLU), summed per particle; node MLP 38 -> 64 -> 64 -> 3 (ReLU); "
            "output corrects the forces stage's acceleration; bf16 weights and activations, "
            "fp32 accumulation")
NN_EDGE_FLOP = 2.0 * 27 * 32
NN_NODE_FLOP = 2.0 * (38 * 64 + 64 * 64 + 64 * 3)
MFMA_BF16_PEAK_TFLOPS = 2500.0
# Two-round network (clsph_gnn_*): per edge and round, 16 -> 128 -> 128;
# per particle, the two projections of H for round 2 and the third edge
# layer of each round (128 x 128 each), node MLP 1 (144 -> 128 -> 128 -> 128)
# and node MLP 2 (256 -> 128 -> 128 -> 32).
GNN_SHAPE = ("2 rounds of message passing over the neighbours within h (at most 128): edge "
             "MLP 16 -> 128 -> 128 (SiLU), summed, then 128 -> 128; node MLPs 144 -> 128 -> "
             "128 -> 128 and 256 -> 128 -> 128 -> 32 (SiLU); bf16 weights and activations, "
             "fp32 accumulation")
GNN_EDGE_FLOP = 2 * 2.0 * (16 * 128 + 128 * 128)
GNN_NODE_FLOP = 2.0 * (4 * 128 * 128 + 144 * 128 + 2 * 128 * 128 + 256 * 128 + 128 * 128
                       + 128 * 32)


def gnn_params(count, seed=0):
    """Seeded weights of the two-round network for timing (values do not
    change its work): U(-1, 1) / sqrt(128) for every entry."""
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1.0, 1.0, count) / np.sqrt(128.0)).astype(np.float32)


def bench_gnn(ctx, state, plain_ms_per_step):
    """Fused inference time of the two-round network on a state whose last
    stage was density, its MFMA rate, and its time at the bf16 peak beside
    the plain method's 4 substeps that one learned macro step replaces."""
    from libclsph_amd import native
    net = native.TwoRoundNetwork(ctx, state.n, e_cap=1024)
    net.set_params(gnn_params(net.count))
    ms, edges = net.profile(state, reps=5)
    net.close()
    total = sum(ms)
    flop = edges * GNN_EDGE_FLOP + state.n * GNN_NODE_FLOP
    tflops = flop / (total * 1e-3) / 1e12
    return {
        "network": GNN_SHAPE,
        "network_parameters": int(native.lib.clsph_gnn_param_count()),
        "ms": total,
        "ms_parts": {"edge_round_1": ms[0], "node_round_1": ms[1], "edge_round_2": ms[2],
                     "node_round_2": ms[3]},
        "edges": edges,
        "tflop_per_evaluation": flop / 1e12,
        "ms_at_mfma_peak": flop / (MFMA_BF16_PEAK_TFLOPS * 1e12) * 1e3,
        "plain_ms_per_4_substeps": 4.0 * plain_ms_per_step,
        "mfma": {"unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS, "achieved": tflops,
                 "frac": tflops / MFMA_BF16_PEAK_TFLOPS},
    }


def bench_learned(ctx, plain, scene, plain_ms_per_step, macro_steps, judge_every=16, tol=0.01):
    """The learned surrogate (trained weights shipped in-tree) against the
    plain method over the same simulated interval from the same state: the
    plain substeps timed over that interval, then the learned macro steps
    without and with the judge. Reports simulated seconds per wall second of
    each, the judge's fallbacks, and the density error against the plain run
    at dt at the matched time (max and mean |rho - rho_plain| / rho0), for
    the learned run and for the plain method at the same 4 dt step (the
    method without the network), beside the literal max |rho / rho0 - 1|."""
    from libclsph_amd import learned as lrn
    from libclsph_amd import native
    if not os.path.exists(lrn.WEIGHTS):
        return {"skipped": f"no trained weights at {lrn.WEIGHTS}"}
    n = plain.n
    weights = np.load(lrn.WEIGHTS)
    if weights.shape != (native.lib.clsph_nn_param_count(),):
        return {"skipped": f"{lrn.WEIGHTS} holds {weights.size} parameters, the network "
                           f"{native.lib.clsph_nn_param_count()}"}
    dt = float(scene.params["time_delta"])
    start = lrn.make_state(ctx, scene)
    learner = lrn.make_state(ctx, scene, dt=lrn.MACRO * dt)
    teacher = lrn.make_state(ctx, scene)
    big = lrn.make_state(ctx, scene, dt=lrn.MACRO * dt)
    net = native.Network(ctx, n, e_cap=1024)
    net.set_params(weights)
    plain.track_ids(True)
    for s in (start, learner, teacher, big):
        s.track_ids(True)
    start.copy_from(plain)
    # the network's own time on the start state
    learner.copy_from(start)
    lrn.forces_stages(learner)
    net_ms, edges = net.profile(learner, reps=5)
    learner.copy_from(start)
    lrn.density_stages(learner)
    gnn = bench_gnn(ctx, learner, plain_ms_per_step)
    # plain substeps over the interval
    ctx.sync()
    t0 = time.perf_counter()
    plain.step(macro_steps * lrn.MACRO)
    ctx.sync()
    plain_wall = time.perf_counter() - t0
    # learned macro steps over the same interval, without the judge
    learner.copy_from(start)
    bare = lrn.run_learned(ctx, learner, teacher, net, macro_steps, dt, judge_every=0)
    # and with it
    learner.copy_from(start)
    r = lrn.run_learned(ctx, learner, teacher, net, macro_steps, dt, judge_every, tol)
    # the plain method at the same 4 dt step (the method without the network)
    big.copy_from(start)
    ctx.sync()
    t0 = time.perf_counter()
    big.step(macro_steps)
    ctx.sync()
    big_wall = time.perf_counter() - t0
    for s in (plain, learner, big):
        lrn.density_stages(s)
    err_max, err_mean = net.compare_density(learner, plain)
    big_max, big_mean = net.compare_density(big, plain)
    rho0 = float(scene.params["fluid_density"])
    lit = {}
    for name, s in (("learned", learner), ("plain", plain)):
        d = s.download_sorted()["density"].astype(np.float64)
        lit[name] = float(np.abs(d / rho0 - 1.0).max())
    for s in (start, learner, teacher, big):
        s.close()
    plain.track_ids(False)
    net.close()
    sim_s = macro_steps * lrn.MACRO * dt
    plain_rate = sim_s / plain_wall
    tflops = (edges * NN_EDGE_FLOP + n * NN_NODE_FLOP) / (net_ms * 1e-3) / 1e12
    return {
        "interval": {"substeps": macro_steps * lrn.MACRO, "sim_s": sim_s,
                     "plain_ms_per_substep": plain_wall / (macro_steps * lrn.MACRO) * 1e3},
        "network": NN_SHAPE,
        "network_parameters": int(native.lib.clsph_nn_param_count()),
        "sim_s_per_wall_s": {"plain": plain_rate, "learned": r.sim_s / r.wall_s,
                             "ratio": (r.sim_s / r.wall_s) / plain_rate,
                             "learned_unjudged": bare.sim_s / bare.wall_s,
                             "ratio_unjudged": (bare.sim_s / bare.wall_s) / plain_rate,
                             "plain_4dt": sim_s / big_wall,
                             "ratio_plain_4dt": (sim_s / big_wall) / plain_rate,
                             "learned_over_plain_4dt": (r.sim_s / r.wall_s) / (sim_s / big_wall)},
        "macro_step_dt": lrn.MACRO * dt,
        "macro_steps": r.macro_steps,
        "unjudged_fraction": 1.0 - r.judged / max(1, r.macro_steps),
        "ms_per_macro_step": {"unjudged": bare.wall_s / macro_steps * 1e3,
                              "judged": r.wall_s / macro_steps * 1e3},
        "judge_every": judge_every,
        "judged": r.judged,
        "fallbacks": r.fallbacks,
        "judge_max_density_err": max(r.judge_err) if r.judge_err else None,
        "judge_mean_density_err": max(r.judge_mean) if r.judge_mean else None,
        "density_err_vs_plain": {"max": err_max, "mean": err_mean},
        "plain_4dt_density_err_vs_plain": {"max": big_max, "mean": big_mean},
        "literal_max_abs_rho_over_rho0_minus_1": lit,
        "network_ms": net_ms,
        "edges": edges,
        "network_kernel_mfma": {
            "unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS, "achieved": tflops,
            "frac": tflops / MFMA_BF16_PEAK_TFLOPS},
        "two_round_network": gnn,
    }


def bench_batched(args):
    from libclsph_amd import autotune, native, scenes
    batch_scenes = scenes.dam_break_batch(64)
    counts = [s.n for s in batch_scenes]
    n = sum(counts)
    steps = args.steps or 200
    warmup = 20 if args.warmup is None else args.warmup
    ctx = native.Context(0)
    b = autotune.make_batch(ctx, batch_scenes)
    sizes = autotune.sizes_for(counts)
    b.set_block_sizes(sizes)
    print(f"# device {ctx.device_name()}; 64 scenes, {n} particles; work-group sizes {sizes}",
          file=sys.stderr, flush=True)
    if args.direct:
        # every substep launched stage by stage, HIP events between stages
        if warmup:
            b.profile_steps(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        prof = b.profile_steps(steps)
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3
    else:
        if warmup:
            b.step(warmup)
        ctx.sync()
        ms = b.time_steps(steps)
        prof = b.profile_steps(args.profile_steps if args.profile_steps > 0 else 50)
    ms_step = ms / steps
    finite = all(bool(np.isfinite(b.download_scene(i)["position"][:, :3]).all())
                 for i in range(0, 64, 7))
    b.close()
    # The same scenes one after another, each in its own single-scene state
    # (graph replay; not run with --direct, whose launches are not comparable).
    seq_ms = None
    if not args.direct:
        seq = []
        for sc in batch_scenes:
            st = native.State(ctx, sc.n)
            st.set_params(sc.params, sc.terms)
            st.scene(sc.normals, sc.vertices, sc.indices)
            st.upload(sc.particles)
            st.step(warmup)
            seq.append(st.time_steps(20) / 20)
            st.close()
        seq_ms = float(sum(seq))
    nbytes = stage_bytes(n, 0)
    dom = max((k for k in prof if k != "grid"), key=lambda k: prof[k])
    achieved = nbytes[dom] / (prof[dom] * 1e-3) / 1e9
    out = {
        "metric": "aggregate particle-steps/s, 64 batched dam-break scenes",
        "value": n * steps / (ms / 1e3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched_dam_break", "scenes": 64, "particles": n,
                   "min_scene": min(counts), "max_scene": max(counts),
                   "block_sizes": sizes, "state_finite": finite,
                   "launch": "stage by stage" if args.direct else "graph replay"},
        "per_scene_latency_ms": ms_step,
        "sequential_ms_per_step": seq_ms,
        "batching_speedup": seq_ms / ms_step if seq_ms else None,
        "stages_ms": {k: round(v, 5) for k, v in prof.items()},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
