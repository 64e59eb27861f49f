"""Work-group size autotuner for batched (and single-scene) substeps.

The per-particle kernels (hash, density, forces, advection) can run with 64,
128, 256 or 512 threads per work-group; the best choice depends on the scene
sizes (how many partial work-groups the scene boundaries leave, occupancy of
the neighbour loops). For each scene-size bucket (log2 of the scene particle
count) the tuner builds a batch of equal dam breaks of that size, measures
graph-replayed substeps for every candidate size of one stage at a time
(coordinate descent, two sweeps) and records the fastest setting. A batch is
then run with the entry of its particle-weighted median scene size.

    python -m libclsph_amd.autotune --out table.json [--buckets 12 14 16 18]
"""
import argparse
import json
import math
import os
import sys

import numpy as np

CANDIDATES = (64, 128, 256, 512)
# Stages that run over the work-group table (stage order of clsph_time_stage).
TUNED = {"hash": 1, "density": 4, "forces": 5, "advect": 6}
DEFAULT = [256, 256, 256, 256, 512, 512, 256]  # the state's own defaults
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mi355x.json")


def bucket_of(counts):
    """log2 bucket of the particle-weighted median scene size."""
    c = np.sort(np.asarray(counts, dtype=np.float64))
    w = np.cumsum(c) / c.sum()
    med = c[np.searchsorted(w, 0.5)]
    return int(round(math.log2(max(med, 1.0))))


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): v for k, v in json.load(f)["buckets"].items()}


def sizes_for(counts, table=None):
    """Work-group sizes for a batch with these scene counts (the defaults
    without a table)."""
    table = load_table() if table is None else table
    if not table:
        return list(DEFAULT)
    b = bucket_of(counts)
    key = min(table, key=lambda k: abs(k - b))
    return list(table[key]["sizes"])


def time_config(state, sizes, steps):
    state.set_block_sizes(sizes)
    state.step(3)  # capture + settle
    return state.time_steps(steps) / steps


def ab_test(state, best, trial, steps):
    """Times trial, best, trial back to back on the (evolving) state and
    returns (best ms/substep of the trial, ms/substep of the incumbent)."""
    t1 = time_config(state, trial, steps)
    t0 = time_config(state, best, steps)
    t2 = time_config(state, trial, steps)
    return min(t1, t2), t0


def tune_state(state, steps=20, sweeps=2, log=print, margin=0.01):
    """Coordinate descent over the tuned stages' work-group sizes. A trial
    replaces the incumbent only when it beats it by `margin` in a trial /
    incumbent / trial sequence, so that drift of the evolving scene and
    timing noise do not decide."""
    sizes = list(state.block_sizes())
    for _ in range(sweeps):
        for name, stage in TUNED.items():
            for c in CANDIDATES:
                if c == sizes[stage]:
                    continue
                trial = list(sizes)
                trial[stage] = c
                t, t_best = ab_test(state, sizes, trial, steps)
                if t < t_best * (1.0 - margin):
                    sizes = trial
                    log(f"  {name} -> {c}: {t:.4f} vs {t_best:.4f} ms/substep")
    state.set_block_sizes(sizes)
    return sizes


def make_batch(ctx, scenes):
    from . import native
    b = native.Batch(ctx, [s.n for s in scenes])
    for i, s in enumerate(scenes):
        b.set_scene_params(i, s.params, s.terms)
        b.mesh_scene(i, s.normals, s.vertices, s.indices)
        b.upload_scene(i, s.particles)
    return b


def tune_buckets(buckets, scenes_per_bucket=16, steps=20, log=print):
    from . import native, scenes
    ctx = native.Context(0)
    spacing = (0.128 / 1048576) ** (1.0 / 3.0)
    out = {}
    for bk in buckets:
        n = 2 ** bk
        k = max(1, min(scenes_per_bucket, (1 << 22) // n))
        batch = make_batch(ctx, [scenes.dam_break(n, seed=7 + i, spacing=spacing)
                                 for i in range(k)])
        batch.step(20)
        sizes = tune_state(batch, steps=steps, log=log)
        best, base = ab_test(batch, list(DEFAULT), sizes, steps)
        log(f"bucket 2^{bk} ({k} scenes): default {base:.4f} ms, tuned {best:.4f} ms, {sizes}")
        out[bk] = {"sizes": sizes, "ms_default": base, "ms_tuned": best, "scenes": k}
        batch.close()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--buckets", type=int, nargs="*", default=[12, 14, 16, 18])
    ap.add_argument("--steps", type=int, default=20)
    a = ap.parse_args()
    res = tune_buckets(a.buckets, steps=a.steps, log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump({"device": "MI355X (gfx950)", "candidates": list(CANDIDATES),
                   "buckets": {str(k): v for k, v in res.items()}}, f, indent=1)
    print(f"wrote {a.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
