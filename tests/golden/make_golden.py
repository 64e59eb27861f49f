"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

  dam4096_step0.npz  : the 4,096-particle dam break (seed 1234): initial
                       positions, then every stage output of one substep
                       (grid fields, cell ids, stable order, cell table,
                       density, pressure, acceleration, advected state)
  dam16k_series.npz  : the 16,384-particle dam break run 1,000 substeps:
                       mean density and kinetic energy after each substep
  pour8k_series.npz  : 8,192 particles poured onto the 32,768-triangle
                       heightfield (dt 2 ms), 50 substeps: mean density and
                       kinetic energy after each substep
  dam1m_series.npz   : the 1,048,576-particle dam break run 1,000 substeps:
                       mean density and kinetic energy after each substep
                       (about 3 hours on 8 host cores; written with --only
                       series1m, which checkpoints every 50 substeps into
                       --checkpoint-dir and resumes from there)

    python tests/golden/make_golden.py [--skip-series]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from libclsph_amd import scenes  # noqa: E402


def step0():
    sc = scenes.dam_break(4096, seed=1234)
    r = oracle.stages(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices, sc.indices)
    p = r.params
    np.savez_compressed(
        os.path.join(HERE, "dam4096_step0.npz"),
        position0=sc.particles["position"][:, :3],
        min_point=p["min_point"][:3], max_point=p["max_point"][:3],
        grid_size=np.array([p["grid_size_x"], p["grid_size_y"], p["grid_size_z"]]),
        grid_cell_count=np.array(p["grid_cell_count"]),
        keys=r.hashed["grid_index"],
        order=np.argsort(r.hashed["grid_index"], kind="stable").astype(np.uint32),
        table=r.table,
        density=r.density["density"], pressure=r.density["pressure"],
        acceleration=r.forces["acceleration"][:, :3],
        position1=r.advected["position"][:, :3],
        velocity1=r.advected["velocity"][:, :3],
        intermediate_velocity1=r.advected["intermediate_velocity"][:, :3],
    )


def series(steps=1000):
    sc = scenes.dam_break(16384)
    t0 = time.time()
    _, rho, ke, _ = oracle.run(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices,
                               sc.indices, steps)
    print(f"series: {steps} substeps in {time.time() - t0:.0f} s")
    np.savez_compressed(os.path.join(HERE, "dam16k_series.npz"), mean_density=rho,
                        kinetic_energy=ke)


def pour_series(steps=50):
    sc = scenes.heightfield_pour(8192, dt=2e-3)
    t0 = time.time()
    _, rho, ke, _ = oracle.run(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices,
                               sc.indices, steps)
    print(f"pour series: {steps} substeps in {time.time() - t0:.0f} s")
    np.savez_compressed(os.path.join(HERE, "pour8k_series.npz"), mean_density=rho,
                        kinetic_energy=ke)


def series_1m(steps=1000, chunk=50, ckpt_dir="/tmp/clsph_golden_1m"):
    sc = scenes.dam_break(1048576)
    os.makedirs(ckpt_dir, exist_ok=True)
    ck = os.path.join(ckpt_dir, "dam1m_ckpt.npz")
    state, p = sc.particles.copy(), sc.params.copy()
    rho, ke = np.zeros(0), np.zeros(0)
    if os.path.exists(ck):
        z = np.load(ck)
        state = z["state"].view(sc.particles.dtype).reshape(-1).copy()
        p = z["params"].view(sc.params.dtype).reshape(()).copy()
        rho, ke = z["mean_density"], z["kinetic_energy"]
        print(f"series1m: resuming after {len(rho)} substeps", flush=True)
    while len(rho) < steps:
        k = min(chunk, steps - len(rho))
        t0 = time.time()
        state, r, e, p = oracle.run(state, p, sc.terms, sc.normals, sc.vertices, sc.indices, k)
        rho, ke = np.concatenate([rho, r]), np.concatenate([ke, e])
        np.savez(ck + ".tmp.npz", state=state.view(np.uint8), params=np.atleast_1d(p).view(np.uint8),
                 mean_density=rho, kinetic_energy=ke)
        os.replace(ck + ".tmp.npz", ck)
        print(f"series1m: {len(rho)} substeps ({time.time() - t0:.0f} s for {k})", flush=True)
    np.savez_compressed(os.path.join(HERE, "dam1m_series.npz"), mean_density=rho,
                        kinetic_energy=ke)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-series", action="store_true")
    ap.add_argument("--only", choices=["step0", "series", "pour", "series1m"])
    ap.add_argument("--checkpoint-dir", default="/tmp/clsph_golden_1m")
    a = ap.parse_args()
    if a.only in (None, "step0"):
        step0()
    if a.only in (None, "series") and not a.skip_series:
        series()
    if a.only in (None, "pour") and not a.skip_series:
        pour_series()
    if a.only == "series1m":
        series_1m(ckpt_dir=a.checkpoint_dir)
