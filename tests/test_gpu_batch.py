"""Batched scenes: every scene of a batch evolves exactly as it does alone,
whatever the work-group sizes."""
import numpy as np
import pytest

import oracle
from libclsph_amd import autotune, native, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def small_batch():
    return scenes.dam_break_batch(n_scenes=6, lo=700, hi=20000, seed=5)


def run_alone(ctx, sc, steps):
    st = native.State(ctx, sc.n, table_cap=1 << 20)
    st.set_params(sc.params, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    st.upload(sc.particles)
    st.step(steps)
    out = st.download()
    st.close()
    return out


FIELDS = ("position", "velocity", "intermediate_velocity", "density", "pressure", "grid_index")


def test_batch_equals_scenes_alone(ctx):
    print("case batch of 6 vs alone, 15 substeps", flush=True)
    scs = small_batch()
    b = autotune.make_batch(ctx, scs)
    b.step(15)
    for i, sc in enumerate(scs):
        got = b.download_scene(i)
        want = run_alone(ctx, sc, 15)
        for f in FIELDS:
            np.testing.assert_array_equal(got[f], want[f], err_msg=f"scene {i} {f}")
    b.close()


def ragged_scene(n, seed):
    """A dam break cut to its first n particles (n may be 0)."""
    sc = scenes.dam_break(max(n, 64), seed=seed)
    p = sc.params.copy()
    p["particles_count"] = n
    return scenes.Scene(sc.particles[:n].copy(), p, sc.terms, sc.normals, sc.vertices,
                        sc.indices)


def test_ragged_batch_with_an_empty_scene(ctx):
    """Scenes of 5,000, 0, 1, 63 and 2,049 particles in one batch (an empty
    scene, a lone particle, less than a wave, one past a sort tile): after 10
    substeps every non-empty scene is bit-identical to itself run alone, and
    the empty scene holds nothing and raises no error."""
    counts = [5000, 0, 1, 63, 2049]
    print(f"case ragged batch {counts}, 10 substeps", flush=True)
    scs = [ragged_scene(c, 31 + i) for i, c in enumerate(counts)]
    b = autotune.make_batch(ctx, scs)
    b.step(10)
    for i, sc in enumerate(scs):
        got = b.download_scene(i)
        assert len(got) == counts[i]
        if counts[i] == 0:
            continue
        want = run_alone(ctx, sc, 10)
        for f in FIELDS:
            np.testing.assert_array_equal(got[f], want[f], err_msg=f"scene {i} {f}")
    b.close()


def test_block_sizes_do_not_change_results(ctx):
    print("case batch work-group sizes", flush=True)
    scs = small_batch()
    ref = autotune.make_batch(ctx, scs)
    ref.step(8)
    a = autotune.make_batch(ctx, scs)
    a.set_block_sizes([0, 64, 0, 0, 512, 128, 64])
    a.step(4)
    a.set_block_sizes([0, 512, 0, 0, 64, 512, 256])
    a.step(4)
    assert a.block_sizes() == [256, 512, 256, 256, 64, 512, 256]
    for i in range(len(scs)):
        x, y = ref.download_scene(i), a.download_scene(i)
        for f in FIELDS:
            np.testing.assert_array_equal(x[f], y[f], err_msg=f"scene {i} {f}")
    ref.close()
    a.close()


def test_batch_grid_matches_oracle_per_scene(ctx):
    print("case batch grid fields vs oracle", flush=True)
    scs = small_batch()
    b = autotune.make_batch(ctx, scs)
    b.stage("grid")
    for i, sc in enumerate(scs):
        p = sc.params.copy()
        assert oracle.grid_params(sc.particles, p) == 0
        g = b.grid_scene(i)
        np.testing.assert_array_equal(g["min_point"][:3], p["min_point"][:3])
        assert int(g["grid_cell_count"]) == int(p["grid_cell_count"])
    b.close()


def test_full_batch_64_scenes_equal_scenes_alone(ctx):
    """The benchmark's batch: 64 dam breaks of 4,096-262,144 particles
    (log-uniform, seeded), viscosity 1e-3-5e-2, with the autotuner's
    work-group sizes. After 5 substeps every scene is bit-identical to the
    same scene run alone; before them every scene's grid fields equal the
    oracle's."""
    scs = scenes.dam_break_batch(64)
    counts = [s.n for s in scs]
    print(f"case 64-scene batch, {sum(counts)} particles, {min(counts)}-{max(counts)}",
          flush=True)
    assert len(scs) == 64 and min(counts) >= 4096 and max(counts) <= 262144
    b = autotune.make_batch(ctx, scs)
    b.set_block_sizes(autotune.sizes_for(counts))
    b.stage("grid")
    for i, sc in enumerate(scs):
        p = sc.params.copy()
        assert oracle.grid_params(sc.particles, p) == 0
        g = b.grid_scene(i)
        for f in ("min_point", "max_point"):
            np.testing.assert_array_equal(g[f][:3], p[f][:3], err_msg=f"scene {i} {f}")
        for f in ("grid_size_x", "grid_size_y", "grid_size_z", "grid_cell_count"):
            assert int(g[f]) == int(p[f]), (i, f)
    b.step(5)
    for i, sc in enumerate(scs):
        got = b.download_scene(i)
        want = run_alone(ctx, sc, 5)
        for f in FIELDS:
            np.testing.assert_array_equal(got[f], want[f], err_msg=f"scene {i} {f}")
    b.close()
