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
