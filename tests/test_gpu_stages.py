"""Stage-by-stage parity of the HIP pipeline against the CPU oracle.

Each case uploads the same particles to the GPU and to the oracle, runs one
substep stage by stage and compares:
  grid parameters, cell ids, stable sorted order, cell table : bit-exact
  density (and pressure)                                      : 1e-5 relative
  acceleration                                                : 1e-4 * max|a|
  advected state                                              : see test
"""
import numpy as np
import pytest

import oracle
from libclsph_amd import native, scenes

pytestmark = pytest.mark.gpu

DENSITY_RTOL = 1e-5
FORCE_TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    print("device:", c.device_name(), flush=True)
    yield c
    c.close()


def cloud(n, seed, dup=0, lo=0.2, hi=0.8):
    """Random particles in [lo, hi]^3 with `dup` exact duplicates appended
    (coincident particles exercise the spiky special case and equal keys)."""
    rng = np.random.default_rng(seed)
    sc = scenes.dam_break(max(n, 1))
    parts = sc.particles.copy()[:n]
    parts["position"][:, :3] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    parts["velocity"][:, :3] = rng.normal(0, 0.05, size=(n, 3)).astype(np.float32)
    parts["intermediate_velocity"][:, :3] = rng.normal(0, 0.05, size=(n, 3)).astype(np.float32)
    if dup:
        parts[n - dup:] = parts[:dup]
    p = sc.params.copy()
    p["particles_count"] = n
    h = np.float32(0.04)
    p["h"] = h
    p["max_velocity"] = np.float32(0.8) * h / p["time_delta"]
    return scenes.Scene(parts, p, native.kernel_terms(float(h)), sc.normals, sc.vertices,
                        sc.indices)


def stage_outputs(st):
    """Runs the six stages and advection of one substep on state `st` from
    its current particles; returns every stage's readback."""
    out = {}
    st.stage("grid")
    out["grid"] = st.grid()
    st.stage("hash")
    out["keys"] = st.keys()
    st.stage("sort")
    out["sorted_keys"], out["perm"] = st.sorted()
    st.stage("cell_table")
    G = int(out["grid"]["grid_cell_count"])
    out["table"] = st.cell_table(G)
    st.stage("density")
    out["density"] = st.download_sorted()
    st.stage("forces")
    out["forces"] = st.download_sorted()
    st.stage("advect")
    out["advected"] = st.download()
    out["capped"] = st.bounce_capped()
    return out


def run_gpu_stages(ctx, sc, table_cap=1 << 22):
    st = native.State(ctx, sc.n, table_cap=table_cap)
    st.set_params(sc.params, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    st.upload(sc.particles)
    out = stage_outputs(st)
    st.close()
    return out


def boundary_cloud(seed, h=0.04):
    """Particles exactly on cell faces of the padded grid and one float32 ulp
    either side of them (the cell-id truncation's edge cases), plus random
    ones; an anchor and a far-corner particle fix the bounding box."""
    rng = np.random.default_rng(seed)
    f = np.float32
    L = f(h) * f(2)
    anchor = rng.uniform(0.1, 0.3, 3).astype(f)
    mn = anchor - L * f(2)
    pts = [anchor, anchor + f(0.5)]
    for _ in range(1500):
        b = (mn + rng.integers(3, 8, 3).astype(f) * L).astype(f)
        side = rng.integers(-1, 2, 3)
        pts.append(np.where(side < 0, np.nextafter(b, f(-1)),
                            np.where(side > 0, np.nextafter(b, f(2)), b)).astype(f))
    pts += list(rng.uniform(anchor, anchor + f(0.5), (1000, 3)).astype(f))
    sc = cloud(len(pts), seed)
    sc.particles["position"][:, :3] = np.array(pts, dtype=f)
    return sc


CASES = {
    "dam16k": lambda: scenes.dam_break(16384),
    "dam4096": lambda: scenes.dam_break(4096, seed=7),
    "cloud1": lambda: cloud(1, 1),
    "cloud2": lambda: cloud(2, 2),
    "cloud63": lambda: cloud(63, 3),
    "cloud1000dup": lambda: cloud(1000, 4, dup=50),
    "cloud4097": lambda: cloud(4097, 5),
    # 300 particles inside one h-sphere: more neighbours than a list holds
    "dense300": lambda: cloud(300, 6, lo=0.49, hi=0.51),
    "boundary2502": lambda: boundary_cloud(8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_stage_parity(ctx, case):
    print(f"case {case}", flush=True)
    sc = CASES[case]()
    ref = oracle.stages(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices, sc.indices)
    assert ref.grid_rc == 0
    g = run_gpu_stages(ctx, sc)
    check_against_oracle(sc.particles, sc.params, g, ref)


def sparse_cloud(n, seed, span=11.0):
    """n particles over span x span x 0.5 (h = 0.04): a grid of about
    140 x 140 x 8 cells whose Morton cell ids need 23 bits."""
    sc = cloud(n, seed)
    rng = np.random.default_rng(seed + 1)
    pos = sc.particles["position"]
    pos[:, 0] = rng.uniform(0.0, span, n).astype(np.float32)
    pos[:, 1] = rng.uniform(0.0, span, n).astype(np.float32)
    pos[:, 2] = rng.uniform(0.0, 0.5, n).astype(np.float32)
    return sc


# A state whose table holds 2^24 cells is captured with three LSD passes
# (two cover cell ids of at most 22 bits). Its third pass runs only when the
# substep's cell ids need it: skipped for the dam break, run for the sparse
# cloud. Both end in the stable order the oracle gives.
THREE_PASS_CASES = {
    "dam4096_two_of_three": (lambda: scenes.dam_break(4096, seed=7), 22),
    "sparse3000_three_of_three": (lambda: sparse_cloud(3000, 11), 23),
    # three passes over 293 tiles of 2,048 keys
    "sparse600k_three_of_three": (lambda: sparse_cloud(600000, 13), 23),
}


@pytest.mark.parametrize("case", list(THREE_PASS_CASES))
def test_three_pass_state_parity(ctx, case):
    print(f"case {case}", flush=True)
    make, id_bits = THREE_PASS_CASES[case]
    sc = make()
    ref = oracle.stages(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices, sc.indices)
    assert ref.grid_rc == 0
    G = int(ref.params["grid_cell_count"])
    assert (G <= 1 << 22) == (id_bits <= 22) and G <= 1 << 24  # the case exercises its path
    g = run_gpu_stages(ctx, sc, table_cap=1 << 24)
    check_against_oracle(sc.particles, sc.params, g, ref)


# Sizes whose tile counts (2,048 keys a tile, 64 slices of tiles per digit in
# the scan) take each form of the sort's scan: at most 16 tiles a slice
# (733 tiles), at most 32 (1,465) and more than 32 (2,198: counts loaded
# twice). The sparse clouds' 23-bit cell ids need the third LSD pass.
LARGE_SORT_CASES = {
    "dam1500k_two_passes": (lambda: scenes.dam_break(1_500_000), 1 << 22),
    "sparse3m_three_passes": (lambda: sparse_cloud(3_000_000, 17), 1 << 24),
    "sparse4500k_three_passes": (lambda: sparse_cloud(4_500_000, 19), 1 << 24),
    # the 4M heightfield pour with its default table (4n cells): three passes
    # captured, two run
    "heightfield4m_two_of_three": (lambda: scenes.heightfield_pour(4_194_304), 1 << 24),
}


@pytest.mark.parametrize("case", list(LARGE_SORT_CASES))
def test_grid_sort_table_parity_at_scale(ctx, case):
    """Grid parameters, cell ids, the stable order and the cell table, all
    bit-exact against the oracle's grid, hash, sort and table stages."""
    print(f"case {case}", flush=True)
    make, cap = LARGE_SORT_CASES[case]
    sc = make()
    p = sc.params.copy()
    assert oracle.grid_params(sc.particles, p) == 0
    hashed = oracle.locate_in_grid(sc.particles, p)
    srt = oracle.sort_particles(hashed)
    G = int(p["grid_cell_count"])
    assert G <= cap
    st = native.State(ctx, sc.n, table_cap=cap)
    try:
        st.set_params(sc.params, sc.terms)
        st.scene(sc.normals, sc.vertices, sc.indices)
        st.upload(sc.particles)
        st.stage("grid")
        g = st.grid()
        st.stage("hash")
        keys = st.keys()
        st.stage("sort")  # also builds the cell table
        skeys, perm = st.sorted()
        table = st.cell_table(G)
    finally:
        st.close()
    for f in ("min_point", "max_point"):
        np.testing.assert_array_equal(g[f][:3], p[f][:3], err_msg=f)
    for f in ("grid_size_x", "grid_size_y", "grid_size_z", "grid_cell_count"):
        assert int(g[f]) == int(p[f]), f
    np.testing.assert_array_equal(keys, hashed["grid_index"])
    np.testing.assert_array_equal(skeys, srt["grid_index"])
    np.testing.assert_array_equal(perm, np.argsort(hashed["grid_index"], kind="stable"))
    np.testing.assert_array_equal(table, oracle.cell_table(srt, G))


def check_against_oracle(parts, params, g, ref):
    """Compares the readbacks of stage_outputs with the oracle's stages of
    the same substep from the same particles `parts`."""
    # grid parameters: bit-exact
    for f in ("min_point", "max_point"):
        np.testing.assert_array_equal(g["grid"][f][:3], ref.params[f][:3], err_msg=f)
    for f in ("grid_size_x", "grid_size_y", "grid_size_z", "grid_cell_count"):
        assert int(g["grid"][f]) == int(ref.params[f]), f

    # cell ids in canonical order: bit-exact
    np.testing.assert_array_equal(g["keys"], ref.hashed["grid_index"])
    # stable sorted order: bit-exact keys and the same permutation
    np.testing.assert_array_equal(g["sorted_keys"], ref.sorted["grid_index"])
    np.testing.assert_array_equal(g["perm"], np.argsort(ref.hashed["grid_index"], kind="stable"))
    np.testing.assert_array_equal(parts["position"][g["perm"]], ref.sorted["position"])
    # cell table: bit-exact
    np.testing.assert_array_equal(g["table"], ref.table)

    # density / pressure: 1e-5 relative
    rd, gd = ref.density["density"], g["density"]["density"]
    assert np.max(np.abs(gd - rd) / np.abs(rd)) <= DENSITY_RTOL
    rp, gp = ref.density["pressure"], g["density"]["pressure"]
    scale = max(np.max(np.abs(rp)), 1e-6)
    assert np.max(np.abs(gp - rp)) <= 1e-4 * scale

    # acceleration: 1e-4 of the largest magnitude
    ra = ref.forces["acceleration"][:, :3].astype(np.float64)
    ga = g["forces"]["acceleration"][:, :3].astype(np.float64)
    amax = np.max(np.linalg.norm(ra, axis=1))
    err = np.max(np.linalg.norm(ga - ra, axis=1))
    assert err <= FORCE_TOL * amax, (err, amax)

    # advection: positions and velocities after the substep
    rpos = ref.advected["position"][:, :3].astype(np.float64)
    gpos = g["advected"]["position"][:, :3].astype(np.float64)
    dt = float(params["time_delta"])
    # a 1e-4*max|a| force difference moves a particle by at most that * dt^2
    tol = FORCE_TOL * amax * dt * dt + 4e-7
    assert np.max(np.abs(gpos - rpos)) <= tol
    rv = ref.advected["velocity"][:, :3].astype(np.float64)
    gv = g["advected"]["velocity"][:, :3].astype(np.float64)
    assert np.max(np.abs(gv - rv)) <= FORCE_TOL * amax * dt + 1e-6
    assert g["capped"] == ref.capped
    # advection leaves acceleration zero, density/pressure/grid_index carried
    assert np.all(g["advected"]["acceleration"] == 0)
    np.testing.assert_array_equal(g["advected"]["grid_index"], ref.advected["grid_index"])


def test_graph_step_matches_stages(ctx):
    """clsph_step (graph-captured) gives the same state as running the stages."""
    sc = scenes.dam_break(16384)
    a = native.State(ctx, sc.n, table_cap=1 << 22)
    b = native.State(ctx, sc.n, table_cap=1 << 22)
    for st in (a, b):
        st.set_params(sc.params, sc.terms)
        st.scene(sc.normals, sc.vertices, sc.indices)
        st.upload(sc.particles)
    for _ in range(3):
        for s in ("grid", "hash", "sort", "cell_table", "density", "forces", "advect"):
            a.stage(s)
    b.step(3)
    pa, pb = a.download(), b.download()
    for f in ("position", "velocity", "intermediate_velocity", "density", "pressure",
              "grid_index"):
        np.testing.assert_array_equal(pa[f], pb[f], err_msg=f)


def test_upload_download_roundtrip(ctx):
    sc = scenes.dam_break(4096)
    parts = sc.particles.copy()
    rng = np.random.default_rng(0)
    for f in ("velocity", "intermediate_velocity", "acceleration"):
        parts[f][:, :3] = rng.normal(size=(len(parts), 3)).astype(np.float32)
    parts["density"] = rng.uniform(900, 1100, len(parts)).astype(np.float32)
    parts["pressure"] = rng.normal(size=len(parts)).astype(np.float32)
    parts["grid_index"] = rng.integers(0, 1 << 20, len(parts)).astype(np.uint32)
    st = native.State(ctx, len(parts), table_cap=1 << 20)
    st.set_params(sc.params, sc.terms)
    st.upload(parts)
    back = st.download()
    for f in parts.dtype.names:
        if parts.dtype[f].shape:
            np.testing.assert_array_equal(back[f][:, :3], parts[f][:, :3], err_msg=f)
        else:
            np.testing.assert_array_equal(back[f], parts[f], err_msg=f)


def density_forces_with_block_size(ctx, sc, bs):
    st = native.State(ctx, sc.n, table_cap=1 << 22)
    st.set_params(sc.params, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    st.set_block_sizes([bs] * 7)
    st.upload(sc.particles)
    for name in ("grid", "hash", "sort", "cell_table", "density"):
        st.stage(name)
    dens = st.download_sorted()
    pairs = st.neighbour_pairs()
    st.stage("forces")
    forc = st.download_sorted()
    st.close()
    return dens, forc, pairs


# dam4096 at 64 threads per work-group stages fewer halos than fit at 512, so
# its particles take both walks; dense4000 puts 4000 particles in one cell,
# beyond any work-group's staging capacity, so every particle walks HBM.
@pytest.mark.parametrize("case", ["dam4096", "cloud4097", "dense4000"])
def test_density_walks_agree_bit_for_bit(ctx, case):
    """The density kernel stages the halo of each cell its work-group's
    particles fall in (as many as fit) in LDS and walks HBM for the rest, in
    the same visiting order: densities, pressures, neighbour lists and hence
    accelerations do not depend on the work-group size, which changes how
    many halos are staged."""
    print(f"case {case}", flush=True)
    sc = (cloud(4000, 8, lo=0.49, hi=0.51) if case == "dense4000" else CASES[case]())
    runs = {bs: density_forces_with_block_size(ctx, sc, bs) for bs in (64, 128, 256, 512)}
    d0, f0, p0 = runs[256]
    for bs, (d, f, p) in runs.items():
        np.testing.assert_array_equal(d["density"], d0["density"], err_msg=f"density bs={bs}")
        np.testing.assert_array_equal(d["pressure"], d0["pressure"], err_msg=f"pressure bs={bs}")
        np.testing.assert_array_equal(f["acceleration"], f0["acceleration"],
                                      err_msg=f"acceleration bs={bs}")
        assert p == p0
    if case == "dense4000":
        ref = oracle.stages(sc.particles, sc.params, sc.terms, sc.normals, sc.vertices,
                            sc.indices)
        rd = ref.density["density"]
        # 4000-term FP32 sums, added in different orders by the two
        assert np.max(np.abs(d0["density"] - rd) / np.abs(rd)) <= 1e-4


def test_table_overflow_is_reported(ctx):
    """A grid with more cells than the state's table holds is reported as
    CLSPH_ETABLE at the next synchronisation. Meanwhile the table kernels
    skip their work and density finds no candidates (its table lookups are
    off), so the substep still runs to its end."""
    print("case table overflow", flush=True)
    sc = scenes.dam_break(4096, seed=7)
    st = native.State(ctx, sc.n, table_cap=64)
    st.set_params(sc.params, sc.terms)
    st.scene(sc.normals, sc.vertices, sc.indices)
    st.upload(sc.particles)
    st.step(1)
    with pytest.raises(native.ClsphError) as err:
        st.download()
    assert err.value.code == -5
    st.close()
