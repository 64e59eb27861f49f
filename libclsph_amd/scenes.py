"""Synthetic scenes: the dam-break configuration in its several sizes, and
the column poured onto a heightfield.

Dam break: a fluid block 0.4 x 0.8 x 0.4 m on a cubic lattice at spacing
s = cbrt(0.128 / N) (so N particles fill it; h = 2 s, i.e. spacing 0.5 h),
coordinates jittered by U(+-1 %) of s from a seeded counter-based generator,
resting in the corner of the axis-aligned box [0, 1]^3 (12 wall triangles).
Rest density 1000 kg/m^3, particle mass 1000 s^3, gas stiffness 3.0,
dynamic viscosity 3.5e-3, surface tension 0.0728 with the water threshold
7.065 (fluid_properties/water.json:6), restitution 0, gravity -9.81, zero
initial velocity, dt = 1e-4 s (target_fps 1e4, simulation_scale 1).
"""
from dataclasses import dataclass

import numpy as np

from . import layout, native


@dataclass
class Scene:
    particles: np.ndarray
    params: np.ndarray
    terms: np.ndarray
    normals: np.ndarray
    vertices: np.ndarray
    indices: np.ndarray

    @property
    def n(self):
        return len(self.particles)


def fluid_params(n, h, mass, dt=1e-4, viscosity=3.5e-3, stiffness=3.0, rest_density=1000.0,
                 surface_tension=0.0728, threshold=7.065, restitution=0.0, gravity=-9.81,
                 steps=100):
    """simulation_parameters for n particles with the derived fields that
    load_settings computes (sph_simulation.cpp:490-497)."""
    p = layout.params()
    f32 = np.float32
    p["particles_count"] = n
    p["fluid_density"] = f32(rest_density)
    p["particle_mass"] = f32(mass)
    p["total_mass"] = f32(n) * f32(mass)
    p["dynamic_viscosity"] = f32(viscosity)
    p["K"] = f32(stiffness)
    p["surface_tension"] = f32(surface_tension)
    p["surface_tension_threshold"] = f32(threshold)
    p["restitution"] = f32(restitution)
    p["constant_acceleration"] = np.array([0.0, gravity, 0.0, 0.0], dtype=np.float32)
    p["h"] = f32(h)
    p["target_fps"] = f32(1.0 / dt)
    p["simulation_scale"] = f32(1.0)
    p["time_delta"] = f32(1.0) / p["target_fps"]
    p["max_velocity"] = f32(0.8) * p["h"] / p["time_delta"]
    p["simulation_time"] = f32(steps) * p["time_delta"]
    return p


def dam_break(n, seed=1234, jitter=0.01, block=(0.4, 0.8, 0.4), box=1.0, spacing=None, **kw):
    """The dam break with n particles. By default the lattice spacing makes n
    particles fill the block; with `spacing` given, the block and box are
    scaled by cbrt(n s^3 / 0.128) instead, keeping that spacing."""
    if spacing is not None:
        spacing = np.float32(spacing)
        k = (n * float(spacing) ** 3 / 0.128) ** (1.0 / 3.0)
        block = (0.4 * k, 0.8 * k, 0.4 * k)
        box = box * k
    else:
        spacing = np.float32((0.128 / n) ** (1.0 / 3.0))
        if block != (0.4, 0.8, 0.4):
            vol = block[0] * block[1] * block[2]
            spacing = np.float32((vol / n) ** (1.0 / 3.0))
    h = np.float32(2.0) * spacing
    mass = np.float32(1000.0 * float(spacing) ** 3)
    parts = native.make_block(n, float(spacing), (0.0, 0.0, 0.0), block, jitter, seed)
    if len(parts) != n:
        raise ValueError(f"block holds {len(parts)} particles, asked for {n}")
    params = fluid_params(n, h, mass, **kw)
    terms = native.kernel_terms(float(h))
    normals, verts, idx = native.make_box((0.0, 0.0, 0.0), (box, box, box))
    return Scene(parts, params, terms, normals, verts, idx)


def merge_meshes(*meshes):
    """Concatenates (normals, vertices, indices) triples into one mesh."""
    normals, verts, idx = [], [], []
    base = 0
    for nrm, v, i in meshes:
        normals.append(np.asarray(nrm, np.float32))
        verts.append(np.asarray(v, np.float32))
        idx.append(np.asarray(i, np.uint32) + base)
        base += len(v) // 3
    return np.concatenate(normals), np.concatenate(verts), np.concatenate(idx)


def terrain(res=128, amplitude=0.1, octaves=4, seed=77, walls=True):
    """A res x res heightfield over [0, 1]^2 (2 res^2 triangles, heights from
    4-octave value noise with amplitude 0.1 m) inside the side walls and
    floor of the box [0, 1] x [-0.1, 2] x [0, 1] (the top is open)."""
    hf = native.make_heightfield(res, 0.0, 0.0, 1.0, amplitude, octaves, seed)
    if not walls:
        return hf
    nrm, v, idx = native.make_box((0.0, -0.1, 0.0), (1.0, 2.0, 1.0))
    tri = idx.reshape(-1, 3)
    V = v.reshape(-1, 3)
    top = np.all(V[tri][:, :, 1] > 1.99, axis=1)  # drop the two top triangles
    keep = ~top
    box = (nrm.reshape(-1, 3)[keep].ravel(), v, tri[keep].ravel())
    return merge_meshes(hf, box)


def heightfield_pour(n, seed=4321, column=(0.3, 0.3), base=0.15, dt=5e-5, **kw):
    """n particles in a column (0.3 x 0.3 m footprint centred over the
    terrain, bottom at 0.15 m) with seeded random velocities
    (U(+-0.25) m/s sideways, 0.5-1.5 m/s downward), at the dam-break lattice
    spacing s = cbrt(0.128 / n), h = 2 s, above the terrain of terrain()."""
    spacing = np.float32((0.128 / n) ** (1.0 / 3.0))
    h = np.float32(2.0) * spacing
    mass = np.float32(1000.0 * float(spacing) ** 3)
    x0 = 0.5 - column[0] / 2
    z0 = 0.5 - column[1] / 2
    parts = native.make_column(n, float(spacing), (x0, base, z0), (column[0], 0.0, column[1]),
                               0.01, 0.25, 0.5, 1.5, seed)
    params = fluid_params(n, h, mass, dt=dt, **kw)
    terms = native.kernel_terms(float(h))
    normals, verts, idx = terrain()
    return Scene(parts, params, terms, normals, verts, idx)


def dam_break_batch(n_scenes=64, lo=4096, hi=262144, seed=99, spacing=None):
    """Independent dam breaks with particle counts log-uniform in [lo, hi] and
    dynamic viscosity uniform in [1e-3, 5e-2] (seeded). Every scene keeps the
    lattice spacing of the 1,048,576-particle dam break (spacing 0.5 h), its
    block and box scaled to its count."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n_scenes)).astype(np.int64)
    visc = rng.uniform(1e-3, 5e-2, n_scenes)
    if spacing is None:
        spacing = (0.128 / 1048576) ** (1.0 / 3.0)
    return [dam_break(int(c), seed=seed * 1000 + i, spacing=spacing, viscosity=float(v))
            for i, (c, v) in enumerate(zip(counts, visc))]
