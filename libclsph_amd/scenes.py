"""Synthetic scenes: the dam-break configuration in its several sizes.

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


def dam_break(n, seed=1234, jitter=0.01, block=(0.4, 0.8, 0.4), box=1.0, **kw):
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
