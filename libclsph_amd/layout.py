"""numpy views of the three structures that cross the libclsph boundary.

Byte layouts are those of the reference's shared structs
(libclsph/common/structures.h:16-54) on a 64-bit host, where cl_float3 is a
16-byte aligned float[4]; include/clsph_types.h declares the same layout in C.
"""
import numpy as np

PARTICLE = np.dtype(
    {
        "names": ["position", "velocity", "intermediate_velocity", "acceleration",
                  "density", "pressure", "grid_index"],
        "formats": [("<f4", 4), ("<f4", 4), ("<f4", 4), ("<f4", 4), "<f4", "<f4", "<u4"],
        "offsets": [0, 16, 32, 48, 64, 68, 72],
        "itemsize": 80,
    }
)

_PARAM_FIELDS = [
    ("particles_count", "<u4", 0),
    ("max_velocity", "<f4", 4),
    ("fluid_density", "<f4", 8),
    ("total_mass", "<f4", 12),
    ("particle_mass", "<f4", 16),
    ("dynamic_viscosity", "<f4", 20),
    ("simulation_time", "<f4", 24),
    ("target_fps", "<f4", 28),
    ("h", "<f4", 32),
    ("simulation_scale", "<f4", 36),
    ("time_delta", "<f4", 40),
    ("surface_tension_threshold", "<f4", 44),
    ("surface_tension", "<f4", 48),
    ("restitution", "<f4", 52),
    ("K", "<f4", 56),
    ("constant_acceleration", ("<f4", 4), 64),
    ("grid_size_x", "<i4", 80),
    ("grid_size_y", "<i4", 84),
    ("grid_size_z", "<i4", 88),
    ("grid_cell_count", "<u4", 92),
    ("min_point", ("<f4", 4), 96),
    ("max_point", ("<f4", 4), 112),
]
PARAMS = np.dtype(
    {
        "names": [f[0] for f in _PARAM_FIELDS],
        "formats": [f[1] for f in _PARAM_FIELDS],
        "offsets": [f[2] for f in _PARAM_FIELDS],
        "itemsize": 128,
    }
)

TERMS = np.dtype(
    {
        "names": ["poly_6", "poly_6_gradient", "poly_6_laplacian", "spiky", "viscosity"],
        "formats": ["<f4"] * 5,
        "offsets": [0, 4, 8, 12, 16],
        "itemsize": 20,
    }
)


def particles(n):
    """Zeroed particle array of length n."""
    return np.zeros(n, dtype=PARTICLE)


def params():
    return np.zeros((), dtype=PARAMS)


def terms():
    return np.zeros((), dtype=TERMS)
