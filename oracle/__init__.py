"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
bench's cpu_baseline leg, always as the checker. The product path
(libclsph_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

from libclsph_amd import layout

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    sig = {
        "orc_morton": (u32, [u32, u32, u32]),
        "orc_morton_decode": (None, [u32, vp]),
        "orc_grid_params": (i32, [vp, u32, vp]),
        "orc_locate_in_grid": (None, [vp, vp, u32, vp]),
        "orc_sort_particles": (None, [vp, vp, vp, u32]),
        "orc_cell_table": (None, [vp, u32, u32, vp]),
        "orc_density_pressure_f32": (None, [vp, vp, u32, vp, vp, vp]),
        "orc_forces_f32": (None, [vp, vp, u32, vp, vp, vp]),
        "orc_advection_collision_f32": (u32, [vp, vp, u32, vp, vp, vp, vp, u32, u32]),
        "orc_to_f64": (None, [vp, vp, u32]),
        "orc_density_pressure_f64": (None, [vp, vp, u32, vp, vp, vp]),
        "orc_forces_f64": (None, [vp, vp, u32, vp, vp, vp]),
        "orc_advection_collision_f64": (u32, [vp, vp, u32, vp, vp, vp, vp, u32, u32]),
        "orc_step_f32": (i32, [vp, vp, vp, vp, u32, vp, vp, vp, vp, vp, u32, u32, vp]),
        "orc_run_f32": (i32, [vp, u32, vp, vp, vp, vp, vp, u32, u32, vp, vp]),
        "orc_threads": (i32, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

PARTICLE64 = np.dtype(
    {
        "names": ["position", "velocity", "intermediate_velocity", "acceleration",
                  "density", "pressure", "grid_index"],
        "formats": [("<f8", 4), ("<f8", 4), ("<f8", 4), ("<f8", 4), "<f8", "<f8", "<u4"],
        "offsets": [0, 32, 64, 96, 128, 136, 144],
        "itemsize": 152,
    }
)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def morton(x, y, z):
    return lib.orc_morton(x, y, z)


def morton_decode(key):
    out = np.zeros(3, dtype=np.uint32)
    lib.orc_morton_decode(key, _p(out))
    return tuple(int(v) for v in out)


def grid_params(parts, params):
    """Fills params' grid fields from parts; returns the reference assert status."""
    return lib.orc_grid_params(_p(parts), len(parts), _p(params))


def locate_in_grid(parts, params):
    out = np.empty_like(parts)
    lib.orc_locate_in_grid(_p(parts), _p(out), len(parts), _p(params))
    return out


def sort_particles(parts):
    out = np.empty_like(parts)
    scratch = np.empty_like(parts)
    lib.orc_sort_particles(_p(parts), _p(out), _p(scratch), len(parts))
    return out


def cell_table(sorted_parts, cell_count):
    t = np.zeros(cell_count, dtype=np.uint32)
    lib.orc_cell_table(_p(sorted_parts), len(sorted_parts), cell_count, _p(t))
    return t


def density_pressure(sorted_parts, params, terms, table):
    out = np.empty_like(sorted_parts)
    lib.orc_density_pressure_f32(_p(sorted_parts), _p(out), len(out), _p(params), _p(terms),
                                 _p(table))
    return out


def forces(parts, params, terms, table):
    out = np.empty_like(parts)
    lib.orc_forces_f32(_p(parts), _p(out), len(out), _p(params), _p(terms), _p(table))
    return out


def advection_collision(parts, params, normals, verts, indices, max_bounces=64):
    out = np.empty_like(parts)
    normals = np.ascontiguousarray(normals, dtype=np.float32)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    capped = lib.orc_advection_collision_f32(_p(parts), _p(out), len(out), _p(params),
                                             _p(normals), _p(verts), _p(indices),
                                             len(indices) // 3, max_bounces)
    return out, capped


def to_f64(parts):
    out = np.zeros(len(parts), dtype=PARTICLE64)
    lib.orc_to_f64(_p(parts), _p(out), len(parts))
    return out


def density_pressure_f64(parts64, params, terms, table):
    out = np.empty_like(parts64)
    lib.orc_density_pressure_f64(_p(parts64), _p(out), len(out), _p(params), _p(terms),
                                 _p(table))
    return out


def forces_f64(parts64, params, terms, table):
    out = np.empty_like(parts64)
    lib.orc_forces_f64(_p(parts64), _p(out), len(out), _p(params), _p(terms), _p(table))
    return out


def advection_collision_f64(parts64, params, normals, verts, indices, max_bounces=64):
    out = np.empty_like(parts64)
    normals = np.ascontiguousarray(normals, dtype=np.float32)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    capped = lib.orc_advection_collision_f64(_p(parts64), _p(out), len(out), _p(params),
                                             _p(normals), _p(verts), _p(indices),
                                             len(indices) // 3, max_bounces)
    return out, capped


class StepResult:
    pass


def stages(parts, params, terms, normals, verts, indices, max_bounces=64):
    """Runs one substep stage by stage (sph_simulation.cpp:173-344) and returns
    every intermediate: grid params, hashed particles, sorted particles, cell
    table, after density, after forces, after advection, bounce-capped count."""
    r = StepResult()
    p = params.copy()
    r.grid_rc = grid_params(parts, p)
    r.params = p
    r.hashed = locate_in_grid(parts, p)
    r.sorted = sort_particles(r.hashed)
    r.table = cell_table(r.sorted, int(p["grid_cell_count"]))
    r.density = density_pressure(r.sorted, p, terms, r.table)
    r.forces = forces(r.density, p, terms, r.table)
    r.advected, r.capped = advection_collision(r.forces, p, normals, verts, indices,
                                               max_bounces)
    return r


def run(parts, params, terms, normals, verts, indices, steps, max_bounces=64):
    """Runs `steps` substeps in place on a copy; returns (state, mean_density[steps],
    kinetic_energy[steps], params)."""
    state = np.ascontiguousarray(parts.copy())
    p = params.copy()
    rho = np.zeros(steps)
    ke = np.zeros(steps)
    normals = np.ascontiguousarray(normals, dtype=np.float32)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    rc = lib.orc_run_f32(_p(state), steps, _p(p), _p(terms), _p(normals), _p(verts),
                         _p(indices), len(indices) // 3, max_bounces, _p(rho), _p(ke))
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc})")
    return state, rho, ke, p


def threads():
    return lib.orc_threads()
