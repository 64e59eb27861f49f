"""ctypes binding of the clsph HIP library (include/clsph_hip.h).

The library is built in-tree (libclsph_amd/build.py) as libclsph_hip.so.
Importing this module without it raises immediately: there is no CPU
fallback for any entry point that computes.
"""
import ctypes
import os

import numpy as np

from . import layout

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclsph_hip.so")

CLSPH_OK = 0
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "EGRID", -5: "ETABLE"}

STAGES = {"grid": 0, "hash": 1, "sort": 2, "cell_table": 3, "density": 4, "forces": 5,
          "advect": 6}

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int


class ClsphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"clsph error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m libclsph_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "clsph_init": (_i32, [_i32, ctypes.POINTER(_vp)]),
        "clsph_free": (_i32, [_vp]),
        "clsph_last_error": (ctypes.c_char_p, [_vp]),
        "clsph_device_name": (_i32, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
        "clsph_sync": (_i32, [_vp]),
        "clsph_state_alloc": (_i32, [_vp, _u32, _u32, ctypes.POINTER(_vp)]),
        "clsph_state_free": (_i32, [_vp]),
        "clsph_set_params": (_i32, [_vp, _vp, _vp, _u32]),
        "clsph_upload_aos": (_i32, [_vp, _vp, _u32]),
        "clsph_download_aos": (_i32, [_vp, _vp, _u32]),
        "clsph_download_sorted_aos": (_i32, [_vp, _vp, _u32]),
        "clsph_scene_upload": (_i32, [_vp, _vp, _vp, _u32, _vp, _u32]),
        "clsph_grid": (_i32, [_vp]),
        "clsph_hash": (_i32, [_vp]),
        "clsph_sort": (_i32, [_vp]),
        "clsph_cell_table": (_i32, [_vp]),
        "clsph_density_pressure": (_i32, [_vp]),
        "clsph_forces": (_i32, [_vp]),
        "clsph_advect_collide": (_i32, [_vp]),
        "clsph_step": (_i32, [_vp, _u32]),
        "clsph_get_grid": (_i32, [_vp, _vp]),
        "clsph_read_keys": (_i32, [_vp, _vp, _u32]),
        "clsph_read_sorted": (_i32, [_vp, _vp, _vp, _u32]),
        "clsph_read_cell_table": (_i32, [_vp, _vp, _u32]),
        "clsph_read_order": (_i32, [_vp, _vp, _vp, _u32]),
        "clsph_bounce_capped": (_i32, [_vp, _vp]),
        "clsph_time_steps": (_i32, [_vp, _u32, _vp]),
        "clsph_time_stage": (_i32, [_vp, _i32, _u32, _vp]),
        "clsph_profile_steps": (_i32, [_vp, _u32, _vp]),
        "clsph_make_block": (_u32, [_vp, _u32, ctypes.c_float, _vp, _vp, ctypes.c_float,
                                    ctypes.c_uint64]),
        "clsph_kernel_terms": (None, [ctypes.c_float, _vp]),
        "clsph_make_box": (None, [_vp, _vp, _vp, _vp, _vp]),
        "clsph_scene_grid": (_i32, [_vp, ctypes.c_float]),
        "clsph_batch_alloc": (_i32, [_vp, _u32, _vp, _u32, ctypes.POINTER(_vp)]),
        "clsph_batch_set_params": (_i32, [_vp, _u32, _vp, _vp, _u32]),
        "clsph_batch_upload": (_i32, [_vp, _u32, _vp, _u32]),
        "clsph_batch_download": (_i32, [_vp, _u32, _vp, _u32]),
        "clsph_batch_scene_upload": (_i32, [_vp, _u32, _vp, _vp, _u32, _vp, _u32]),
        "clsph_batch_get_grid": (_i32, [_vp, _u32, _vp]),
        "clsph_batch_bounce_capped": (_i32, [_vp, _u32, _vp]),
        "clsph_set_block_sizes": (_i32, [_vp, _vp]),
        "clsph_get_block_sizes": (_i32, [_vp, _vp]),
        "clsph_make_heightfield": (None, [_u32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_float, _u32, ctypes.c_uint64, _vp, _vp, _vp]),
        "clsph_make_column": (_u32, [_vp, _u32, ctypes.c_float, _vp, _vp, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_uint64]),
        "clsph_nn_param_count": (_u32, []),
        "clsph_nn_create": (_i32, [_vp, _u32, _u32, ctypes.POINTER(_vp)]),
        "clsph_nn_free": (_i32, [_vp]),
        "clsph_nn_set_params": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_get_params": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_forward": (_i32, [_vp, _vp, _vp]),
        "clsph_nn_infer": (_i32, [_vp, _vp]),
        "clsph_neighbour_pairs": (_i32, [_vp, _vp]),
        "clsph_learned_run": (_i32, [_vp, _vp, _vp, _u32, _u32, ctypes.c_float, _vp, _vp, _vp,
                                     _vp]),
        "clsph_nn_profile": (_i32, [_vp, _vp, _u32, _vp, _vp]),
        "clsph_nn_read_graph": (_i32, [_vp, _vp, _vp, _u32, _u32]),
        "clsph_nn_read_output": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_set_target": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_read_target": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_target_from": (_i32, [_vp, _vp, _vp, ctypes.c_float]),
        "clsph_nn_target_mean_acc": (_i32, [_vp, _vp, _vp, _u32]),
        "clsph_nn_backward": (_i32, [_vp, _vp]),
        "clsph_nn_read_grad": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_adam": (_i32, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float]),
        "clsph_learned_step": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_compare_density": (_i32, [_vp, _vp, _vp, _vp, _vp]),
        "clsph_state_copy": (_i32, [_vp, _vp]),
        "clsph_track_ids": (_i32, [_vp, _i32]),
        "clsph_read_ids": (_i32, [_vp, _vp, _u32]),
        "clsph_nn_gemm": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _u32, _u32, _u32]),
        "clsph_nn_gemm_noalloc": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _u32, _u32,
                                         _u32]),
        "clsph_gnn_param_count": (_u32, []),
        "clsph_gnn_create": (_i32, [_vp, _u32, _u32, ctypes.POINTER(_vp)]),
        "clsph_gnn_free": (_i32, [_vp]),
        "clsph_gnn_set_params": (_i32, [_vp, _vp, _u32]),
        "clsph_gnn_get_params": (_i32, [_vp, _vp, _u32]),
        "clsph_gnn_forward": (_i32, [_vp, _vp, _vp]),
        "clsph_gnn_infer": (_i32, [_vp, _vp]),
        "clsph_gnn_profile": (_i32, [_vp, _vp, _u32, _vp, _vp]),
        "clsph_gnn_read_graph": (_i32, [_vp, _vp, _vp, _u32, _u32]),
        "clsph_gnn_read_output": (_i32, [_vp, _vp, _u32]),
        "clsph_gnn_set_target": (_i32, [_vp, _vp, _u32]),
        "clsph_gnn_read_target": (_i32, [_vp, _vp, _u32]),
        "clsph_gnn_backward": (_i32, [_vp, _vp]),
        "clsph_gnn_read_grad": (_i32, [_vp, _vp, _u32]),
        "clsph_gnn_adam": (_i32, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------- host-only helpers (no device needed) ----------------

def kernel_terms(h):
    t = layout.terms()
    lib.clsph_kernel_terms(ctypes.c_float(h), _ptr(t))
    return t


def make_block(n, spacing, origin, extent, jitter, seed):
    out = layout.particles(n)
    o = np.asarray(origin, dtype=np.float32)
    e = np.asarray(extent, dtype=np.float32)
    got = lib.clsph_make_block(_ptr(out), n, ctypes.c_float(spacing), _ptr(o), _ptr(e),
                               ctypes.c_float(jitter), ctypes.c_uint64(seed))
    return out[:got]


def make_box(lo, hi):
    lo = np.asarray(lo, dtype=np.float32)
    hi = np.asarray(hi, dtype=np.float32)
    verts = np.zeros(24, dtype=np.float32)
    idx = np.zeros(36, dtype=np.uint32)
    nrm = np.zeros(36, dtype=np.float32)
    lib.clsph_make_box(_ptr(lo), _ptr(hi), _ptr(verts), _ptr(idx), _ptr(nrm))
    return nrm, verts, idx


def make_heightfield(res, x0, z0, size, amplitude, octaves, seed):
    verts = np.zeros(3 * (res + 1) ** 2, dtype=np.float32)
    idx = np.zeros(6 * res * res, dtype=np.uint32)
    nrm = np.zeros(6 * res * res, dtype=np.float32)
    lib.clsph_make_heightfield(res, ctypes.c_float(x0), ctypes.c_float(z0), ctypes.c_float(size),
                               ctypes.c_float(amplitude), octaves, ctypes.c_uint64(seed),
                               _ptr(verts), _ptr(idx), _ptr(nrm))
    return nrm, verts, idx


def make_column(n, spacing, origin, extent, jitter, vh, v0, v1, seed):
    out = layout.particles(n)
    o = np.asarray(origin, dtype=np.float32)
    e = np.asarray(extent, dtype=np.float32)
    got = lib.clsph_make_column(_ptr(out), n, ctypes.c_float(spacing), _ptr(o), _ptr(e),
                                ctypes.c_float(jitter), ctypes.c_float(vh), ctypes.c_float(v0),
                                ctypes.c_float(v1), ctypes.c_uint64(seed))
    return out[:got]


# ---------------- device context / state ----------------

class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib.clsph_init(device, ctypes.byref(h))
        if rc != CLSPH_OK:
            raise ClsphError(rc, lib.clsph_last_error(None).decode())
        self.h = h

    def error(self):
        return lib.clsph_last_error(self.h).decode()

    def check(self, rc):
        if rc != CLSPH_OK:
            raise ClsphError(rc, self.error())

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        self.check(lib.clsph_device_name(self.h, buf, 256))
        return buf.value.decode()

    def sync(self):
        self.check(lib.clsph_sync(self.h))

    def close(self):
        if self.h:
            lib.clsph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class State:
    """A device-resident particle system (one simulation)."""

    def __init__(self, ctx, n_cap, table_cap=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(lib.clsph_state_alloc(ctx.h, n_cap, table_cap, ctypes.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        # freeing needs the context's stream: after the context is closed the
        # buffers are left to the end of the process
        if self.h and self.ctx.h:
            lib.clsph_state_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, terms, max_bounces=0):
        self.ctx.check(lib.clsph_set_params(self.h, _ptr(params), _ptr(terms), max_bounces))

    def block_sizes(self):
        b = np.zeros(7, dtype=np.uint32)
        self.ctx.check(lib.clsph_get_block_sizes(self.h, _ptr(b)))
        return [int(v) for v in b]

    def set_block_sizes(self, sizes):
        b = np.asarray(sizes, dtype=np.uint32)
        assert b.shape == (7,)
        self.ctx.check(lib.clsph_set_block_sizes(self.h, _ptr(b)))

    def upload(self, parts):
        parts = np.ascontiguousarray(parts, dtype=layout.PARTICLE)
        self.ctx.check(lib.clsph_upload_aos(self.h, _ptr(parts), len(parts)))
        self.n = len(parts)

    def download(self):
        out = layout.particles(self.n)
        self.ctx.check(lib.clsph_download_aos(self.h, _ptr(out), self.n))
        return out

    def download_sorted(self):
        out = layout.particles(self.n)
        self.ctx.check(lib.clsph_download_sorted_aos(self.h, _ptr(out), self.n))
        return out

    def scene(self, normals, vertices, indices):
        normals = np.ascontiguousarray(normals, dtype=np.float32)
        vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.ctx.check(lib.clsph_scene_upload(self.h, _ptr(normals), _ptr(vertices),
                                              len(vertices) // 3, _ptr(indices),
                                              len(indices) // 3))

    def scene_grid(self, cell_size):
        """Triangle grid for collisions: > 0 cell size, 0 automatic, < 0 none."""
        self.ctx.check(lib.clsph_scene_grid(self.h, ctypes.c_float(cell_size)))

    def stage(self, name):
        fn = {"grid": lib.clsph_grid, "hash": lib.clsph_hash, "sort": lib.clsph_sort,
              "cell_table": lib.clsph_cell_table, "density": lib.clsph_density_pressure,
              "forces": lib.clsph_forces, "advect": lib.clsph_advect_collide}[name]
        self.ctx.check(fn(self.h))

    def step(self, nsteps=1):
        self.ctx.check(lib.clsph_step(self.h, nsteps))

    def grid(self):
        p = layout.params()
        self.ctx.check(lib.clsph_get_grid(self.h, _ptr(p)))
        return p

    def keys(self):
        k = np.zeros(self.n, dtype=np.uint32)
        self.ctx.check(lib.clsph_read_keys(self.h, _ptr(k), self.n))
        return k

    def sorted(self):
        k = np.zeros(self.n, dtype=np.uint32)
        p = np.zeros(self.n, dtype=np.uint32)
        self.ctx.check(lib.clsph_read_sorted(self.h, _ptr(k), _ptr(p), self.n))
        return k, p

    def order(self):
        """(slot, perm) of the kernels' internal order: internal position k
        is reference sorted position slot[k] and canonical index perm[k]."""
        slot = np.zeros(self.n, dtype=np.uint32)
        perm = np.zeros(self.n, dtype=np.uint32)
        self.ctx.check(lib.clsph_read_order(self.h, _ptr(slot), _ptr(perm), self.n))
        return slot, perm

    def cell_table(self, count):
        t = np.zeros(count, dtype=np.uint32)
        self.ctx.check(lib.clsph_read_cell_table(self.h, _ptr(t), count))
        return t

    def bounce_capped(self):
        c = np.zeros(1, dtype=np.uint32)
        self.ctx.check(lib.clsph_bounce_capped(self.h, _ptr(c)))
        return int(c[0])

    def time_steps(self, nsteps):
        ms = np.zeros(1, dtype=np.float32)
        self.ctx.check(lib.clsph_time_steps(self.h, nsteps, _ptr(ms)))
        return float(ms[0])

    def profile_steps(self, nsteps):
        """Average ms per substep of each stage, keyed by stage name."""
        ms = np.zeros(7, dtype=np.float32)
        self.ctx.check(lib.clsph_profile_steps(self.h, nsteps, _ptr(ms)))
        return {name: float(ms[i]) for name, i in STAGES.items()}

    def neighbour_pairs(self):
        """Ordered neighbour pairs (r < h) found by the last density stage."""
        e = np.zeros(1, dtype=np.uint64)
        self.ctx.check(lib.clsph_neighbour_pairs(self.h, _ptr(e)))
        return int(e[0])

    def time_stage(self, name, reps):
        ms = np.zeros(1, dtype=np.float32)
        self.ctx.check(lib.clsph_time_stage(self.h, STAGES[name], reps, _ptr(ms)))
        return float(ms[0])

    def copy_from(self, other):
        """Take the particles (and per-particle fields) of another state."""
        self.ctx.check(lib.clsph_state_copy(self.h, other.h))
        self.n = other.n

    def track_ids(self, enable=True):
        """Reset identities to the current canonical index and keep them
        through substeps (False stops tracking)."""
        self.ctx.check(lib.clsph_track_ids(self.h, 1 if enable else 0))

    def ids(self):
        out = np.zeros(self.n, dtype=np.uint32)
        self.ctx.check(lib.clsph_read_ids(self.h, _ptr(out), self.n))
        return out

    def learned_step(self, net, nsteps=1):
        """Substeps with the network in place of the forces stage."""
        self.ctx.check(lib.clsph_learned_step(self.h, net.h, nsteps))


def nn_gemm(ctx, a_bits, b_bits, c, m, n, k, trans_a=False, trans_b=False, accumulate=False,
            bias=None, noalloc=False):
    """The network's bf16 GEMM on host arrays (bf16 as uint16 bit patterns);
    returns the float32 result (c updated in place). `noalloc` stages the
    operands in the library's static device workspace (clsph_nn_gemm_noalloc)
    instead of device allocations."""
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint16)
    b_bits = np.ascontiguousarray(b_bits, dtype=np.uint16)
    c = np.ascontiguousarray(c, dtype=np.float32)
    bias = None if bias is None else np.ascontiguousarray(bias, dtype=np.float32)
    fn = lib.clsph_nn_gemm_noalloc if noalloc else lib.clsph_nn_gemm
    ctx.check(fn(ctx.h, int(trans_a), int(trans_b), int(accumulate), _ptr(a_bits), _ptr(b_bits),
                 _ptr(c), _ptr(bias), m, n, k))
    return c


class Network:
    """The learned force surrogate's device workspace and parameters."""

    def __init__(self, ctx, n_cap, e_cap=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(lib.clsph_nn_create(ctx.h, n_cap, e_cap, ctypes.byref(h)))
        self.h = h
        self.count = int(lib.clsph_nn_param_count())
        self.n = 0
        self.e = 0

    def close(self):
        if self.h and self.ctx.h:
            lib.clsph_nn_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.shape == (self.count,)
        self.ctx.check(lib.clsph_nn_set_params(self.h, _ptr(flat), self.count))

    def params(self):
        out = np.zeros(self.count, dtype=np.float32)
        self.ctx.check(lib.clsph_nn_get_params(self.h, _ptr(out), self.count))
        return out

    def forward(self, state):
        """Training evaluation on a state whose last stage was forces (the
        state is not changed); returns the edge count."""
        e = np.zeros(1, dtype=np.uint32)
        self.ctx.check(lib.clsph_nn_forward(self.h, state.h, _ptr(e)))
        self.n, self.e = state.n, int(e[0])
        return self.e

    def infer(self, state):
        """Fused inference on a state whose last stage was forces: adds the
        correction to its accelerations (once per forces stage)."""
        self.ctx.check(lib.clsph_nn_infer(self.h, state.h))
        self.n = state.n

    def profile(self, state, reps=5):
        """Average ms of one fused evaluation over `reps` runs (on a copy of
        the state's accelerations), and the edges per pass."""
        ms = np.zeros(1, dtype=np.float32)
        e = np.zeros(1, dtype=np.uint64)
        self.ctx.check(lib.clsph_nn_profile(self.h, state.h, reps, _ptr(ms), _ptr(e)))
        self.n = state.n
        return float(ms[0]), int(e[0])

    def graph(self):
        off = np.zeros(self.n + 1, dtype=np.uint32)
        src = np.zeros(self.e, dtype=np.uint32)
        self.ctx.check(lib.clsph_nn_read_graph(self.h, _ptr(off), _ptr(src), self.n, self.e))
        return off, src

    def output(self):
        out = np.zeros((self.n, 32), dtype=np.float32)
        self.ctx.check(lib.clsph_nn_read_output(self.h, _ptr(out), self.n))
        return out

    def set_target(self, target):
        t = np.ascontiguousarray(target, dtype=np.float32)
        assert t.shape == (self.n, 3)
        self.ctx.check(lib.clsph_nn_set_target(self.h, _ptr(t), self.n))

    def target(self, n=None):
        n = self.n if n is None else n
        t = np.zeros((n, 3), dtype=np.float32)
        self.ctx.check(lib.clsph_nn_read_target(self.h, _ptr(t), n))
        return t

    def target_from(self, learner, teacher, dt):
        self.ctx.check(lib.clsph_nn_target_from(self.h, learner.h, teacher.h, ctypes.c_float(dt)))

    def target_mean_acc(self, learner, teacher, nsub):
        """Target: (mean forces-stage acceleration over nsub plain substeps of
        `teacher`, which they advance, minus the learner's own) / 2."""
        self.ctx.check(lib.clsph_nn_target_mean_acc(self.h, learner.h, teacher.h, nsub))
        self.n = learner.n

    def backward(self, read_loss=True):
        loss = np.zeros(1, dtype=np.float32)
        self.ctx.check(lib.clsph_nn_backward(self.h, _ptr(loss) if read_loss else None))
        return float(loss[0]) if read_loss else None

    def grad(self):
        out = np.zeros(self.count, dtype=np.float32)
        self.ctx.check(lib.clsph_nn_read_grad(self.h, _ptr(out), self.count))
        return out

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.check(lib.clsph_nn_adam(self.h, ctypes.c_float(lr), ctypes.c_float(beta1),
                                         ctypes.c_float(beta2), ctypes.c_float(eps)))

    def learned_run(self, learner, judge, macro_steps, judge_every=8, tol=0.01):
        """Judged learned run (clsph_learned_run); returns (judged,
        fallbacks, [max error of each judged step], [mean error of each])."""
        nj = (macro_steps + judge_every - 1) // judge_every if judge_every else 0
        errs = np.zeros(max(nj, 1), dtype=np.float32)
        means = np.zeros(max(nj, 1), dtype=np.float32)
        counts = np.zeros(2, dtype=np.uint32)
        self.ctx.check(lib.clsph_learned_run(
            learner.h, judge.h if judge is not None else None, self.h, macro_steps,
            judge_every, ctypes.c_float(tol), _ptr(counts[0:1]), _ptr(counts[1:2]), _ptr(errs),
            _ptr(means)))
        k = int(counts[0])
        return (k, int(counts[1]), [float(e) for e in errs[:k]], [float(e) for e in means[:k]])

    def compare_density(self, a, b):
        """(max, mean) of |rho_a - rho_b| / rho0 over the same particles."""
        mx = np.zeros(1, dtype=np.float32)
        mn = np.zeros(1, dtype=np.float32)
        self.ctx.check(lib.clsph_nn_compare_density(self.h, a.h, b.h, _ptr(mx), _ptr(mn)))
        return float(mx[0]), float(mn[0])


class TwoRoundNetwork:
    """The two-round message-passing network's device workspace and
    parameters (clsph_gnn_*)."""

    def __init__(self, ctx, n_cap, e_cap=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(lib.clsph_gnn_create(ctx.h, n_cap, e_cap, ctypes.byref(h)))
        self.h = h
        self.count = int(lib.clsph_gnn_param_count())
        self.n = 0
        self.e = 0

    def close(self):
        if self.h and self.ctx.h:
            lib.clsph_gnn_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.shape == (self.count,)
        self.ctx.check(lib.clsph_gnn_set_params(self.h, _ptr(flat), self.count))

    def params(self):
        out = np.zeros(self.count, dtype=np.float32)
        self.ctx.check(lib.clsph_gnn_get_params(self.h, _ptr(out), self.count))
        return out

    def forward(self, state):
        """Training evaluation on a state whose last stage was density: its
        accelerations become the network's; returns the edge count."""
        e = np.zeros(1, dtype=np.uint32)
        self.ctx.check(lib.clsph_gnn_forward(self.h, state.h, _ptr(e)))
        self.n, self.e = state.n, int(e[0])
        return self.e

    def infer(self, state):
        """Fused inference on a state whose last stage was density."""
        self.ctx.check(lib.clsph_gnn_infer(self.h, state.h))
        self.n = state.n

    def profile(self, state, reps=5):
        """Average ms of edge round 1, node round 1, edge round 2 and node
        round 2 of the fused evaluation over `reps` runs, and the edges per
        pass."""
        ms = np.zeros(4, dtype=np.float32)
        e = np.zeros(1, dtype=np.uint64)
        self.ctx.check(lib.clsph_gnn_profile(self.h, state.h, reps, _ptr(ms), _ptr(e)))
        self.n = state.n
        return [float(x) for x in ms], int(e[0])

    def graph(self):
        off = np.zeros(self.n + 1, dtype=np.uint32)
        src = np.zeros(self.e, dtype=np.uint32)
        self.ctx.check(lib.clsph_gnn_read_graph(self.h, _ptr(off), _ptr(src), self.n, self.e))
        return off, src

    def output(self):
        out = np.zeros((self.n, 32), dtype=np.float32)
        self.ctx.check(lib.clsph_gnn_read_output(self.h, _ptr(out), self.n))
        return out

    def set_target(self, target):
        t = np.ascontiguousarray(target, dtype=np.float32)
        assert t.shape == (self.n, 3)
        self.ctx.check(lib.clsph_gnn_set_target(self.h, _ptr(t), self.n))

    def target(self, n=None):
        n = self.n if n is None else n
        t = np.zeros((n, 3), dtype=np.float32)
        self.ctx.check(lib.clsph_gnn_read_target(self.h, _ptr(t), n))
        return t

    def backward(self, read_loss=True):
        loss = np.zeros(1, dtype=np.float32)
        self.ctx.check(lib.clsph_gnn_backward(self.h, _ptr(loss) if read_loss else None))
        return float(loss[0]) if read_loss else None

    def grad(self):
        out = np.zeros(self.count, dtype=np.float32)
        self.ctx.check(lib.clsph_gnn_read_grad(self.h, _ptr(out), self.count))
        return out

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.check(lib.clsph_gnn_adam(self.h, ctypes.c_float(lr), ctypes.c_float(beta1),
                                          ctypes.c_float(beta2), ctypes.c_float(eps)))


class Batch(State):
    """Independent scenes of fixed particle counts advanced together: every
    stage runs as one launch over all scenes."""

    def __init__(self, ctx, counts, table_cap=0):
        self.ctx = ctx
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        h = ctypes.c_void_p()
        ctx.check(lib.clsph_batch_alloc(ctx.h, len(counts), _ptr(counts), table_cap,
                                        ctypes.byref(h)))
        self.h = h
        self.counts = [int(c) for c in counts]
        self.n = int(counts.sum())

    def set_scene_params(self, scene, params, terms, max_bounces=0):
        self.ctx.check(lib.clsph_batch_set_params(self.h, scene, _ptr(params), _ptr(terms),
                                                  max_bounces))

    def upload_scene(self, scene, parts):
        parts = np.ascontiguousarray(parts, dtype=layout.PARTICLE)
        self.ctx.check(lib.clsph_batch_upload(self.h, scene, _ptr(parts), len(parts)))

    def download_scene(self, scene):
        out = layout.particles(self.counts[scene])
        self.ctx.check(lib.clsph_batch_download(self.h, scene, _ptr(out), len(out)))
        return out

    def mesh_scene(self, scene, normals, vertices, indices):
        normals = np.ascontiguousarray(normals, dtype=np.float32)
        vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.ctx.check(lib.clsph_batch_scene_upload(self.h, scene, _ptr(normals), _ptr(vertices),
                                                    len(vertices) // 3, _ptr(indices),
                                                    len(indices) // 3))

    def grid_scene(self, scene):
        p = layout.params()
        self.ctx.check(lib.clsph_batch_get_grid(self.h, scene, _ptr(p)))
        return p

    def bounce_capped_scene(self, scene):
        c = np.zeros(1, dtype=np.uint32)
        self.ctx.check(lib.clsph_batch_bounce_capped(self.h, scene, _ptr(c)))
        return int(c[0])
