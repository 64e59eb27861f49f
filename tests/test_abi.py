"""The C ABI library loads, exports every entry point include/*.h declares,
and its host-only helpers (no device involved) are correct."""
import glob
import math
import os
import re
import subprocess

import numpy as np

from libclsph_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(clsph_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    for n in names:
        getattr(native.lib, n)  # resolvable through ctypes


def kernel_terms_f64(h):
    """sph_simulation.cpp:499-505 as g++/libstdc++ compiles it: pow(h, 9.f)
    unqualified is ::pow(double, double), the arithmetic is double, and each
    value is rounded once to float (numpy's float32 cast rounds to nearest)."""
    hd = float(np.float32(h))
    h9, h6 = math.pow(hd, 9.0), math.pow(hd, 6.0)
    return {
        "poly_6": np.float32(315.0 / (64.0 * math.pi * h9)),
        "poly_6_gradient": np.float32(-945.0 / (32.0 * math.pi * h9)),
        "poly_6_laplacian": np.float32(-945.0 / (32.0 * math.pi * h9)),
        "spiky": np.float32(-45.0 / (math.pi * h6)),
        "viscosity": np.float32(45.0 / (math.pi * h6)),
    }


def test_kernel_terms_bit_exact_over_many_h():
    rng = np.random.default_rng(3)
    hs = np.concatenate([rng.uniform(1e-3, 0.2, 4000), np.float32([0.0123, 0.04, 9.921e-3])])
    powf_differs = 0
    for h in hs.astype(np.float32):
        t = native.kernel_terms(float(h))
        want = kernel_terms_f64(h)
        for k, v in want.items():
            assert t[k].tobytes() == v.tobytes(), (float(h), k, float(t[k]), float(v))
        # float pow would round h^9 to float first: it differs for many h,
        # so this test tells the two readings of the reference apart
        h9f = np.float32(math.pow(float(h), 9.0))
        powf_differs += np.float32(315.0 / (64.0 * math.pi * float(h9f))) != want["poly_6"]
    assert powf_differs > len(hs) // 10


def test_make_box_inward_unit_normals():
    lo, hi = (0.0, -1.0, 2.0), (1.0, 3.0, 2.5)
    normals, verts, idx = native.make_box(lo, hi)
    V = verts.reshape(-1, 3)
    N = normals.reshape(-1, 3)
    centre = (np.array(lo) + np.array(hi)) / 2
    assert idx.max() == 7 and len(idx) == 36
    np.testing.assert_allclose(np.linalg.norm(N, axis=1), 1.0, rtol=1e-6)
    area = 0.0
    for f, t in enumerate(idx.reshape(-1, 3)):
        a, b, c = V[t]
        n = np.cross(b - a, c - a)
        np.testing.assert_allclose(n / np.linalg.norm(n), N[f], atol=1e-6)
        assert N[f] @ (centre - (a + b + c) / 3) > 0  # points into the box
        area += np.linalg.norm(n) / 2
    d = np.array(hi) - np.array(lo)
    np.testing.assert_allclose(area, 2 * (d[0] * d[1] + d[1] * d[2] + d[0] * d[2]), rtol=1e-6)


def test_make_block_lattice_and_jitter():
    s = 0.02
    parts = native.make_block(1000, s, (0.1, 0.2, 0.3), (0.2, 0.0, 0.1), 0.01, 42)
    assert len(parts) == 1000
    P = parts["position"][:, :3].astype(np.float64)
    # 10 sites along x, 5 along z, layers stacked in y
    site = np.round((P - np.array([0.1, 0.2, 0.3])) / s - 0.5)
    assert site[:, 0].max() == 9 and site[:, 2].max() == 4 and site[:, 1].max() == 19
    off = P - (np.array([0.1, 0.2, 0.3]) + (site + 0.5) * s)
    assert np.abs(off).max() <= 0.01 * s * (1 + 1e-4)
    assert np.abs(off).max() > 0.005 * s  # jitter present
    # x fastest, then z, then y
    np.testing.assert_array_equal(site[:12, 0], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 0, 1])
    assert site[10, 2] == 1 and site[50, 1] == 1
    again = native.make_block(1000, s, (0.1, 0.2, 0.3), (0.2, 0.0, 0.1), 0.01, 42)
    np.testing.assert_array_equal(again["position"], parts["position"])
    other = native.make_block(1000, s, (0.1, 0.2, 0.3), (0.2, 0.0, 0.1), 0.01, 43)
    assert not np.array_equal(other["position"], parts["position"])


def test_context_without_device_fails_loudly():
    # On a machine with no GPU, creating a context raises instead of falling back.
    import ctypes
    h = ctypes.c_void_p()
    rc = native.lib.clsph_init(0, ctypes.byref(h))
    if rc == 0:  # a GPU is present: nothing to check here
        native.lib.clsph_free(h)
        return
    assert rc == -2
    assert native.lib.clsph_last_error(None)
