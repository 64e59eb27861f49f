"""clsph_nn_gemm_noalloc (the GEMM entry point staged in the library's static
8 MiB device workspace instead of device allocations) against clsph_nn_gemm
on the same host operands: bit-identical results on two calls in a row, the
second smaller than the first, with and without bias and accumulation, and
a clear error for operands beyond the workspace."""
import numpy as np
import pytest

from libclsph_amd import native
from oracle import learned as L

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    return L.bf16_bits(x)


def operands(rng, m, n, k, ta, tb, acc, with_bias):
    a = rng.normal(size=(k, m) if ta else (m, k)).astype(np.float32)
    b = rng.normal(size=(n, k) if tb else (k, n)).astype(np.float32)
    c0 = rng.normal(size=(m, n)).astype(np.float32) if acc else np.zeros((m, n), np.float32)
    bias = rng.normal(size=n).astype(np.float32) if with_bias else None
    return bf16_bits(a), bf16_bits(b), c0, bias


def test_noalloc_gemm_matches_allocating_gemm():
    ctx = native.Context(0)
    try:
        rng = np.random.default_rng(42)
        # (m, n, k, trans_a, trans_b, accumulate, bias): the second of each
        # pair is smaller than the first
        calls = [(640, 256, 1024, False, False, False, True), (64, 32, 96, False, False, False, True),
                 (512, 128, 256, True, False, True, False), (40, 16, 72, True, False, True, False),
                 (256, 384, 200, False, True, False, False), (8, 8, 8, False, True, False, False)]
        for m, n, k, ta, tb, acc, wb in calls:
            print(f"case m={m} n={n} k={k} ta={ta} tb={tb} acc={acc} bias={wb}", flush=True)
            ab, bb, c0, bias = operands(rng, m, n, k, ta, tb, acc, wb)
            want = native.nn_gemm(ctx, ab, bb, c0.copy(), m, n, k, ta, tb, acc, bias)
            got = native.nn_gemm(ctx, ab, bb, c0.copy(), m, n, k, ta, tb, acc, bias, noalloc=True)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        # operands beyond the 8 MiB workspace: a clear error, nothing run
        m = n = k = 1024  # a, b 2 MiB each, c 4 MiB, bias 4 KiB: 8 MiB + 4 KiB
        ab, bb, c0, bias = operands(rng, m, n, k, False, False, False, True)
        with pytest.raises(native.ClsphError, match="workspace"):
            native.nn_gemm(ctx, ab, bb, c0, m, n, k, bias=bias, noalloc=True)
        # and the workspace still serves a call afterwards
        ab, bb, c0, bias = operands(rng, 16, 16, 16, False, False, False, True)
        want = native.nn_gemm(ctx, ab, bb, c0.copy(), 16, 16, 16, bias=bias)
        got = native.nn_gemm(ctx, ab, bb, c0.copy(), 16, 16, 16, bias=bias, noalloc=True)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    finally:
        ctx.close()
