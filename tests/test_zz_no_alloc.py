"""sqbem_gemm16_noalloc: the 16-bit GEMM that allocates nothing (its workspace is the
library's fixed per-device scratch) against the allocating sqbem_gemm16, through the C ABI.

Both entry points run the same kernels; with the tile autotuner in its fixed-rule mode they
pick the same tile, so the products must agree bit for bit. Two calls in a row, the second
smaller (the scratch keeps the first call's contents), for BF16 and FP16 (whose row absmax
words live in the scratch and must be cleared on every call), on the converted-operand path
(m, n >= 256, k >= 128) and the small path; plus a product whose workspace exceeds the
scratch, which must be rejected without touching C. Also checked against an FP32 reference
at the 16-bit input rounding.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    t = torch.empty_strided(a.shape, (1, a.shape[0]), dtype=torch.float32, device="cuda")
    t.copy_(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)))
    return t


def call(sq, fn, kind, alpha, A, B, beta, C):
    m, k = A.shape
    n = B.shape[1]
    return fn(kind, m, n, k, ctypes.c_float(alpha), ctypes.c_void_p(A.data_ptr()), A.stride(1),
              ctypes.c_void_p(B.data_ptr()), B.stride(1), ctypes.c_float(beta), ctypes.c_void_p(C.data_ptr()),
              C.stride(1))


@pytest.mark.parametrize("kind", [3, 2])  # SQBEM_BF16_IR, SQBEM_FP16_IR
def test_gemm16_noalloc_matches_gemm16(sq, kind):
    lib = sq.lib()
    sq.tune(0)  # fixed rule: both entry points choose the same tile
    try:
        rng = np.random.default_rng(kind)
        for (m, n, k) in [(1100, 900, 700), (300, 260, 130), (200, 70, 50)]:
            print(f"gemm16_noalloc kind={kind} m={m} n={n} k={k}", flush=True)
            A = rng.standard_normal((m, k)) * np.exp(rng.uniform(-3, 3, (m, 1)))
            B = rng.standard_normal((k, n))
            C0 = rng.standard_normal((m, n))
            dA, dB = dev(A), dev(B)
            c_ref, c_new = dev(C0), dev(C0)
            torch.cuda.synchronize()
            assert call(sq, lib.sqbem_gemm16, kind, -0.75, dA, dB, 0.5, c_ref) == 0, sq.last_error()
            assert call(sq, lib.sqbem_gemm16_noalloc, kind, -0.75, dA, dB, 0.5, c_new) == 0, sq.last_error()
            sq.sync()
            got, ref = c_new.cpu().numpy(), c_ref.cpu().numpy()
            assert np.array_equal(got, ref), np.abs(got - ref).max()
            exact = -0.75 * (A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)) \
                + 0.5 * C0.astype(np.float32)
            scale = 0.75 * np.abs(A) @ np.abs(B) + 0.5 * np.abs(C0)
            tol = {3: 2 ** -7, 2: 2 ** -10}[kind]
            assert (np.abs(got - exact) / scale).max() < tol
    finally:
        sq.tune(1)


def test_gemm16_noalloc_rejects_oversized(sq):
    lib = sq.lib()
    m = n = k = 6000  # (6144 + 6144) * 6016 * 2 bytes > 96 MiB
    need = lib.sqbem_gemm16_noalloc_bytes(3, m, n, k)
    assert need > (96 << 20)
    A = torch.zeros((1, 1), dtype=torch.float32, device="cuda")
    rc = lib.sqbem_gemm16_noalloc(3, m, n, k, ctypes.c_float(1.0), ctypes.c_void_p(A.data_ptr()), m,
                                  ctypes.c_void_p(A.data_ptr()), k, ctypes.c_float(0.0),
                                  ctypes.c_void_p(A.data_ptr()), m)
    assert rc == 1 and "scratch" in sq.last_error()
