"""gtf_gemm_split_noalloc (the split-BF16 product with its scratch in a
static device array, for callers that cannot allocate) against the
allocating entry point gtf_gemm_split_host: bitwise equal results on two
calls in a row, the second with a smaller matrix, through the C ABI on
device buffers the caller owns (allocated with the library's own HIP
runtime)."""
import ctypes

import numpy as np
import pytest

gpu = pytest.mark.gpu


@gpu
def test_split_gemm_noalloc_matches_allocating_entry_point(hip):
    import gtfock_amd as g
    L = g.lib()
    L.gtf_gemm_split_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.gtf_gemm_split_noalloc.restype = ctypes.c_int
    rng = np.random.default_rng(77)
    for n, transB, ns in ((846, 0, 5), (250, 1, 6)):
        ld = (n + 63) // 64 * 64
        A = rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-4, 1, size=(n, 1))
        B = rng.normal(size=(n, n))
        ref = g.dgemm(A, B, transB=bool(transB), split=ns)
        Ap = np.zeros((ld, ld))
        Bp = np.zeros((ld, ld))
        Ap[:n, :n] = A
        Bp[:n, :n] = B
        Cp = np.full((ld, ld), np.nan)
        dA, dB, dC = hip.malloc(Ap.nbytes), hip.malloc(Bp.nbytes), hip.malloc(Cp.nbytes)
        try:
            hip.upload(dA, Ap)
            hip.upload(dB, Bp)
            hip.upload(dC, Cp)
            st = L.gtf_gemm_split_noalloc(n, ld, dA, dB, dC, transB, ns, None)
            assert st == 0
            hip.sync()
            hip.download(Cp, dC)
        finally:
            for p in (dA, dB, dC):
                hip.free(p)
        C = Cp[:n, :n]
        print("n %d transB %d slices %d: max |C_noalloc - C_alloc| %.1e" % (n, transB, ns, np.abs(C - ref).max()))
        assert np.array_equal(C, ref)
    # sizes beyond the static scratch are refused, nothing launched
    big = hip.malloc(64 * 64 * 8)
    try:
        assert L.gtf_gemm_split_noalloc(2000, 2048, big, big, big, 0, 5, None) == -3
    finally:
        hip.free(big)
