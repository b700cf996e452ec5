"""The k-sum contraction kernels against numpy (matrix-core lane layouts,
split-K, precision of the split-BF16x3 scheme, the streamed FP32 pair
kernel's 16-byte loads and tails in M, N and K)."""
import numpy as np
import pytest

import grtm_cuda_amd as g

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(256, 256, 512), (200, 40, 300), (8192, 8, 64), (64, 256, 4096),
                                   (130, 37, 1001), (1024, 256, 8192)])
@pytest.mark.parametrize("mode", [-1, 0, 1, 2, 3, 4])
def test_gemm_matches_numpy(mode, shape):
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    print(f"gpu: gemm mode {mode} shape {shape}", flush=True)
    C = g.selftest_gemm(mode, A, B)
    if mode in (1, 2, 3):
        ref = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)
    else:
        ref = A @ B
    err = np.abs(C - ref).max() / (np.abs(A).max() * np.abs(B).max() * np.sqrt(K))
    tol = {-1: 1e-14, 0: 1e-14, 1: 1e-6, 2: 2e-5, 3: 1e-6, 4: 1e-14}[mode]
    assert err < tol, err


def test_fp64_mfma_lane_layout():
    """v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
    B[l / 16][l % 16]; accumulator r of lane l is D[l / 16 + 4 r][l % 16]."""
    import ctypes
    from grtm_cuda_amd import ffi
    for l0 in (0, 1, 15, 16, 33, 63):
        out = np.zeros(256)
        ffi.check(ffi._lib().grtm_selftest_mfma_probe(l0, out.ctypes.data_as(ctypes.c_void_p)))
        i, k = l0 % 16, l0 // 16
        expect = np.zeros(256)
        for lane in range(64):
            for r in range(4):
                if lane // 16 + 4 * r == i:
                    expect[lane * 4 + r] = 1000 + 16 * k + lane % 16
        assert np.array_equal(out, expect), l0
