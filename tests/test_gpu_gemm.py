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


@pytest.mark.parametrize("shape", [(8192, 256, 2048), (1024, 64, 1000), (400, 72, 333), (4, 8, 5), (2048, 130, 4096)])
@pytest.mark.parametrize("prec", [1, 0, 6, 3, 16])
def test_learned_grid_contraction(prec, shape):
    """The learned-grid contraction (FP32: stream-K over the tiles' K slabs;
    FP64: split-K limited to each tile's extent): every operand row read up to
    its frequency row's cut, tiles without K (cut 0 throughout), rows cut at
    full K, unaligned cuts, column tiles of 64, 128 and 256, shapes with
    ragged tails.  Against numpy on the masked operands; the split-BF16 forms
    (prec 6: three bf16 terms per operand, six products; 3: two terms, three
    products) to their truncation, 2^-24 and 2^-16 of each product, and the
    scaled FP16 split (prec 16: two fp16 terms, three products, about 2^-22;
    B in [-1, 1] as the Bessel tables, rows of A over eight decades so the
    per-row scales differ)."""
    M, N, K = shape
    rng = np.random.default_rng(3 * M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    if prec == 16:
        B = rng.uniform(-1.0, 1.0, (K, N))
        A *= np.repeat(10.0 ** rng.uniform(-4, 4, M // 4), 4)[:, None]
    nw = M // 4
    nk_w = rng.integers(0, K + 1, nw)
    nk_w[: min(nw, 40)] = 0                    # leading tiles without K at all
    nk_w[rng.random(nw) < 0.1] = K             # rows at full extent
    nk_w[rng.random(nw) < 0.1] = 1             # single-wavenumber rows
    print(f"gpu: learned-grid contraction prec {prec} shape {shape}", flush=True)
    C = g.selftest_gemm_cut(prec, A, B, nk_w)
    if prec != 0:
        A = A.astype(np.float32).astype(np.float64)
        B = B.astype(np.float32).astype(np.float64)
    mask = np.arange(K)[None, :] < np.repeat(nk_w, 4)[:, None]
    Am = np.where(mask, A, 0.0)
    ref = Am @ B
    bound = np.abs(Am) @ np.abs(B)  # per-element scale of the rounding
    # FP32 / FP64 accumulation over up to 4096 terms; two-term splits drop
    # products near 3 2^-16 of each term
    tol = {0: 1e-13, 1: 2e-5, 6: 2e-5, 3: 1.5e-4, 16: 2e-5}[prec]
    assert np.all(np.abs(C - ref) <= tol * bound + 1e-300), np.max(np.abs(C - ref) / (bound + 1e-300))
    assert np.all(C[np.repeat(nk_w, 4) == 0] == 0.0)


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
