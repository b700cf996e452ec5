"""The FP64 oracle against the reference's own output files (demo/)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

DEMO = os.path.join(GOLDEN, "demo_input.txt")


@pytest.fixture(scope="module")
def demo_rows(orc):
    rows = np.arange(1024)
    return orc.green_rows(DEMO, rows, compat=True)[:, :15]


def test_demo_spectrum_matches_reference(demo_rows):
    g = np.load(os.path.join(GOLDEN, "demo_spec.npy"))
    o = demo_rows
    rel = np.linalg.norm(o - g, axis=1) / np.linalg.norm(g, axis=1)
    # rows >= 3: agreement to the FP64 noise of the reference's formulation
    assert np.linalg.norm(o[3:] - g[3:]) / np.linalg.norm(g[3:]) < 1e-7
    assert rel[3:8].max() < 3e-6
    assert rel[8:64].max() < 5e-8
    assert rel[64:].max() < 1e-10


def test_demo_low_rows_are_fp64_noise(orc):
    """Rows 0-2 (|w| ~ omega_I = 4.9e-4 rad/s) are ill-conditioned in the
    reference's 4x4 formulation: two backward-stable FP64 solves of the same
    systems (direct LU vs explicit inverse then product, the reference's
    cuBLAS sequence) disagree by as much as either disagrees with the golden
    file, so those rows cannot be pinned by any FP64 implementation."""
    g = np.load(os.path.join(GOLDEN, "demo_spec.npy"))[:3]
    try:
        orc.set_inverse_mode(0)
        a = orc.green_rows(DEMO, [0, 1, 2])[:, :15]
        orc.set_inverse_mode(1)
        b = orc.green_rows(DEMO, [0, 1, 2])[:, :15]
    finally:
        orc.set_inverse_mode(0)
    noise = np.linalg.norm(a - b, axis=1) / np.linalg.norm(a, axis=1)
    dev = np.linalg.norm(a - g, axis=1) / np.linalg.norm(g, axis=1)
    assert noise[0] > 0.1  # row 0: O(1) disagreement between two FP64 solves
    assert np.all(dev < 5 * noise + 1e-9)


QUAD_ROWS = [0, 1, 2, 3, 4, 5]


@pytest.fixture(scope="module")
def demo_quad_rows(orc):
    """Demo rows 0-5 from the binary128 build of the oracle (same restatement,
    same FP64 grid): the truth the FP64 formulations are judged against."""
    return orc.green_rows(DEMO, QUAD_ROWS, compat=True, quad=True)[:, :15]


def _rel(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def test_quad_oracle_is_converged_and_on_the_fp64_grid(orc, demo_quad_rows):
    """The binary128 evaluation is its own check: the direct LU and the
    reference's inverse-then-product sequence agree to ~1e-19 in binary128,
    where in FP64 they differ by O(1) at row 0; both builds see the same
    ragged grid (the grid is formed in FP64 in both)."""
    q = demo_quad_rows
    try:
        orc.set_inverse_mode(1, quad=True)
        q1 = orc.green_rows(DEMO, [0, 1, 2], compat=True, quad=True)[:, :15]
    finally:
        orc.set_inverse_mode(0, quad=True)
    assert _rel(q1, q[:3]).max() < 1e-15
    st, run = orc.parse_input(DEMO)
    assert st == 0
    d64 = orc.ref_grid(run)
    d128 = orc.ref_grid(run, quad=True)
    assert d64[0] == d128[0] and d64[1] == d128[1] and np.array_equal(d64[2], d128[2])


def test_demo_low_rows_reference_is_fp64_noise(orc, demo_quad_rows):
    """Rows 0-2 of the reference's oname_w against the binary128 truth: the
    reference's row 0 is 41% off, and so are both FP64 solves of the same
    restatement (20%, 28%) -- the reference's rows 0-2 are the rounding noise
    of an FP64 4x4 formulation at |w| ~ w_I, and no FP64 implementation can
    reproduce them.  From row 3 on the reference agrees with the truth to
    1e-6 and better."""
    g = np.load(os.path.join(GOLDEN, "demo_spec.npy"))[QUAD_ROWS]
    q = demo_quad_rows
    try:
        orc.set_inverse_mode(0)
        a = orc.green_rows(DEMO, QUAD_ROWS)[:, :15]
        orc.set_inverse_mode(1)
        b = orc.green_rows(DEMO, QUAD_ROWS)[:, :15]
    finally:
        orc.set_inverse_mode(0)
    dev_ref, dev_lu, dev_inv = _rel(g, q), _rel(a, q), _rel(b, q)
    print("row err vs binary128: reference", dev_ref, "FP64 LU", dev_lu, "FP64 inverse", dev_inv)
    assert dev_ref[0] > 0.1 and dev_lu[0] > 0.05 and dev_inv[0] > 0.05
    # the reference's deviation is of the size of FP64 noise, row by row
    noise = np.maximum(dev_lu, dev_inv)
    assert np.all(dev_ref[:3] < 5 * noise[:3])
    assert dev_ref[3:].max() < 2e-6


def _trace_from_spec(spec):
    nw = spec.shape[0]
    N = 2 * nw
    full = np.zeros((N, spec.shape[1]), complex)
    full[:nw] = spec
    full[nw + 1:] = np.conj(spec[1:][::-1])
    dt, t0 = 2.0, 949.04
    omega_i = 2.0 / (N * dt)
    return np.fft.ifft(full, axis=0).real * np.exp(omega_i * (np.arange(N) * dt + t0))[:, None]


def test_demo_traces_against_reference_and_truth(demo_rows, demo_quad_rows):
    """Time domain.  'Truth' = the reference's oname_w with rows 0-2 replaced by
    the binary128 rows, transformed like green.cu:2222-2237.  The reference's
    own oname_t misses it by up to 2.6e-2 per column (all from rows 0-2);
    the FP64 oracle with the same substitution matches it to 1e-6."""
    g = np.load(os.path.join(GOLDEN, "demo_spec.npy"))
    t = np.load(os.path.join(GOLDEN, "demo_trace.npy"))
    truth_spec = g.copy()
    truth_spec[:3] = demo_quad_rows[:3]
    truth = _trace_from_spec(truth_spec)
    ref_vs_truth = _rel(t.T, truth.T)
    print("reference oname_t vs truth, per column:", ref_vs_truth)
    assert 1e-3 < ref_vs_truth.max() < 5e-2
    o = demo_rows.copy()
    o[:3] = demo_quad_rows[:3]
    assert _rel(_trace_from_spec(o).T, truth.T).max() < 1e-6


def test_demo_trace_is_closed_form_of_spectrum():
    """oname_t = Re(IFFT(oname_w)) / N * exp(omega_I (j dt + t0)) with the
    Hermitian extension -- pins the post/IFFT stage (green.cu:1925-1938,
    2222-2237) independently of the recursion."""
    g = np.load(os.path.join(GOLDEN, "demo_spec.npy"))
    t = np.load(os.path.join(GOLDEN, "demo_trace.npy"))
    nw = g.shape[0]
    N = 2 * nw
    full = np.zeros((N, 15), complex)
    full[:nw] = g
    full[nw + 1:] = np.conj(g[1:][::-1])
    dt, t0 = 2.0, 949.04
    omega_i = 2.0 / (N * dt)
    x = np.fft.ifft(full, axis=0).real * np.exp(omega_i * (np.arange(N) * dt + t0))[:, None]
    rel = np.linalg.norm(x - t, axis=0) / np.linalg.norm(t, axis=0)
    assert rel.max() < 1e-12


def test_oracle_post_ifft_reproduces_reference_trace(orc):
    """The oracle's post+IFFT on the golden spectrum (undoing the post step)
    gives the golden trace."""
    g = np.load(os.path.join(GOLDEN, "demo_spec.npy"))
    t = np.load(os.path.join(GOLDEN, "demo_trace.npy"))
    nw = g.shape[0]
    dt, t0, taper = 2.0, 949.04, 0.3
    N = 2 * nw
    df = 1.0 / (N * dt)
    omega_i = 2.0 / (N * dt)
    wc = int(nw * (1 - taper))
    iw = np.arange(nw)
    st0 = np.where(iw > wc, 0.5 * (1 + np.cos((iw - wc + 1) * np.pi / (nw - wc + 1))), 1.0)
    post = np.exp(1j * 2 * np.pi * df * iw * t0) * st0 * df * nw / np.pi
    for c in (0, 4, 10):
        raw = g[:, c] / post
        sp, tr = orc.post_ifft(raw, df, omega_i, dt, t0, taper)
        assert np.allclose(sp, g[:, c], rtol=1e-12, atol=1e-25)
        assert np.linalg.norm(tr - t[:, c]) / np.linalg.norm(t[:, c]) < 1e-12


def test_example_input_parses(orc):
    st, run = orc.parse_input(os.path.join(GOLDEN, "example_input.txt"))
    assert st == 0
    assert run.n_layer == 5 and run.s_layer == 0 and abs(run.zs - 8.0) < 1e-12
    dk, oi, nk, nkw = orc.ref_grid(run)
    assert nkw == nk.sum() and nk.min() >= 2


def test_c1_fp64_oracle_against_binary128(orc):
    """The FP64 oracle on the 3-layer workload against its binary128 build
    (the same restatement in __float128): the lowest rows, where the P and S
    columns of the basis grow parallel, within 2e-8 of each receiver's
    largest spectrum value; a middle band within 1e-12."""
    from grtm_cuda_amd import models
    cfg = models.config("c1")
    for (lo, hi), tol in (((0, 8), 2e-8), ((200, 204), 1e-12)):
        for kind in (0, 1):
            args = (cfg["layers"], cfg["src_depth"], cfg["rcv_depth"], cfg["dists"], kind, cfg["grid"])
            _, s64 = orc.traces(*args, w_range=(lo, hi))
            _, sq = orc.traces(*args, w_range=(lo, hi), quad=True)
            a, b = s64[..., lo:hi], sq[..., lo:hi]
            assert (np.abs(a - b).max(axis=-1) / np.abs(b).max(axis=-1)).max() < tol, (lo, kind)
