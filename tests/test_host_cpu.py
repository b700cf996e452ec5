"""CPU-only checks of libgrtm.so: exports, error paths, and the
host-compiled copy of the device cell arithmetic against the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import ffi as _lib
from conftest import GOLDEN, ROOT


def test_library_exports_every_declared_symbol():
    lib = _lib._lib()
    hdr = open(os.path.join(ROOT, "include", "grtm.h")).read()
    code = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)  # declarations only, not comments
    declared = set(re.findall(r"\b(green|grtm_\w+)\s*\(", code))
    assert "green" in declared and "grtm_plan_execute" in declared
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_status_strings():
    lib = _lib._lib()
    assert b"gfx950" in lib.grtm_version()
    for s in range(-7, 1):
        assert lib.grtm_status_str(s)


def test_green_error_paths(tmp_path):
    with pytest.raises(g.GrtmError) as e:
        g.green(str(tmp_path / "missing.txt"), str(tmp_path / "w"), str(tmp_path / "t"))
    assert e.value.status == _lib.GRTM_E_IO
    bad = tmp_path / "bad.txt"
    bad.write_text("8.0\n0.0\n4\n0,6.1,3.5\n")
    with pytest.raises(g.GrtmError) as e:
        g.green(str(bad), str(tmp_path / "w"), str(tmp_path / "t"))
    assert e.value.status == _lib.GRTM_E_PARSE
    # the reference's green() has no status: the drop-in symbol records it
    lib = _lib._lib()
    lib.green(str(bad).encode(), b"/dev/null", b"/dev/null")
    assert lib.grtm_last_status() == _lib.GRTM_E_PARSE


def test_compat_rejects_source_on_interface(tmp_path):
    """Source exactly on a layer top: the reference inserts no layer but still
    counts one (its last layer stays uninitialised); compat mode refuses."""
    from grtm_cuda_amd import inputfile
    rows = np.array([[0, 6.1, 3.5, 2.7, 600, 300], [10, 6.6, 3.8, 2.9, 800, 400], [30, 8.0, 4.5, 3.3, 1200, 600]])
    p = tmp_path / "in.txt"
    inputfile.write_input(str(p), rows, 10.0, 0.0, 0.5, 50.0, 8, 0.0)
    with pytest.raises(g.GrtmError) as e:
        g.green_mem(str(p), compat=True)
    assert e.value.status == _lib.GRTM_E_UNSUPPORTED


def _random_case(rng, orc, compat):
    L = int(rng.integers(3, 9))
    w = 2 * np.pi * rng.uniform(0.01, 3.0) - 1j * rng.uniform(0.01, 0.2)
    vs = np.sort(rng.uniform(2.5, 4.5, L))
    vp = vs * rng.uniform(1.7, 1.8, L)
    rho = 1.74 * vp ** 0.25
    lay = np.zeros((L, 7))
    for j in range(L):
        a = orc.dispersion(w, vp[j], 100 * vs[j] ** 2)
        b = orc.dispersion(w, vs[j], 50 * vs[j] ** 2)
        mu = vs[j] ** 2 * rho[j] if compat else rho[j] * b * b
        lay[j] = [rng.uniform(1, 8), a.real, a.imag, b.real, b.imag, np.real(mu), np.imag(mu)]
    s = int(rng.integers(0, L - 1))
    # compat: the reference's receiver search ends at a layer <= s
    rl = int(rng.integers(0, s + 1 if compat else L))
    zs = rng.uniform(0.05, lay[s, 0] - 0.05)
    zr = rng.uniform(0.0, lay[rl, 0]) if rl < L - 1 else rng.uniform(0, 5)
    if rl == s and abs(zr - zs) < 1e-3:
        zr = 0.5 * zs
    k = rng.uniform(0, 3.0) * abs(w) / 2.5
    return L, w, lay, s, zs, rl, zr, k


@pytest.mark.parametrize("compat", [1, 0])
@pytest.mark.parametrize("src", [0, 1])
def test_cell_solver_matches_oracle(orc, compat, src):
    """Closed-form 2x2 block R/T + streaming sweeps (the device code, compiled
    for the host) against the oracle's 4x4 LU restatement, all four
    receiver/source branches."""
    rng = np.random.default_rng(100 + 10 * compat + src)
    seen = set()
    p0 = np.array([0.01, 0.02, 0.015, 0.03])
    for _ in range(80):
        L, w, lay, s, zs, rl, zr, k = _random_case(rng, orc, compat)
        a = orc.cell(compat, src, k, w, lay, s, zs, rl, zr, p0)
        b = g.selftest_cell(compat, src, k, w, lay, s, zs, rl, zr, p0)
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()
        seen.add(1 if rl < s else (2 if rl == s and zr <= zs else (3 if rl == s else 4)))
    assert seen == ({1, 2, 3} if compat else {1, 2, 3, 4})


def test_compat_refuses_a_receiver_below_the_source_layer(orc):
    """A receiver layer below the source layer cannot come out of the
    reference's search (green.cu:2092-2101), so compat mode refuses it on the
    device path and in the oracle alike; physics mode solves it."""
    rng = np.random.default_rng(5)
    L, w, lay, s, zs, _, _, k = _random_case(rng, orc, 0)
    while s >= L - 2:
        L, w, lay, s, zs, _, _, k = _random_case(rng, orc, 0)
    rl, zr = s + 1, 0.5 * lay[s + 1, 0]
    with pytest.raises(RuntimeError, match="status -4"):
        orc.cell(1, 0, k, w, lay, s, zs, rl, zr, np.zeros(4))
    with pytest.raises(g.GrtmError):
        g.selftest_cell(1, 0, k, w, lay, s, zs, rl, zr, np.zeros(4))
    a = orc.cell(0, 0, k, w, lay, s, zs, rl, zr, np.zeros(4))
    b = g.selftest_cell(0, 0, k, w, lay, s, zs, rl, zr, np.zeros(4))
    assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def test_cell_solver_depth_truncation(orc):
    """Thick layers and strongly evanescent wavenumbers, where the device
    template cuts its upward sweep below the deeper of source and receiver
    (or drops it): against the oracle's full recursion, receiver above the
    source layer and above / below the source in its layer, both modes and
    sources.  With the cut measured from the source instead of a deeper
    receiver, the branch-3 cases miss by about e^-18."""
    L, h, zs, s = 8, 6.0, 2.0, 2
    w = 2 * np.pi * 1.0 - 0.05j
    vs = np.linspace(3.0, 4.4, L)
    lay = np.zeros((L, 7))
    for compat in (1, 0):
        for j in range(L):
            a = orc.dispersion(w, 1.75 * vs[j], 100 * vs[j] ** 2)
            b = orc.dispersion(w, vs[j], 50 * vs[j] ** 2)
            rho = 1.74 * (1.75 * vs[j]) ** 0.25
            mu = vs[j] ** 2 * rho if compat else rho * b * b
            lay[j] = [h, a.real, a.imag, b.real, b.imag, np.real(mu), np.imag(mu)]
        fired = 0
        for rl, zr in ((0, 0.0), (s, 1.0), (s, 4.5)):
            z0 = max(zs, zr) if rl == s else zs
            for k in (3.0, 6.3, 12.0):
                nu = np.sqrt(k ** 2 - (w / (lay[s, 3] + 1j * lay[s, 4])) ** 2)
                fired += nu.real * (h - z0) + nu.real * h >= 18  # the sweep ends at layer s or s + 1
                for src in (0, 1):
                    a = orc.cell(compat, src, k, w, lay, s, zs, rl, zr, np.zeros(4))
                    b = g.selftest_cell(compat, src, k, w, lay, s, zs, rl, zr, np.zeros(4))
                    assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max(), (compat, rl, zr, k, src)
        assert fired >= 6


def test_cell_solver_free_surface_receiver(orc):
    """Receiver on the free surface in physics mode, where the device template
    forms the d/dz response from the traction-free conditions: against the
    oracle's general formula, both sources."""
    rng = np.random.default_rng(7)
    p0 = np.zeros(4)
    for _ in range(20):
        L, w, lay, s, zs, rl, zr, k = _random_case(rng, orc, 0)
        if s == 0:
            continue
        for src in (0, 1):
            a = orc.cell(0, src, k, w, lay, s, zs, 0, 0.0, p0)
            b = g.selftest_cell(0, src, k, w, lay, s, zs, 0, 0.0, p0)
            assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def _demo_cells(orc, row, compat):
    """Layer rows (h, alpha, beta, mu) of the demo model at its frequency row
    `row`, its geometry and wavenumbers (the reference's ragged grid)."""
    st, run = orc.parse_input(os.path.join(GOLDEN, "demo_input.txt"), compat=compat)
    assert st == 0
    dk, oi, nk, _ = orc.ref_grid(run, compat=compat)
    df = 1.0 / ((1 << run.log2n) * run.dt)
    w = complex(2 * np.pi * df * row, -oi)
    L = run.n_layer
    lay = np.zeros((L, 7))
    for j in range(L):
        a = orc.dispersion(w, run.a0[j], run.qp[j])
        b = orc.dispersion(w, run.b0[j], run.qs[j])
        mu = run.b0[j] ** 2 * run.rho[j] if compat else run.rho[j] * b * b
        lay[j] = [run.h[j], a.real, a.imag, b.real, b.imag, np.real(mu), np.imag(mu)]
    w0 = complex(0, -oi)
    a0 = orc.dispersion(w0, run.a0[run.r_layer], run.qp[run.r_layer])
    b0 = orc.dispersion(w0, run.b0[run.r_layer], run.qs[run.r_layer])
    r0, v0 = np.sqrt(-(w0 / a0) ** 2), np.sqrt(-(w0 / b0) ** 2)
    r0, v0 = (r0 if r0.real >= 0 else -r0), (v0 if v0.real >= 0 else -v0)
    p0 = np.array([r0.real, r0.imag, v0.real, v0.imag])
    return run, w, lay, p0, dk * np.arange(nk[row])


@pytest.mark.parametrize("compat", [1, 0])
def test_cell_solver_double_double_at_the_lowest_rows(orc, compat):
    """Demo rows 0-2 (|w| = 1, 3.3, 6.4 w_I): the device cell solve compiled in
    FP64 is rounding noise at row 0 (tens of % per cell, like the oracle's FP64
    4x4 solve); compiled in double-double (dd.h) it agrees with the binary128
    oracle to 1e-13, from the same FP64 layer constants."""
    worst64 = 0.0
    for row in (0, 1, 2):
        run, w, lay, p0, ks = _demo_cells(orc, row, compat)
        for src in (0, 1):
            for k in ks[::211]:
                q = orc.cell(compat, src, k, w, lay, run.s_layer, run.zs, run.r_layer, run.zr, p0, quad=True)
                d = g.selftest_cell(compat, src, k, w, lay, run.s_layer, run.zs, run.r_layer, run.zr, p0, dd=True)
                f = g.selftest_cell(compat, src, k, w, lay, run.s_layer, run.zs, run.r_layer, run.zr, p0)
                m = np.abs(q).max()
                assert np.abs(d - q).max() <= 1e-13 * m, (row, src, k)
                if row == 0:
                    worst64 = max(worst64, np.abs(f - q).max() / m)
    assert worst64 > 1e-2  # the FP64 solve at row 0: noise


def test_noalloc_rejects_a_run_beyond_its_scratch(tmp_path):
    """A run whose device memory exceeds the no-allocation entry point's fixed
    32 MiB scratch is refused with GRTM_E_TOO_LARGE before any device work
    (so also without a GPU): 8192 frequencies of the demo model need about
    190 MB of k-sum partials."""
    from grtm_cuda_amd import inputfile
    rows = np.array([[0, 6.1, 3.5, 2.7, 600, 300], [10, 6.6, 3.8, 2.9, 800, 400], [30, 8.0, 4.5, 3.3, 1200, 600]])
    p = tmp_path / "big.txt"
    inputfile.write_input(str(p), rows, 5.0, 0.0, 2.0, 50.0, 14, 0.0)
    lib = _lib._lib()
    nw = ctypes.c_int(0)
    spec = np.zeros((30, 8192, 2))
    st = lib.grtm_green_mem_noalloc(str(p).encode(), 1, spec.ctypes.data_as(ctypes.c_void_p), None,
                                    ctypes.byref(nw), None)
    assert st == _lib.GRTM_E_TOO_LARGE and nw.value == 8192
    assert b"scratch" in lib.grtm_status_str(st)
