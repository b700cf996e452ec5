"""GPU parity of the drop-in path (green) through the C ABI."""
import os

import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import inputfile
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
DEMO = os.path.join(GOLDEN, "demo_input.txt")


@pytest.fixture(scope="module")
def demo_gpu():
    print("gpu: green_mem(demo, compat)", flush=True)
    return g.green_mem(DEMO, compat=True)


def test_demo_matches_reference_output(demo_gpu):
    spec, tr = demo_gpu
    gold = np.load(os.path.join(GOLDEN, "demo_spec.npy"))
    s = spec[:15].T
    rel = np.linalg.norm(s - gold, axis=1) / np.linalg.norm(gold, axis=1)
    assert np.linalg.norm(s[3:] - gold[3:]) / np.linalg.norm(gold[3:]) < 1e-7
    assert rel[3:8].max() < 3e-6
    assert rel[8:64].max() < 5e-8
    assert rel[64:].max() < 1e-10
    assert np.all(spec[15:] == 0) and np.all(tr[15:] == 0)


def _trace_from_spec(spec):
    """oname_t from oname_w rows 0..Nw-1 (green.cu:1925-1938, 2222-2237)."""
    nw = spec.shape[0]
    N = 2 * nw
    full = np.zeros((N, spec.shape[1]), complex)
    full[:nw] = spec
    full[nw + 1:] = np.conj(spec[1:][::-1])
    dt, t0 = 2.0, 949.04
    omega_i = 2.0 / (N * dt)
    return np.fft.ifft(full, axis=0).real * np.exp(omega_i * (np.arange(N) * dt + t0))[:, None]


def test_demo_low_rows_match_binary128(demo_gpu, orc):
    """Rows 0-5 run in double-double on the GPU (the FP64 solve is rounding
    noise at |w| ~ w_I): against the binary128 oracle they agree to 1e-10,
    where the reference's own rows 0-2 are off by 41%, 5e-4 and 7e-6."""
    spec, _ = demo_gpu
    rows = [0, 1, 2, 3, 4, 5]
    q = orc.green_rows(DEMO, rows, compat=True, quad=True)[:, :15]
    s = spec[:15, rows].T
    rel = np.linalg.norm(s - q, axis=1) / np.linalg.norm(q, axis=1)
    print("GPU rows 0-5 vs binary128:", rel)
    assert rel.max() < 1e-10


def test_demo_traces_match_reference(demo_gpu, orc):
    """The drop-in's oname_t against the reference's (demo/oname_t.txt), per
    column.  The reference's rows 0-2 are FP64 noise, so the comparison is
    made twice: against its traces as shipped, where the GPU is exactly as
    far from them as the binary128 truth is (up to 2.2e-2 per column, all
    from the reference's rows 0-2), and against the truth (the reference's
    spectrum with rows 0-7 from binary128), where every column agrees to
    1e-6 rel-L2 and peak."""
    _, tr = demo_gpu
    gold_t = np.load(os.path.join(GOLDEN, "demo_trace.npy"))
    gold_w = np.load(os.path.join(GOLDEN, "demo_spec.npy"))
    truth_w = gold_w.copy()
    # the reference's rows 0-7 carry FP64 noise above 1e-8 (41% at row 0)
    truth_w[:8] = orc.green_rows(DEMO, list(range(8)), compat=True, quad=True)[:, :15]
    truth = _trace_from_spec(truth_w).T
    t = tr[:15]
    g = gold_t.T

    def rel(a, b):
        return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)

    def peak(a, b):
        return np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)

    vs_ref, vs_truth, ref_vs_truth = rel(t, g), rel(t, truth), rel(g, truth)
    print("per column rel-L2: GPU vs reference", vs_ref, "GPU vs truth", vs_truth, "reference vs truth", ref_vs_truth)
    print("per column peak: GPU vs reference", peak(t, g), "GPU vs truth", peak(t, truth))
    assert vs_truth.max() < 1e-6 and peak(t, truth).max() < 1e-6
    assert np.all(np.abs(vs_ref - ref_vs_truth) <= 1e-6 + 1e-3 * ref_vs_truth)


def test_demo_matches_oracle(demo_gpu, orc):
    spec, _ = demo_gpu
    rows = np.arange(3, 1024, 7)
    o = orc.green_rows(DEMO, rows, compat=True)
    s = spec[:, rows].T
    rel = np.linalg.norm(s - o, axis=1) / np.linalg.norm(o, axis=1)
    assert rel.max() < 2e-6 and np.median(rel) < 1e-11


def test_demo_trace_is_ifft_of_spectrum(demo_gpu):
    spec, tr = demo_gpu
    nw = spec.shape[1]
    N = 2 * nw
    full = np.zeros((15, N), complex)
    full[:, :nw] = spec[:15]
    full[:, 0] = full[:, 0].real
    full[:, nw + 1:] = np.conj(spec[:15, 1:][:, ::-1])
    dt, t0 = 2.0, 949.04
    oi = 2.0 / (N * dt)
    x = np.fft.ifft(full, axis=1).real * np.exp(oi * (np.arange(N) * dt + t0))[None, :]
    assert np.linalg.norm(x - tr[:15]) / np.linalg.norm(tr[:15]) < 1e-12


def test_green_writes_reference_format(tmp_path, demo_gpu):
    ow, ot = str(tmp_path / "w.txt"), str(tmp_path / "t.txt")
    g.green(DEMO, ow, ot)
    w = inputfile.read_output(ow)
    t = inputfile.read_output(ot)
    spec, tr = demo_gpu
    nw = spec.shape[1]
    assert w.shape == (2 * nw, 60) and t.shape == (2 * nw, 30)
    assert np.allclose(w[:nw, 0::2] + 1j * w[:nw, 1::2], spec.T, rtol=1e-15, atol=0)
    assert np.all(w[nw] == 0)
    assert np.allclose(w[nw + 1:, 0] + 1j * w[nw + 1:, 1], np.conj(spec[0, 1:][::-1]))
    assert np.allclose(t, tr.T, rtol=1e-15, atol=0)


THIN_TOP = np.array([[0, 5.0, 2.9, 2.5, 400, 200], [2, 6.1, 3.5, 2.75, 600, 300], [17, 6.6, 3.8, 2.9, 800, 400],
                     [35, 8.0, 4.5, 3.3, 1200, 600]])
THICK_TOP = np.array([[0, 5.8, 3.3, 2.6, 500, 250], [20, 6.4, 3.7, 2.8, 700, 350], [30, 6.9, 3.9, 3.0, 900, 450],
                      [45, 8.1, 4.6, 3.3, 1200, 600]])


def _write(tmp_path, name, hs, hr, dist, src_type, log2n=8, dt=0.5, rows=THIN_TOP):
    p = str(tmp_path / name)
    inputfile.write_input(p, rows, hs, hr, dt, dist, log2n, 2.0, src_type=src_type)
    return p


@pytest.mark.parametrize("compat", [True, False])
@pytest.mark.parametrize("src_type", [0, 1])
@pytest.mark.parametrize("geom", ["r_above", "r_above_thick", "same_above", "same_below", "r_below"])
def test_branches_match_oracle(tmp_path, orc, compat, src_type, geom):
    """Every receiver/source branch, both source types, both modes."""
    hs, hr = {"r_above": (10.0, 0.0), "r_above_thick": (25.0, 0.0), "same_above": (10.0, 4.0),
              "same_below": (5.0, 8.0), "r_below": (5.0, 25.0)}[geom]
    rows = THICK_TOP if geom == "r_above_thick" else THIN_TOP
    p = _write(tmp_path, "in.txt", hs, hr, 30.0, src_type, rows=rows)
    print(f"gpu: branch {geom} src {src_type} compat {compat}", flush=True)
    spec, tr = g.green_mem(p, compat=compat)
    rows = np.arange(2, spec.shape[1], 5)
    o = orc.green_rows(p, rows, compat=compat)
    # FP64 noise floor of the formulation: the oracle's two solves of the
    # interface systems (direct LU vs inverse-then-product)
    try:
        orc.set_inverse_mode(1)
        o2 = orc.green_rows(p, rows, compat=compat)
    finally:
        orc.set_inverse_mode(0)
    s = spec[:, rows].T
    live = np.abs(o).max(axis=0) > 0
    rel = np.linalg.norm(s[:, live] - o[:, live]) / np.linalg.norm(o[:, live])
    noise = np.linalg.norm(o2[:, live] - o[:, live]) / np.linalg.norm(o[:, live])
    # compat "r_below" puts the receiver 20 km below the bottom of its layer
    # (the reference's receiver search keys on the source depth): e^{+20 nu}
    # growth makes that output rounding noise in any FP64 implementation
    assert rel < max(1e-9, 5 * noise), (rel, noise)


def test_c_caller_writes_the_reference_files(tmp_path, demo_gpu):
    """The C caller (tests/c_caller/green_main, built by build() against
    include/grtm.h) run as the reference's ./test is: green("input.txt",
    "out4.txt", "signal.txt") with paths relative to its working directory."""
    import shutil
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "c_caller", "green_main")
    shutil.copy(DEMO, tmp_path / "input.txt")
    print("gpu: C caller green()", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    w = inputfile.read_output(str(tmp_path / "out4.txt"))
    t = inputfile.read_output(str(tmp_path / "signal.txt"))
    spec, tr = demo_gpu
    nw = spec.shape[1]
    assert np.allclose(w[:nw, 0::2] + 1j * w[:nw, 1::2], spec.T, rtol=1e-15, atol=0)
    assert np.allclose(t, tr.T, rtol=1e-15, atol=0)
