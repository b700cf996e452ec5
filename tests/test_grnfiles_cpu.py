"""CPU checks of the Green's-function file tools: binary SAC layout and round
trip, the grn mapping of the reference's driver and its misfit."""
import numpy as np

from grtm_cuda_amd import grnfiles


def test_sac_layout_and_round_trip(tmp_path):
    d = np.sin(np.arange(300) * 0.05).astype(np.float32)
    p = str(tmp_path / "x.sac")
    grnfiles.write_sac(p, d, 0.05, b=2.5, dist=30.0, evdp=10.0)
    raw = open(p, "rb").read()
    assert len(raw) == 632 + 4 * 300
    f = np.frombuffer(raw, "<f4", 70, 0)
    i = np.frombuffer(raw, "<i4", 40, 280)
    # the SAC header words: DELTA 0, B 5, E 6, DIST 50; NVHDR 76, NPTS 79, IFTYPE 85, LEVEN 105 (word numbers)
    assert f[0] == np.float32(0.05) and f[5] == 2.5 and np.isclose(f[6], 2.5 + 0.05 * 299) and f[50] == 30.0
    assert i[76 - 70] == 6 and i[79 - 70] == 300 and i[85 - 70] == 1 and i[105 - 70] == 1
    assert raw[440:448] == b"-12345  "  # KSTNM, first string field
    data, h = grnfiles.read_sac(p)
    assert np.array_equal(data, d) and h["npts"] == 300 and h["evdp"] == 10.0
    # big-endian files (written by other tools) read the same
    big = (f.astype(">f4").tobytes() + i.astype(">i4").tobytes() + raw[440:632] + d.astype(">f4").tobytes())
    q = tmp_path / "big.sac"
    q.write_bytes(big)
    data2, h2 = grnfiles.read_sac(str(q))
    assert np.array_equal(data2, d) and h2["dist"] == 30.0


def test_grn_mapping_and_files(tmp_path):
    cols = np.arange(6)[:, None] * np.ones((6, 8)) + 1.0  # column c holds c + 1
    g = grnfiles.grn_traces(cols)
    assert [g[i, 0] for i in range(6)] == [5.0, 2.0, 0.0, 4.0, 1.0, 3.0]
    assert np.array_equal(grnfiles.grn_traces(cols[:5]), g)  # five columns: no d/dz needed
    paths = grnfiles.write_grn(str(tmp_path), 30, cols, 0.1, t0=1.0, src_depth=12.0)
    assert [p.rsplit("/", 1)[1] for p in paths] == [f"30.grn.{i}" for i in range(6)]
    d, h = grnfiles.read_sac(paths[4])
    assert np.all(d == 1.0) and h["b"] == 1.0 and h["dist"] == 30.0


def test_reference_misfit():
    rng = np.random.default_rng(1)
    refs = [rng.standard_normal(50) for _ in range(6)]
    assert grnfiles.misfit(refs, refs) == 0.0
    assert grnfiles.misfit([-r for r in refs], refs) == 0.0  # it compares magnitudes
    # one pair with d = 0 against a non-zero reference contributes 1; the sum is divided by 5
    d = [r.copy() for r in refs]
    d[2][:] = 0.0
    assert np.isclose(grnfiles.misfit(d, refs), 1.0 / 5)
    refs[2][:] = 0.0  # all-zero references are skipped
    assert grnfiles.misfit(d, refs) == 0.0
    a, b = np.array([1.0, 2.0]), np.array([2.0, 2.0])
    assert np.isclose(grnfiles.misfit([a], [b], divisor=1), 1.0 / 7.0)
