"""Pins of the oracle's compat-only paths (the reference's quirks), which no
reference output covers: the ragged grid's cal_kmax branches against a table
worked out by hand from the geometry, the receiver search, and compat against
physics mode wherever a quirk is inactive or its effect has a closed form.

Coefficient order of oracle.cell: a0[0], a0[1], a1[0], a1[1], a2[0], a2[1],
b1, b2, then the same with d/dz (a0t ... b2t)."""
import math

import numpy as np
import pytest

from grtm_cuda_amd import inputfile, models

# Input layers (top km, vp, vs, rho, qp, qs).  Q = 1e12 makes the dispersed
# velocities equal to the reference ones to 1e-12 (so |w / b| = |w| / vs), and
# the reference's real mu (Q3) equal to rho b(w)^2.
Q_INF = 1e12
VS = [1.8, 2.6, 2.0, 3.6, 4.5]
TOPS = [0.0, 2.0, 5.0, 9.0, 14.0]
MODEL = np.array([[t, 1.75 * v, v, 2.5, Q_INF, Q_INF] for t, v in zip(TOPS, VS)])

# (mode, source depth, receiver depth) -> the stack after the source split, the
# located (s, zs, r, zr) and, by cal_kmax's branch (green.cu:112-147), the
# slowest shear velocity entering pmin and dzmax.  Worked out by hand:
#  - physics mode locates by layer tops; compat compares the source's depth
#    inside its layer with each thickness (Q1, green.cu:2092-2101);
#  - rl < s - 1 takes b[s], b[s - 1] and b[lay + 1] for lay = rl .. s - 2 (so
#    never layer rl) and the thicknesses h[rl .. s - 2];
#  - rl > s + 1 takes b[s], b[s + 1] and b[lay] for lay = s + 1 .. rl - 2 (so
#    never layer rl - 1) and the thicknesses h[s + 1 .. rl - 2];
#  - the source sits at the bottom of layer s (the split), so h[s] - zs = 0
#    and every rl > s case is near field.
CASES = [
    # name, compat, hs, hr, (s, zs, r, zr), v_slow, dzmax
    ("rl == s", 0, 3.5, 2.5, (1, 1.5, 1, 0.5), 2.6, 1.0),
    ("rl == s, near field", 0, 3.5, 3.4, (1, 1.5, 1, 1.4), 2.6, 0.1),
    ("rl == s + 1", 0, 3.5, 4.0, (1, 1.5, 2, 0.5), 2.6, 0.0),
    # layers s .. rl: 2.6, 2.6, 2.0, 3.6; the 2.0 layer is rl - 1
    ("rl > s + 1", 0, 3.5, 12.0, (1, 1.5, 4, 3.0), 2.6, 0.0),
    # dzmax = min(zs = 2.5, h[rl] - zr = 3 - 2.2)
    ("rl == s - 1", 0, 7.5, 4.2, (2, 2.5, 1, 2.2), 2.0, 0.8),
    # dzmax = min(zs = 1, hs - hr, h[0] = 2); layer 0 (1.8 km/s) is rl
    ("rl < s - 1", 0, 6.0, 0.5, (2, 1.0, 0, 0.5), 2.0, 1.0),
    # dzmax = min(zs = 3, hs - hr, h[0] = 2, h[1] = 3)
    ("rl < s - 1, two layers between", 0, 12.0, 0.5, (3, 3.0, 0, 0.5), 2.0, 2.0),
    ("compat rl == s - 1", 1, 7.5, 4.2, (2, 2.5, 1, 2.2), 2.0, 0.8),
    ("compat rl < s - 1", 1, 6.0, 0.5, (2, 1.0, 0, 0.5), 2.0, 1.0),
    # Q1: the source's residual depth 1.5 fits in layer 0, so the receiver is
    # placed in layer 0 at 2.5 km, below that layer's 2 km bottom:
    # dzmax = min(zs, h[0] - zr) = -0.5 (near field), pmin over layers 1 and 0
    ("compat receiver past its layer", 1, 3.5, 2.5, (1, 1.5, 0, 2.5), 1.8, -0.5),
    ("compat rl == s", 1, 1.5, 0.0, (0, 1.5, 0, 0.0), 1.8, 1.5),
]


def _input(tmp_path, hs, hr, name="in.txt"):
    p = tmp_path / name
    inputfile.write_input(str(p), MODEL, hs, hr, 0.1, 50.0, 8, 0.0)
    return str(p)


def _hand_nk(n, v, dz, df, oi, dk):
    """nk of row n from cal_kmax's final steps (green.cu:149-160, 2155-2163)."""
    w = complex(2 * math.pi * df * max(n, 3), -oi)  # rows below 3 are sized at 3 df
    p = abs(w) / v
    if dz < 0.2:
        km = min(p + 2.0, (20.0 if n < 10 else 2.0) * p)
    else:
        km = math.sqrt((3.0 / dz) ** 2 + p * p)
    floor = km - p < 0.5
    if floor:
        km = p + 0.5
    return int(1.1 * km / dk) + 2, floor


def test_cal_kmax_branch_table(orc, tmp_path):
    hit = set()
    for name, compat, hs, hr, (s, zs, r, zr), v, dz in CASES:
        st, run = orc.parse_input(_input(tmp_path, hs, hr), compat=bool(compat))
        assert st == 0, name
        assert (run.s_layer, run.r_layer) == (s, r), name
        assert abs(run.zs - zs) < 1e-12 and abs(run.zr - zr) < 1e-12, name
        dk, oi, nk, _ = orc.ref_grid(run, compat=bool(compat))
        N = 1 << run.log2n
        df = 1.0 / (N * run.dt)
        assert abs(oi - 2.0 / (N * run.dt)) < 1e-15  # Q13: omega_I = 2/T
        for n in range(len(nk)):
            want, floor = _hand_nk(n, v, dz, df, oi, dk)
            assert nk[n] == want, (name, n, nk[n], want)
            hit.add("near" if dz < 0.2 else "far")
            hit.add(("near<10" if n < 10 else "near>=10") if dz < 0.2 else "sqrt")
            if floor:
                hit.add("floor")
    assert hit == {"near", "far", "near<10", "near>=10", "sqrt", "floor"}


def test_compat_receiver_search_stays_at_or_above_the_source_layer(orc, tmp_path):
    """The reference's receiver search compares the source's depth inside its
    layer (at most h[s]) with each thickness, so it can only stop at a layer
    <= s; physics mode places deeper receivers below.  Where the source's
    residual depth fits in the top layer and the receiver is in it, both
    searches give the same receiver."""
    rng = np.random.default_rng(3)
    deeper = 0
    for i in range(200):
        hs, hr = rng.uniform(0.1, 13.9), rng.uniform(0.0, 20.0)
        if any(abs(hs - t) < 1e-3 for t in TOPS):
            continue
        st, run = orc.parse_input(_input(tmp_path, hs, hr, f"r{i}.txt"), compat=True)
        assert st == 0
        assert run.r_layer <= run.s_layer
        zsum = sum(run.h[j] for j in range(run.r_layer))
        assert abs(run.zr - (hr - zsum)) < 1e-12
        st, phys = orc.parse_input(_input(tmp_path, hs, hr, f"r{i}.txt"), compat=False)
        deeper += phys.r_layer > phys.s_layer
        if run.zs <= run.h[0] and hr <= run.h[0]:
            assert (phys.r_layer, phys.zr) == (run.r_layer, run.zr)
    assert deeper > 20


def _cells(rng, n):
    """Random cells with every compat receiver branch (rl <= s)."""
    out = []
    for _ in range(n):
        L = int(rng.integers(3, 8))
        w = 2 * np.pi * rng.uniform(0.02, 2.0) - 1j * rng.uniform(0.02, 0.2)
        vs = np.sort(rng.uniform(2.5, 4.5, L))
        lay = np.zeros((L, 7))
        for j in range(L):
            from oracle import oracle
            a = oracle.dispersion(w, 1.75 * vs[j], 100 * vs[j] ** 2)
            b = oracle.dispersion(w, vs[j], 50 * vs[j] ** 2)
            mu = 2.6 * b * b
            lay[j] = [rng.uniform(1, 6), a.real, a.imag, b.real, b.imag, mu.real, mu.imag]
        s = int(rng.integers(0, L - 1))
        rl = int(rng.integers(0, s + 1))
        zs = rng.uniform(0.1, lay[s, 0] - 0.1)
        zr = rng.uniform(0.0, lay[rl, 0])
        if rl == s and abs(zr - zs) < 0.05:
            zr = 0.5 * zs
        k = rng.uniform(0.0, 2.5) * abs(w) / vs[0]
        out.append((w, lay, s, zs, rl, zr, k))
    return out


def _nu(k, w, v):
    z = np.sqrt(complex(k * k) - (w / v) ** 2)
    return z if z.real >= 0 else -z


def _close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def test_compat_displacement_coefficients_equal_physics(orc):
    """On the same layer constants (Q3 is the caller's mu) and geometry, the
    quirks Q4/Q5 and the r < s d/dz product touch only d/dz coefficients:
    every displacement coefficient of the single force, and the P-SV ones of
    the moment tensor, are physics mode's."""
    rng = np.random.default_rng(11)
    for w, lay, s, zs, rl, zr, k in _cells(rng, 60):
        p0 = np.array([0.3, -0.1, 0.2, 0.05])
        for src, idx in ((0, range(0, 7)), (1, range(0, 6))):
            c = orc.cell(1, src, k, w, lay, s, zs, rl, zr, p0)
            p = orc.cell(0, src, k, w, lay, s, zs, rl, zr, p0)
            m = np.abs(p[:8]).max()
            for i in idx:
                assert abs(c[i] - p[i]) <= 1e-13 * m, (src, i, s, rl)


def test_compat_tensor_sh_factors_in_closed_form(orc):
    """Q17 (green.cu:1408-1411): the reference gives both SH orders the factor
    1 + R above the source and R - 1 below.  So above the source compat's n = 1
    coefficient is the n = 2 one times v_s / k (physics: 1 - R for n = 1), and
    below compat's n = 2 is the n = 1 one times k / v_s (physics: 1 + R for
    n = 2); the other order is physics mode's, which reciprocity pins."""
    rng = np.random.default_rng(12)
    up = down = differs = 0
    for w, lay, s, zs, rl, zr, k in _cells(rng, 80):
        if k < 1e-3:
            continue
        c = orc.cell(1, 1, k, w, lay, s, zs, rl, zr, np.zeros(4))
        p = orc.cell(0, 1, k, w, lay, s, zs, rl, zr, np.zeros(4))
        vs = _nu(k, w, lay[s, 3] + 1j * lay[s, 4])
        if rl < s or zr < zs:
            up += 1
            assert _close(c[7], p[7]) and _close(c[6], p[7] * vs / k, 1e-11)
            differs += not _close(c[6], p[6], 1e-3)
        else:
            down += 1
            assert _close(c[6], p[6]) and _close(c[7], p[6] * k / vs, 1e-11)
            differs += not _close(c[7], p[7], 1e-3)
    # the quirk acts wherever the SH reflection at the source is not negligible
    assert up > 20 and down > 5 and differs > 20


def test_compat_single_force_dz_coefficients_in_closed_form(orc):
    """The single force's d/dz coefficients in compat (columns 5-9):
    - receiver above the source in its layer (Q5, green.cu:720): the upgoing SH
      term carries an extra factor v, so b1t = b1t_phys + (v - 1)(v b1 + b1t_phys)/2
      (b1 = X(P + Q), b1t_phys = X v(Q - P) with Q the direct upgoing part);
    - receiver above the source layer: with the first cell's wavenumbers (Q4,
      green.cu:674-675) set to the cell's own, the P-SV d/dz coefficients are
      physics mode's, and the SH one is the physics one times the receiver
      factor u0 = A + B of green.cu:842, which b1 and b1t_phys fix:
      b1 = Y(A + B), b1t_phys = Y v(B - A), B = e^{-v (h_r - z_r)}.
    - receiver below the source in its layer: no quirk, physics mode's."""
    rng = np.random.default_rng(13)
    seen = set()
    for w, lay, s, zs, rl, zr, k in _cells(rng, 90):
        ar, br = lay[rl, 1] + 1j * lay[rl, 2], lay[rl, 3] + 1j * lay[rl, 4]
        rr, vr = _nu(k, w, ar), _nu(k, w, br)
        p0 = np.array([rr.real, rr.imag, vr.real, vr.imag])
        c = orc.cell(1, 0, k, w, lay, s, zs, rl, zr, p0)
        p = orc.cell(0, 0, k, w, lay, s, zs, rl, zr, p0)
        m = np.abs(p[8:12]).max()
        if rl == s and zr < zs:
            seen.add("above in layer")
            want = p[14] + (vr - 1) * (vr * p[6] + p[14]) / 2
            assert abs(c[14] - want) <= 1e-10 * max(abs(want), abs(p[14])), (s, rl, k)
        elif rl < s:
            seen.add("above layer")
            B = np.exp(-vr * (lay[rl, 0] - zr))
            u0 = p[6] * 2 * B / (p[6] + p[14] / vr)
            assert abs(c[14] - p[14] * u0) <= 1e-9 * abs(c[14]), (s, rl, k)
        else:
            seen.add("below in layer")
            assert abs(c[14] - p[14]) <= 1e-13 * abs(p[14])
        for i in range(8, 12):
            assert abs(c[i] - p[i]) <= 1e-12 * m, (i, s, rl)
    assert seen == {"above in layer", "above layer", "below in layer"}


@pytest.mark.parametrize("hs,hr", [(6.0, 0.5), (1.5, 0.0)], ids=["receiver above the source layer", "same layer"])
def test_compat_kernels_equal_physics_at_infinite_q(orc, hs, hr):
    """Where the two receiver searches agree and Q is infinite (so the
    reference's real mu, Q3, is exact): the vertical force's Z and R kernels
    and the explosion's Z kernel are physics mode's, and the explosion's R
    kernel is minus its Z kernel (Q9, green.cu:1653-1657: the n = 0 R columns
    take the vertical coefficients)."""
    layers = MODEL.copy()
    layers[:-1, 0] = np.diff(MODEL[:, 0])  # thicknesses
    layers[-1, 0] = 0.0
    grid = models.dense_grid(32, 0.1, 256, min(VS))
    for kind in (0, 1):
        c = orc.kernels(layers, hs, hr, kind, grid, 0, 32, compat=True)
        p = orc.kernels(layers, hs, hr, kind, grid, 0, 32, compat=False)
        m = np.abs(p).max()
        assert np.abs(c[..., 0] - p[..., 0]).max() <= 1e-9 * m
        if kind == 1:
            assert np.abs(c[..., 1] - p[..., 1]).max() <= 1e-9 * m
        else:
            assert np.abs(c[..., 1] + c[..., 0]).max() <= 1e-15 * m
            assert np.abs(c[..., 1] - p[..., 1]).max() > 1e-3 * m


def test_compat_traces_equal_physics_at_infinite_q(orc):
    """End to end where no quirk acts (infinite Q, the searches agree, the
    vertical force and the explosion's Z): compat and physics traces differ
    only by compat's A&S Bessel polynomials (absolute error <= 5e-8)."""
    layers = MODEL.copy()
    layers[:-1, 0] = np.diff(MODEL[:, 0])
    layers[-1, 0] = 0.0
    grid = models.dense_grid(64, 0.1, 1024, min(VS))
    dists = [10.0, 25.0]
    for kind, comps in ((1, (0, 1)), (0, (0,))):
        c = orc.traces(layers, 6.0, 0.5, dists, kind, grid, compat=True)[0]
        p = orc.traces(layers, 6.0, 0.5, dists, kind, grid, compat=False)[0]
        for ci in comps:
            a, b = c[:, ci], p[:, ci]
            err = np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)
            assert err.max() < 2e-6, (kind, ci, err)
