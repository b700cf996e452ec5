"""Peak-trough averaging (PTAM) of the k-sum tail in the oracle, pinned by a
closed-form integral and by convergence to the long dense grid.  The
reference reads n_ptam (green.cu:2059) and leaves the method as a sketch
(integral.cu:38-88) that never runs, so it has no output to compare with."""
import numpy as np
import scipy.special as sp

from grtm_cuda_amd import models


def test_ptam_lipschitz_integral(orc):
    """int_0^inf e^{-a k} J0(k r) dk = 1/sqrt(a^2 + r^2) (Lipschitz), complex a
    with a slow decay: midpoint sums truncated at k = 20 miss by ~1e-2, the
    PTAM continuation of the same sums recovers the integral to the
    quadrature's own error (which a finer step lowers by h^2)."""
    a, r = 0.05 + 0.05j, 10.0
    exact = 1 / np.sqrt(a * a + r * r)
    errs = []
    for h in (0.004, 0.002):
        n0 = int(round(20.0 / h))
        k = (np.arange(n0) + 0.5) * h
        s0 = np.sum(np.exp(-a * k) * sp.j0(k * r)) * h
        kt = (np.arange(n0, n0 + int(round(6.0 / h))) + 0.5) * h
        terms = np.exp(-a * kt) * sp.j0(kt * r) * h
        est, found = orc.ptam(s0, terms, 10)
        assert found == 10
        assert abs(s0 - exact) > 1e-2 * abs(exact)
        errs.append(abs(est - exact) / abs(exact))
    assert errs[1] < 2e-7 and errs[1] < 0.35 * errs[0], errs


def test_ptam_short_tail_is_the_extended_sum(orc):
    """Fewer extrema than asked for in the terms given: the plain extended sum."""
    s0 = 1.0 + 2.0j
    terms = np.array([0.1 + 0.1j, 0.05 - 0.02j, -0.01 + 0.03j])
    est, found = orc.ptam(s0, terms, 10)
    assert found < 10 and abs(est - (s0 + terms.sum())) < 1e-15


def test_ptam_traces_converge_to_the_long_grid(orc):
    """A source 0.5 km below surface receivers: the 3-layer model's kernels
    decay only as e^{-0.5 k}, so the dense grid's k_max leaves 1e-4..1e-2 of
    each trace out.  With 10-extremum PTAM over the following wavenumbers
    (same dk) the traces come within 1e-5 of the dense sum on a grid 32 times
    longer, per receiver up to 30 km; at 60 km (6.6 grid points per J0
    period) the gain is still tenfold."""
    layers = models.C1_LAYERS
    g1 = models.dense_grid(64, 0.1, 512, float(layers[:, 2].min()))
    dk = g1["kmax"] / g1["nk"]
    g2 = dict(g1, nk=512 * 32, kmax=dk * 512 * 32)
    dists = np.array([5.0, 10.0, 30.0, 60.0])
    nt = int(np.ceil((10 / 2 + 2) * 2 * np.pi / (dists.min() * dk)))
    for kind in (0, 1):
        long_ = orc.traces(layers, 0.5, 0.0, dists, kind, g2)[0]
        plain = orc.traces(layers, 0.5, 0.0, dists, kind, g1)[0]
        pt = orc.traces(layers, 0.5, 0.0, dists, kind, g1, ptam=(10, nt))[0]
        e_plain = np.linalg.norm(plain - long_, axis=-1) / np.linalg.norm(long_, axis=-1)
        e_pt = np.linalg.norm(pt - long_, axis=-1) / np.linalg.norm(long_, axis=-1)
        print(kind, e_plain.max(axis=1), e_pt.max(axis=1))
        assert (e_plain[:3] > 1e-4).any() and e_pt[:3].max() < 1e-5
        assert e_pt[3].max() < 0.1 * e_plain[3].max()
