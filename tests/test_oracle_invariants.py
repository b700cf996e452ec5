"""Pins of the oracle beyond the golden files: library Bessel values,
closed-form physics (whole-space explosion, Rayleigh pole) and invariances."""
import numpy as np
import pytest
import scipy.special as sp

from grtm_cuda_amd import models


def test_bessel_physics_mode_matches_scipy(orc):
    xs = np.concatenate([np.linspace(0, 10, 101), np.geomspace(10, 2e4, 80)])
    for n in (0, 1, 2):
        for x in xs:
            assert abs(orc.bessel(1, 0, n, x) - sp.jv(n, x)) < 1e-14 + 1e-12 * abs(sp.jv(n, x))
            assert abs(orc.bessel(1, 1, n, x) - sp.jvp(n, x, 1)) < 1e-13
            assert abs(orc.bessel(1, 2, n, x) - sp.jvp(n, x, 2)) < 1e-13


def test_bessel_reference_mode_within_as_error(orc):
    """A&S 9.4.1-9.4.6 polynomials: absolute error below 5e-8 (their published
    bound is 5e-8 for J0/J1 on x <= 3 and 1.6e-8 / 1.1e-8 in amplitude/phase above)."""
    xs = np.concatenate([np.linspace(0, 10, 201), np.geomspace(10, 2e4, 100)])
    err0 = max(abs(orc.bessel(0, 0, 0, x) - sp.j0(x)) for x in xs)
    err1 = max(abs(orc.bessel(0, 0, 1, x) - sp.j1(x)) for x in xs)
    err2 = max(abs(orc.bessel(0, 0, 2, x) - sp.jv(2, x)) for x in xs)
    errd = max(abs(orc.bessel(0, 1, 1, x) - sp.jvp(1, x)) for x in xs[1:])
    # J2 below x = 2 comes from Miller's recurrence with iacc = 40 (single-precision
    # accuracy by design), J1' = J1/x - J2 inherits it
    assert err0 < 5e-8 and err1 < 5e-8 and err2 < 3e-7 and errd < 3e-7
    # the reference's "J''" (green.cu:1552) is not the second derivative
    # (it differs by 2 J0/x - 4 J1/x^2 for n = 1)
    x = 1.0
    assert abs(orc.bessel(0, 2, 1, x) - sp.jvp(1, x, 2)) > 0.2


def _lay(w, rows, compat=False):
    out = np.zeros((len(rows), 7))
    for j, (h, vp, vs, rho, qp, qs) in enumerate(rows):
        a = orc_disp(w, vp, qp)
        b = orc_disp(w, vs, qs)
        mu = vs * vs * rho if compat else rho * b * b
        out[j] = [h, a.real, a.imag, b.real, b.imag, np.real(mu), np.imag(mu)]
    return out


def orc_disp(w, v, q):
    from oracle import oracle
    return oracle.dispersion(w, v, q)


def test_identical_layers_are_transparent(orc):
    w = 2 * np.pi * 0.7 - 0.05j
    a, b, mu = orc_disp(w, 6.1, 600), orc_disp(w, 3.5, 300), 2.75 * 3.5 ** 2
    for k in (0.0, 0.3, 1.2, 5.0):
        rt = orc.interface_rt(k, w, (a, b, mu), (a, b, mu))
        assert np.allclose(rt["rud"], 0, atol=1e-12) and np.allclose(rt["rdu"], 0, atol=1e-12)
        assert np.allclose(rt["td"], np.eye(2), atol=1e-12) and np.allclose(rt["tu"], np.eye(2), atol=1e-12)
        assert abs(rt["rud0"]) < 1e-14 and abs(rt["td0"] - 1) < 1e-14


def test_sh_interface_energy_conservation(orc):
    """Lossless SH interface, propagating regime: 1 - R^2 = (mu_d nu_d)/(mu_u nu_u) T^2."""
    w = 2 * np.pi * 1.0 + 0j
    bu, bd, mu_u, mu_d = 3.0, 3.6, 3.0 ** 2 * 2.7, 3.6 ** 2 * 2.9
    for k in (0.1, 0.5, 1.0, 1.5):
        rt = orc.interface_rt(k, w, (6.0 + 0j, bu + 0j, mu_u + 0j), (6.5 + 0j, bd + 0j, mu_d + 0j))
        nu_u = np.sqrt((w / bu) ** 2 - k * k).real
        nu_d = np.sqrt((w / bd) ** 2 - k * k).real
        R, T = rt["rdu0"], rt["td0"]
        assert abs(1 - abs(R) ** 2 - (mu_d * nu_d) / (mu_u * nu_u) * abs(T) ** 2) < 1e-12


def test_split_layer_invariance(orc):
    """Cutting a layer into two identical layers changes nothing."""
    rows = models.C1_LAYERS.copy()
    split = np.vstack([rows[:1], [[6.0] + list(rows[1, 1:])], [[9.0] + list(rows[1, 1:])], rows[2:]])
    grid = dict(nw=64, dt=0.4, nk=512, kmax=3.0, omega_i=np.pi / 51.2, t0=0.0, taper=0.3)
    for kind in (0, 1):
        k1 = orc.kernels(rows, 10.0, 0.0, kind, grid, 5, 9)
        k2 = orc.kernels(split, 10.0, 0.0, kind, grid, 5, 9)
        assert np.abs(k1 - k2).max() < 1e-11 * np.abs(k1).max()


def test_explosion_radiates_only_p(orc):
    """Homogeneous medium, source and receiver 500 km down (surface
    reflections below e^-1000 for k > w/beta): the explosion's Z/R kernels are
    the direct P wave, Kz/Kr = nu_alpha/k exactly (an S wave would give
    k/nu_beta; the sign reflects Z positive upward with R = sum Kr J1)."""
    vp, vs, rho = 6.0, 3.46, 2.7
    rows = np.array([[1000.0, vp, vs, rho, 1e9, 1e9], [0.0, vp, vs, rho, 1e9, 1e9]])
    grid = dict(nw=32, dt=0.5, nk=400, kmax=8.0, omega_i=np.pi / 32.0, t0=0.0, taper=0.0)
    kern = orc.kernels(rows, 500.0, 495.0, 0, grid, 4, 12)
    df = 1.0 / (2 * 32 * 0.5)
    for i, iw in enumerate(range(4, 12)):
        w = 2 * np.pi * df * iw - 1j * grid["omega_i"]
        a = orc_disp(w, vp, 1e9)
        ks = np.arange(400) * (8.0 / 400)
        sel = ks > 1.5 * abs(w) / vs + 0.05
        nua = np.sqrt((ks[sel] - w / a) * (ks[sel] + w / a))
        ratio = kern[i, sel, 0] / kern[i, sel, 1]
        assert np.allclose(ratio, nua / ks[sel], rtol=1e-9)


def test_rayleigh_pole_location(orc):
    """Poisson half-space, nearly real frequency: the surface Z kernel of a
    buried vertical force peaks at the Rayleigh wavenumber w / (0.9194 beta)."""
    vs = 3.5
    vp = vs * np.sqrt(3.0)
    rows = np.array([[20.0, vp, vs, 2.7, 1e9, 1e9], [0.0, vp, vs, 2.7, 1e9, 1e9]])
    nw, dt = 64, 0.25
    T = 2 * nw * dt
    grid = dict(nw=nw, dt=dt, nk=4000, kmax=4.0, omega_i=0.02 / T, t0=0.0, taper=0.0)
    iw = 20
    kern = orc.kernels(rows, 2.0, 0.0, 1, grid, iw, iw + 1)[0, :, 0]
    ks = np.arange(4000) * 1e-3
    # Rayleigh speed of a Poisson solid: c/beta = sqrt(2 - 2/sqrt(3))
    cr = vs * np.sqrt(2 - 2 / np.sqrt(3))
    w = 2 * np.pi * iw / T
    kpk = ks[np.argmax(np.abs(kern))]
    assert abs(kpk - w / cr) < 2e-3


# (force depth, receiver depth) pairs covering every receiver branch and its mirror:
# receiver above the source in another layer, below it, in the source layer, at the surface
RECIPROCAL_PAIRS = ((10.0, 0.0), (20.0, 5.0), (5.0, 25.0), (3.0, 10.0))


def test_vertical_force_reciprocity(orc):
    """Elastodynamic reciprocity: the vertical displacement at depth b from a
    vertical force at depth a equals that at a from the force at b (physics
    mode; surface receivers and surface forces included).  A wrong sign or
    index in one receiver branch breaks the equality with its mirror."""
    grid = models.dense_grid(128, 0.1, 1024, float(models.C1_LAYERS[:, 2].min()))
    dists = [15.0, 40.0]
    for a, b in RECIPROCAL_PAIRS:
        z1 = orc.traces(models.C1_LAYERS, a, b, dists, 1, grid)[0][:, 0]
        z2 = orc.traces(models.C1_LAYERS, b, a, dists, 1, grid)[0][:, 0]
        assert (np.linalg.norm(z1 - z2, axis=-1) / np.linalg.norm(z1, axis=-1)).max() < 1e-11, (a, b)


@pytest.mark.parametrize("compat", [False, True])
def test_columns_reduce_to_zr_traces(orc, compat):
    """The oracle's displacement columns (the reference's integrand formulas
    with direct Bessel values) against its Z/R trace path (k-kernels times
    J0 / J1 tables): explosion Z = column 8 - column 9, R = column 3 -
    column 2; vertical force Z = single-force column 4, R = column 1.  Both
    modes, so compat's index slips in columns 2 and 3 cancel as they should."""
    grid = models.dense_grid(32, 0.2, 256, float(models.C1_LAYERS[:, 2].min()))
    dists = [15.0, 40.0]
    cols = orc.columns(models.C1_LAYERS, 10.0, 0.0, dists, grid, compat=compat)
    assert cols.shape == (15, 2, 64) and np.all(np.isfinite(cols))
    ex = orc.traces(models.C1_LAYERS, 10.0, 0.0, dists, 0, grid, compat=compat)[0]
    vf = orc.traces(models.C1_LAYERS, 10.0, 0.0, dists, 1, grid, compat=compat)[0]
    pairs = [(cols[8] - cols[9], ex[:, 0]), (cols[3] - cols[2], ex[:, 1]),
             (cols[14], vf[:, 0]), (cols[11], vf[:, 1])]
    for got, want in pairs:
        scale = np.abs(want).max(axis=-1, keepdims=True)
        assert np.abs(got - want).max() < 1e-11 * scale.max()


# Source-receiver reciprocity: the displacement at receiver x from a moment
# tensor at x' is the strain at x' of the displacement from a unit force at x.
# Tensor column c (source at depth a, receiver at depth b) against the
# single-force columns with the depths swapped (physics mode: columns 0-4
# displacements R_h, R_v, T_h, Z_h, Z_v; 5-9 their d/dz; 10-14 their d/dr;
# h: horizontal force, v: vertical force), r the distance.  The n = 0 terms
# are the vertical dipole (column 8: d/dz of Z_v) and the horizontal
# dilatation (column 9: d/dr R_v + R_v / r); column 6 carries the horizontal
# shear of the vertical force, the n = 1 and n = 2 terms those of the
# horizontal force.
RECIPROCITY = {
    0: [(-1, 10, 0), (1, 0, -1), (-1, 2, -1)],
    1: [(-1, 5, 0), (1, 13, 0)],
    2: [(1, 10, 0), (1, 0, -1), (-1, 2, -1)],
    3: [(1, 8, 0)],
    4: [(-1, 12, 0), (-1, 0, -1), (1, 2, -1)],
    5: [(-1, 7, 0), (1, 3, -1)],
    6: [(-1, 11, 0), (1, 1, -1)],
    7: [(-1, 6, 0), (1, 14, 0)],
    8: [(1, 9, 0)],
    9: [(1, 11, 0), (1, 1, -1)],
}  # column: [(sign, single-force column, power of r)]


def _reciprocity_residuals(orc, a, b):
    grid = models.dense_grid(32, 0.2, 512, float(models.C1_LAYERS[:, 2].min()))
    dists = np.array([12.0, 30.0])
    T = orc.columns(models.C1_LAYERS, a, b, dists, grid, all30=True)[0]
    F = orc.columns(models.C1_LAYERS, b, a, dists, grid, all30=True)[1]
    r = dists[:, None]
    out = {}
    for c, terms in RECIPROCITY.items():
        rhs = sum(s * F[fc] * r ** p for s, fc, p in terms)
        out[c] = np.linalg.norm(T[c] - rhs) / np.linalg.norm(T[c])
    return out


@pytest.mark.parametrize("a,b", [(10.0, 3.0), (5.0, 25.0), (10.0, 14.0), (10.0, 4.0), (3.0, 10.0)])
def test_tensor_columns_reciprocity(orc, a, b):
    """All ten tensor columns against the swapped single-force strains, every
    receiver branch of the tensor solve (physics mode)."""
    res = _reciprocity_residuals(orc, a, b)
    assert max(res.values()) < 1e-10, res


def test_reference_tensor_sh_factors_break_reciprocity(orc):
    """The reference's SH factors of the tensor source (1 + R above, R - 1
    below, for both orders; green.cu:1408-1411; compat mode keeps them),
    planted in the oracle's physics mode, break the n = 1 relations above the
    source and the n = 2 relations below it, and nothing else."""
    try:
        orc.set_tensor_sh_reference(1)
        above = _reciprocity_residuals(orc, 10.0, 3.0)
        below = _reciprocity_residuals(orc, 5.0, 25.0)
    finally:
        orc.set_tensor_sh_reference(0)
    assert above[1] > 1e-3 and above[5] > 1e-3 and below[0] > 1e-3 and below[4] > 1e-3
    assert max(v for c, v in above.items() if c not in (1, 5)) < 1e-10
    assert max(v for c, v in below.items() if c not in (0, 4)) < 1e-10


def test_free_surface_traction(orc):
    """Receiver on the free surface (physics mode): zero traction fixes the
    d/dz coefficients of both sources, a.t[0] = k a[1] (horizontal, from
    tau_rz = 0), a.t[1] = -k (1 - 2 b^2/a^2) a[0] (vertical, from tau_zz = 0)
    and b.t = 0 (SH, from tau_zphi = 0), for every order."""
    w = 2 * np.pi * 0.7 - 0.05j
    rows = [(2.0, 5.0, 2.9, 2.5, 400, 200), (6.0, 6.1, 3.5, 2.75, 600, 300), (20.0, 6.6, 3.8, 2.9, 800, 400)]
    lay = np.zeros((3, 7))
    for j, (h, vp, vs, rho, qp, qs) in enumerate(rows):
        a = orc.dispersion(w, vp, qp)
        b = orc.dispersion(w, vs, qs)
        mu = rho * b * b
        lay[j] = [h, a.real, a.imag, b.real, b.imag, mu.real, mu.imag]
    ratio = complex(lay[0, 3], lay[0, 4]) / complex(lay[0, 1], lay[0, 2])
    kg = 1 - 2 * ratio ** 2
    for src in (0, 1):
        for k in (0.05, 0.4, 1.5, 4.0):
            c = orc.cell(0, src, k, w, lay, 1, 3.0, 0, 0.0)
            for i in range(0, 6 if src else 4, 2):  # a0, a1 (, a2)
                a, at = c[i:i + 2], c[8 + i:10 + i]
                scale = abs(a).max() * k
                assert abs(at[0] - k * a[1]) < 1e-10 * scale
                assert abs(at[1] + k * kg * a[0]) < 1e-10 * scale
            assert c[14] == 0 and c[15] == 0
