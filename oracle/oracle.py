"""ctypes wrapper of the FP64 CPU oracle (oracle/libgrtm_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker.  The product (grtm_cuda_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrtm_oracle.so")
# the same restatement computed in binary128 (__float128); same API, in double
LIB_PATH_QUAD = os.path.join(_HERE, "libgrtm_oracle_q.so")
P = ctypes.c_void_p
D = ctypes.c_double
_LIBS = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(quad=False):
    """The FP64 oracle, or with quad=True its binary128 build."""
    if quad not in _LIBS:
        path = LIB_PATH_QUAD if quad else LIB_PATH
        if not os.path.exists(path):
            build()
        h = ctypes.CDLL(path)
        h.or_bessel.restype = D
        h.or_bessel.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, D]
        _LIBS[quad] = h
    return _LIBS[quad]


class Run(ctypes.Structure):
    _fields_ = [("n_layer", ctypes.c_int), ("h", D * 64), ("a0", D * 64), ("b0", D * 64), ("rho", D * 64),
                ("qp", D * 64), ("qs", D * 64), ("hs", D), ("hr", D), ("s_layer", ctypes.c_int),
                ("r_layer", ctypes.c_int), ("zs", D), ("zr", D), ("dt", D), ("dist", D), ("twin", D), ("t0", D),
                ("taper", D), ("nf1", ctypes.c_int), ("nf2", ctypes.c_int), ("log2n", ctypes.c_int),
                ("n_ptam", ctypes.c_int), ("src_type", ctypes.c_int), ("split_done", ctypes.c_int)]


def parse_input(path, compat=True):
    r = Run()
    st = lib().or_parse_input(path.encode(), int(compat), ctypes.byref(r))
    return st, r


def ref_grid(run, compat=True, quad=False):
    dk, oi, nkw = D(), D(), ctypes.c_longlong()
    nw = 1 << (run.log2n - 1)
    nk = np.zeros(nw, dtype=np.int32)
    lib(quad).or_ref_grid(ctypes.byref(run), int(compat), ctypes.byref(dk), ctypes.byref(oi), nk.ctypes.data_as(P),
                      ctypes.byref(nkw))
    return dk.value, oi.value, nk, nkw.value


def green_rows(iname, rows, compat=True, nthreads=0, quad=False):
    """Post-processed spectrum rows: array [len(rows), 30] complex."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((len(rows), 60))
    st = lib(quad).or_green_rows(iname.encode(), int(compat), rows.ctypes.data_as(P), len(rows),
                             out.ctypes.data_as(P), int(nthreads))
    if st:
        raise RuntimeError(f"oracle green_rows status {st}")
    return out[:, 0::2] + 1j * out[:, 1::2]


def green_files(iname, oname_w, oname_t, compat=True, nthreads=0):
    return lib().or_green_files(iname.encode(), oname_w.encode(), oname_t.encode(), int(compat), int(nthreads))


def set_inverse_mode(m, quad=False):
    lib(quad).or_set_inverse_mode(int(m))


def set_tensor_sh_reference(on):
    """Test switch: the reference's SH factors of the tensor source in physics mode too."""
    lib().or_set_tensor_sh_reference(int(on))


def dispersion(w, v0, q):
    o = np.zeros(2)
    lib().or_dispersion(D(w.real), D(w.imag), D(v0), D(q), o.ctypes.data_as(P))
    return o[0] + 1j * o[1]


def cell(compat, src_type, k, w, lay, s_layer, zs, r_layer, zr, par0_rv=None, quad=False):
    lay = np.ascontiguousarray(lay, dtype=np.float64)
    out = np.zeros(36)
    p0 = None if par0_rv is None else np.ascontiguousarray(par0_rv, dtype=np.float64)
    st = lib(quad).or_cell(int(compat), int(src_type), D(k), D(w.real), D(w.imag), lay.shape[0], lay.ctypes.data_as(P),
                       int(s_layer), D(zs), int(r_layer), D(zr), None if p0 is None else p0.ctypes.data_as(P),
                       out.ctypes.data_as(P))
    if st:
        raise RuntimeError(f"oracle cell status {st}")
    return out[0::2] + 1j * out[1::2]


def interface_rt(k, w, up, dn):
    """up/dn: (alpha, beta, mu) complex.  Returns dict of 2x2 blocks + SH."""
    u = np.array([up[0].real, up[0].imag, up[1].real, up[1].imag, up[2].real, up[2].imag])
    d = np.array([dn[0].real, dn[0].imag, dn[1].real, dn[1].imag, dn[2].real, dn[2].imag])
    out = np.zeros(40)
    lib().or_interface_rt(D(k), D(w.real), D(w.imag), u.ctypes.data_as(P), d.ctypes.data_as(P),
                          out.ctypes.data_as(P))
    z = out[0::2] + 1j * out[1::2]
    return dict(td=z[0:4].reshape(2, 2), rud=z[4:8].reshape(2, 2), rdu=z[8:12].reshape(2, 2),
                tu=z[12:16].reshape(2, 2), td0=z[16], rud0=z[17], rdu0=z[18], tu0=z[19])


def bessel(mode, which, n, x):
    return lib().or_bessel(int(mode), int(which), int(n), D(x))


def traces(layers, src_depth, rcv_depth, dists, src_kind, grid, compat=False, w_range=None, nthreads=0,
           quad=False, ptam=None):
    """Z/R traces [nr, 2, N] and post-processed spectra [nr, 2, nw] of one source kind
    (0 explosion, 1 vertical force) on the dense grid.  ptam=(n_ptam, nt): the
    tail of every k-sum from nt further wavenumbers by peak-trough averaging."""
    layers = np.ascontiguousarray(layers, dtype=np.float64)
    dists = np.ascontiguousarray(dists, dtype=np.float64)
    nw = int(grid["nw"])
    lo, hi = (0, nw) if w_range is None else w_range
    tr = np.zeros((len(dists), 2, 2 * nw))
    sp = np.zeros((len(dists), 2, nw, 2))
    n_ptam, nt = ptam if ptam else (0, 0)
    st = lib(quad).or_traces_ptam(int(compat), layers.ctypes.data_as(P), layers.shape[0], D(src_depth),
                                  D(rcv_depth), dists.ctypes.data_as(P), len(dists), int(src_kind), nw,
                                  D(grid["dt"]), int(grid["nk"]), D(grid["kmax"]), D(grid["omega_i"]),
                                  D(grid.get("t0", 0.0)), D(grid.get("taper", 0.0)), int(lo), int(hi), int(n_ptam),
                                  int(nt), tr.ctypes.data_as(P), sp.ctypes.data_as(P), int(nthreads))
    if st:
        raise RuntimeError(f"oracle traces status {st}")
    return tr, sp[..., 0] + 1j * sp[..., 1]


def ptam(s0, terms, n):
    """PTAM estimate of one complex k-sum: s0 the sum so far, terms the further
    complex terms; returns (estimate, extrema seen)."""
    t = np.ascontiguousarray(np.stack([np.real(terms), np.imag(terms)], axis=-1).ravel(), dtype=np.float64)
    out = np.zeros(2)
    found = ctypes.c_int()
    st = lib().or_ptam(D(np.real(s0)), D(np.imag(s0)), t.ctypes.data_as(P), len(terms), int(n), out.ctypes.data_as(P),
                       ctypes.byref(found))
    if st:
        raise RuntimeError(f"oracle ptam status {st}")
    return out[0] + 1j * out[1], found.value


N_COLUMNS = 15


def columns(layers, src_depth, rcv_depth, dists, grid, compat=False, nthreads=0, all30=False, quad=False):
    """Displacement-column traces [15, nr, N]: tensor columns 0-9, then
    single-force columns 0-4, on the dense grid.  all30: every integrand
    column of both sources instead, [2 (tensor, single force), 30, nr, N]."""
    layers = np.ascontiguousarray(layers, dtype=np.float64)
    dists = np.ascontiguousarray(dists, dtype=np.float64)
    nw = int(grid["nw"])
    tr = np.zeros((60 if all30 else N_COLUMNS, len(dists), 2 * nw))
    st = lib(quad).or_columns_all(int(compat), int(all30), layers.ctypes.data_as(P), layers.shape[0], D(src_depth),
                              D(rcv_depth), dists.ctypes.data_as(P), len(dists), nw, D(grid["dt"]), int(grid["nk"]),
                              D(grid["kmax"]), D(grid["omega_i"]), D(grid.get("t0", 0.0)), D(grid.get("taper", 0.0)),
                              tr.ctypes.data_as(P), int(nthreads))
    if st:
        raise RuntimeError(f"oracle columns status {st}")
    return tr.reshape(2, 30, len(dists), 2 * nw) if all30 else tr


def kernels(layers, src_depth, rcv_depth, src_kind, grid, w_lo, w_hi, compat=False, nthreads=0, quad=False):
    """Unweighted Z/R k-kernels [w_hi-w_lo, nk, 2] complex."""
    layers = np.ascontiguousarray(layers, dtype=np.float64)
    nk = int(grid["nk"])
    out = np.zeros((w_hi - w_lo, nk, 4))
    st = lib(quad).or_kernels(int(compat), layers.ctypes.data_as(P), layers.shape[0], D(src_depth), D(rcv_depth),
                          int(src_kind), int(grid["nw"]), D(grid["dt"]), nk, D(grid["kmax"]), D(grid["omega_i"]),
                          int(w_lo), int(w_hi), out.ctypes.data_as(P), int(nthreads))
    if st:
        raise RuntimeError(f"oracle kernels status {st}")
    return out[..., 0::2] + 1j * out[..., 1::2]


def post_ifft(spec_raw, df, omega_i, dt, t0, taper):
    spec_raw = np.ascontiguousarray(spec_raw, dtype=np.complex128)
    nw = spec_raw.shape[0]
    raw = np.zeros(2 * nw)
    raw[0::2] = spec_raw.real
    raw[1::2] = spec_raw.imag
    post = np.zeros(2 * nw)
    tr = np.zeros(2 * nw)
    lib().or_post_ifft(nw, raw.ctypes.data_as(P), D(df), D(omega_i), D(dt), D(t0), D(taper), post.ctypes.data_as(P),
                       tr.ctypes.data_as(P))
    return post[0::2] + 1j * post[1::2], tr
