"""Green's-function files for the reference's Python workflow without obspy:
binary SAC traces, the fk-style ``{dist}.grn.{i}`` set the reference's
drivers compare against (green1.py:104-110) and their L1 misfit
(green1.py:124-135).

SAC is the IRIS/LLNL binary format: a 632-byte header of 70 float32, 40
int32 and 192 bytes of 8/16-character strings, then NPTS float32 samples
(little-endian here; readers detect the byte order from NVHDR = 6)."""
import os

import numpy as np

UNDEF = -12345.0
# float header words (index into the 70 floats)
_F = {"delta": 0, "depmin": 1, "depmax": 2, "b": 5, "e": 6, "o": 7, "stla": 31, "stlo": 32, "evla": 35,
      "evlo": 36, "evdp": 38, "dist": 50, "az": 51, "baz": 52, "gcarc": 53, "depmen": 56}
# integer header words (index into the 40 ints)
_I = {"nzyear": 0, "nzjday": 1, "nzhour": 2, "nzmin": 3, "nzsec": 4, "nzmsec": 5, "nvhdr": 6, "npts": 9,
      "iftype": 15, "idep": 16, "iztype": 17, "leven": 35, "lpspol": 36, "lovrok": 37, "lcalda": 38}
ITIME = 1  # iftype: time series, evenly spaced
IB = 9     # iztype: reference time = begin time

# single-force columns (0: horizontal force R, 1: vertical force R,
# 2: horizontal force T, 3: horizontal force Z, 4: vertical force Z, 5: first
# d/dz column) written as grn.0 .. grn.5, grn.2 zeroed, as the reference's
# driver reads them (green1.py:104-110)
GRN_COLUMNS = (4, 1, 5, 3, 0, 2)
GRN_ZERO = (2,)


def write_sac(path, data, delta, b=0.0, **header):
    """Write one evenly spaced trace as binary SAC.  ``header``: any of the
    float or integer fields named in this module (dist, evdp, ...)."""
    data = np.ascontiguousarray(data, dtype="<f4")
    hf = np.full(70, UNDEF, dtype="<f4")
    hi = np.full(40, int(UNDEF), dtype="<i4")
    hs = np.frombuffer((b"-12345  " * 24), dtype="S1").copy()
    hf[_F["delta"]] = delta
    hf[_F["b"]] = b
    hf[_F["e"]] = b + delta * (len(data) - 1)
    if len(data):
        hf[_F["depmin"]], hf[_F["depmax"]], hf[_F["depmen"]] = data.min(), data.max(), data.mean()
    hi[_I["nvhdr"]] = 6
    hi[_I["npts"]] = len(data)
    hi[_I["iftype"]] = ITIME
    hi[_I["iztype"]] = IB
    hi[_I["leven"]] = 1
    hi[_I["lpspol"]] = 1
    hi[_I["lovrok"]] = 1
    hi[_I["lcalda"]] = 0
    for k, v in header.items():
        if k in _F:
            hf[_F[k]] = v
        elif k in _I:
            hi[_I[k]] = v
        else:
            raise KeyError(f"unknown SAC header field {k}")
    with open(path, "wb") as f:
        f.write(hf.tobytes())
        f.write(hi.tobytes())
        f.write(hs.tobytes())
        f.write(data.tobytes())


def read_sac(path):
    """(data float32, header dict of the named fields) from a binary SAC file
    of either byte order."""
    raw = open(path, "rb").read()
    if len(raw) < 632:
        raise ValueError(f"{path}: not a SAC file (shorter than its header)")
    for order in ("<", ">"):
        hi = np.frombuffer(raw, dtype=order + "i4", count=40, offset=280)
        if hi[_I["nvhdr"]] == 6:
            break
    else:
        raise ValueError(f"{path}: not a SAC file (NVHDR is not 6)")
    hf = np.frombuffer(raw, dtype=order + "f4", count=70, offset=0)
    npts = int(hi[_I["npts"]])
    if len(raw) != 632 + 4 * npts:
        raise ValueError(f"{path}: {len(raw)} bytes for NPTS = {npts}")
    data = np.frombuffer(raw, dtype=order + "f4", count=npts, offset=632).astype(np.float32)
    head = {k: float(hf[i]) for k, i in _F.items()}
    head.update({k: int(hi[i]) for k, i in _I.items()})
    return data, head


def grn_traces(columns):
    """The six grn traces from single-force column traces [N] each (the
    columns of green()'s oname_t, or a column plan's five single-force
    columns): grn.i = column GRN_COLUMNS[i], grn.2 zero."""
    columns = np.asarray(columns, dtype=np.float64)
    n = columns.shape[-1]
    out = np.zeros((len(GRN_COLUMNS), n))
    for i, c in enumerate(GRN_COLUMNS):
        if i not in GRN_ZERO:
            out[i] = columns[c]
    return out


def write_grn(directory, dist, columns, delta, t0=0.0, src_depth=None):
    """Write ``{dist}.grn.0`` .. ``.grn.5`` (SAC) under ``directory`` from
    single-force column traces; the first sample is at t0.  Returns the paths."""
    os.makedirs(directory, exist_ok=True)
    tr = grn_traces(columns)
    name = f"{dist:g}" if float(dist) != int(dist) else f"{int(dist)}"
    paths = []
    for i, d in enumerate(tr):
        p = os.path.join(directory, f"{name}.grn.{i}")
        extra = {"dist": float(dist)}
        if src_depth is not None:
            extra["evdp"] = float(src_depth)
        write_sac(p, d, delta, b=t0, **extra)
        paths.append(p)
    return paths


def misfit(traces, refs, divisor=None):
    """The reference drivers' L1 measure (green1.py:124-135): per pair
    E = sum | |d| - |d_ref| | / (sum |d_ref| + sum |d|), pairs whose reference
    is all zero skipped, the sum divided by len(pairs) - 1 (5 for the six grn
    files, as there) unless a divisor is given."""
    total = 0.0
    for d, r in zip(traces, refs):
        ad, ar = np.abs(np.asarray(d, dtype=np.float64)), np.abs(np.asarray(r, dtype=np.float64))
        n = min(len(ad), len(ar))
        ad, ar = ad[:n], ar[:n]
        if ar.sum() == 0:
            continue
        total += np.abs(ad - ar).sum() / (ar.sum() + ad.sum())
    return total / (divisor if divisor is not None else max(1, len(refs) - 1))
