"""input.txt reading/writing and Earth flattening, as the reference's Python
drivers do it (green1.py:8-48 readtxt/readpar, green1.py:54-93 writer)."""
import math

import numpy as np

R_EARTH = 6371.0


def read_output(path):
    """Read an oname_t / oname_w file (rows of comma-terminated numbers)."""
    rows = []
    with open(path) as f:
        for line in f:
            vals = [float(v) for v in line.strip().split(",") if v.strip()]
            if vals:
                rows.append(vals)
    return np.array(rows)


def read_model(path, flatten=False):
    """Depth-to-top model rows (z, vp, vs, rho, qp, qs) -> input.txt rows,
    optionally Earth-flattened with the layer-centre radius factor
    s = R / (r + h/2) applied to thickness and velocities (green1.py:14-48)."""
    src = []
    with open(path) as f:
        for line in f:
            vals = [float(v) for v in line.strip().split(",") if v.strip()]
            if vals:
                src.append(vals)
    out = []
    r = R_EARTH
    thick = []
    for i, (z, vp, vs, rho, qp, qs) in enumerate(src):
        sh = 0.0 if i == len(src) - 1 else src[i + 1][0] - z
        r -= sh
        s = R_EARTH / (r + 0.5 * sh) if flatten else 1.0
        thick.append(sh * s)
        out.append([0.0, vp * s, vs * s, rho, qp, qs])
    z = 0.0
    for i in range(len(out)):
        if i > 0:
            z += thick[i - 1]
        out[i][0] = z
    return np.array(out)


def flatten_depth(d):
    """Earth-flattening transform of a source / receiver depth."""
    return R_EARTH * math.log(R_EARTH / (R_EARTH - d))


def write_input(path, rows, src_depth, rcv_depth, dt, dist, log2n, t0, taper=0.3, src_type=0,
                n_ptam=7, nf1=1, nf2=1000, twin=33.33):
    """Write input.txt in the reference's format (parsed by green.cu:1989-2060)."""
    with open(path, "w") as f:
        f.write(f"{src_depth} \n{rcv_depth}\n{len(rows)}\n")
        for r in rows:
            f.write(",".join(repr(float(v)) for v in r) + "\n")
        f.write(f"{dt}\n{dist} \n{nf1}\n{nf2}\n{twin} \n{log2n}\n{t0}\n{taper}\n{n_ptam}\n{src_type}\n\n")
