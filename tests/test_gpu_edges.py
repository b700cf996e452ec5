"""Geometry and grid edge cases of the trace path against the FP64 oracle
(physics mode), with per-trace rel-L2 and peak error: one layer over a
half-space, a source in the half-space, buried receivers above and below the
source and in the source layer, and small or odd grids (one receiver, 4
frequencies, K not a multiple of the contraction tiles); and source-receiver
reciprocity of the vertical-force response on the GPU path."""
import numpy as np
import pytest

import grtm_cuda_amd as g
from grtm_cuda_amd import models

pytestmark = pytest.mark.gpu

ONE_LAYER = np.array([[5.0, 6.0, 3.5, 2.7, 600.0, 300.0], [0.0, 8.0, 4.5, 3.3, 1200.0, 600.0]])

# name, layers, source depth, receiver depth, distances, (nw, dt, nk)
CASES = [
    ("one layer over a half-space", ONE_LAYER, 2.0, 0.0, [10.0, 30.0], (64, 0.1, 512)),
    ("source in the half-space", models.C1_LAYERS, 40.0, 0.0, [20.0, 60.0], (64, 0.1, 512)),
    ("buried receiver above the source", models.C1_LAYERS, 20.0, 5.0, [15.0, 40.0], (64, 0.1, 512)),
    ("buried receiver below the source", models.C1_LAYERS, 5.0, 20.0, [15.0, 40.0], (64, 0.1, 512)),
    ("receiver below the source in its layer", models.C1_LAYERS, 3.0, 10.0, [12.0], (64, 0.1, 512)),
    ("one receiver, odd wavenumber count", models.C1_LAYERS, 10.0, 0.0, [7.5], (16, 0.2, 300)),
    ("four frequencies", models.C1_LAYERS, 10.0, 0.0, [25.0, 50.0, 75.0], (4, 0.5, 256)),
]


def errors(tr, ref):
    """Per trace: rel-L2 and peak error max|u - u_ref| / max|u_ref|."""
    d = tr - ref
    return (np.linalg.norm(d, axis=-1) / np.linalg.norm(ref, axis=-1),
            np.abs(d).max(axis=-1) / np.abs(ref).max(axis=-1))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_edge_geometry_matches_oracle(orc, case):
    name, layers, zs, zr, dists, (nw, dt, nk) = case
    grid = models.dense_grid(nw, dt, nk, float(layers[:, 2].min()))
    print(f"gpu: edge case '{name}'", flush=True)
    ref = np.stack([orc.traces(layers, zs, zr, dists, kind, grid, compat=False)[0].transpose(1, 0, 2)
                    for kind in (0, 1)])  # [q, c, r, N]
    # Four frequencies, 256 wavenumbers and 8 samples: the far receivers' k-sums
    # cancel by about 1e4, so FP32 accumulation in the contraction (every row's
    # recursion is already FP64 here: at least 4 rows are) ends near 5e-4.
    tol32 = 1e-3 if nw == 4 else 1e-4
    for prec, tol in ((g.PREC_FP64, 1e-7), (g.PREC_FP32, tol32)):
        tr = g.TracePlan(layers, zs, zr, dists, grid, precision=prec).traces()
        rl, pk = errors(tr, ref)
        print(f"  precision {prec}: max rel-L2 {rl.max():.2e}, max peak error {pk.max():.2e}")
        assert rl.max() < tol and pk.max() < tol, (name, prec)


@pytest.mark.parametrize("prec,tol", [(g.PREC_FP64, 1e-11), (g.PREC_FP32, 1e-5)])
def test_vertical_force_reciprocity(prec, tol):
    """Vertical displacement at depth b from a vertical force at depth a equals
    that at a from the force at b, on the GPU path (every receiver branch
    against its mirror)."""
    grid = models.dense_grid(128, 0.1, 1024, float(models.C1_LAYERS[:, 2].min()))
    dists = [15.0, 40.0]
    print(f"gpu: reciprocity, precision {prec}", flush=True)
    for a, b in ((10.0, 0.0), (20.0, 5.0), (5.0, 25.0), (3.0, 10.0)):
        z1 = g.TracePlan(models.C1_LAYERS, a, b, dists, grid, precision=prec).traces()[1, 0]
        z2 = g.TracePlan(models.C1_LAYERS, b, a, dists, grid, precision=prec).traces()[1, 0]
        err = (np.linalg.norm(z1 - z2, axis=-1) / np.linalg.norm(z1, axis=-1)).max()
        print(f"  force {a} km, receiver {b} km: {err:.2e}")
        assert err < tol, (a, b)
