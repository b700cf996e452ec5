"""grtm_green_mem_noalloc, the drop-in run without device allocation (its
device memory is the library's fixed per-GPU scratch), against
grtm_green_mem: bit for bit, on consecutive calls that reuse the scratch --
the 2048-sample demo, then a smaller run (the reference's example input:
fewer wavenumbers per row, so a smaller carve of the scratch), then the demo
again."""
import os

import numpy as np
import pytest

import grtm_cuda_amd as g
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
DEMO = os.path.join(GOLDEN, "demo_input.txt")
EXAMPLE = os.path.join(GOLDEN, "example_input.txt")


@pytest.mark.parametrize("compat", [True, False])
def test_noalloc_matches_allocating_entry_point(compat):
    for path in (DEMO, EXAMPLE, DEMO):
        print(f"gpu: green_mem / green_mem_noalloc {os.path.basename(path)} compat {compat}", flush=True)
        sa, ta = g.green_mem(path, compat=compat)
        sb, tb = g.green_mem(path, compat=compat, noalloc=True)
        assert sa.shape == sb.shape and ta.shape == tb.shape
        assert np.array_equal(sa, sb) and np.array_equal(ta, tb), path
        assert np.isfinite(ta).all() and np.abs(ta).max() > 0
