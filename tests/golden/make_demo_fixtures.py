"""Writes the golden fixtures of the drop-in path from the reference's own
output files (demo/input.txt -> demo/oname_w.txt, demo/oname_t.txt) and its
example inputs.  Data only: inputs and expected outputs.

Run in the build container: python tests/golden/make_demo_fixtures.py
"""
import os
import shutil

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    w = np.loadtxt(os.path.join(REF, "demo/oname_w.txt"), delimiter=",", usecols=range(60))
    t = np.loadtxt(os.path.join(REF, "demo/oname_t.txt"), delimiter=",", usecols=range(30))
    nw = w.shape[0] // 2
    spec = w[:, 0::2] + 1j * w[:, 1::2]
    # live columns 0-14 (15-29 are exactly zero in the reference output)
    assert np.all(spec[:, 15:] == 0) and np.all(t[:, 15:] == 0)
    np.save(os.path.join(HERE, "demo_spec.npy"), spec[:nw, :15].astype(np.complex128))
    np.save(os.path.join(HERE, "demo_trace.npy"), t[:, :15].astype(np.float64))
    shutil.copy(os.path.join(REF, "demo/input.txt"), os.path.join(HERE, "demo_input.txt"))
    shutil.copy(os.path.join(REF, "cusrc/example/crust1.0.txt"), os.path.join(HERE, "example_input.txt"))
    shutil.copy(os.path.join(REF, "demo/crust1.0.txt"), os.path.join(HERE, "demo_crust_model.txt"))


if __name__ == "__main__":
    main()
