"""CPU checks of the drivers' input tooling and of a C caller of the drop-in
symbol."""
import os
import re
import subprocess

import numpy as np

from grtm_cuda_amd import inputfile
from conftest import GOLDEN, ROOT


def _numbers(path):
    return [float(x) for x in re.findall(r"[-+0-9.eE]+", open(path).read())]


def test_flattening_and_writer_reproduce_the_reference_input(tmp_path):
    """green6.py:50-93 turns demo/crust1.0.txt (depth-to-top rows) into
    demo/input.txt with Earth flattening (green1.py:14-48, r_earth 6371) and
    flattened source / receiver depths: read_model + flatten_depth +
    write_input give the reference's file number for number."""
    rows = inputfile.read_model(os.path.join(GOLDEN, "demo_crust_model.txt"), flatten=True)
    p = tmp_path / "input.txt"
    inputfile.write_input(str(p), rows, inputfile.flatten_depth(8.0), inputfile.flatten_depth(0.0), 2, 10000, 11,
                          949.04, taper=0.3, src_type=0, n_ptam=7)
    assert _numbers(str(p)) == _numbers(os.path.join(GOLDEN, "demo_input.txt"))


def test_read_model_without_flattening_keeps_the_rows():
    src = np.array(_numbers(os.path.join(GOLDEN, "demo_crust_model.txt"))).reshape(-1, 6)
    rows = inputfile.read_model(os.path.join(GOLDEN, "demo_crust_model.txt"), flatten=False)
    assert np.array_equal(rows, src)
    # flattening stretches every layer and velocity by R / (r + h / 2) > 1
    flat = inputfile.read_model(os.path.join(GOLDEN, "demo_crust_model.txt"), flatten=True)
    assert np.all(flat[1:, 0] > src[1:, 0]) and np.all(flat[:, 1:3] > src[:, 1:3])
    assert np.array_equal(flat[:, 3:], src[:, 3:])


def test_c_caller_builds_against_the_header_and_reports_status(tmp_path):
    """tests/c_caller/green_main.c, the reference's main.cu call compiled
    against include/grtm.h with gcc and linked to libgrtm.so; without a GPU
    green() records GRTM_E_DEVICE, which the program reports."""
    exe = tmp_path / "green_main"
    subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "c_caller", "green_main.c"), "-L", os.path.join(ROOT, "grtm_cuda_amd"),
                    "-lgrtm", "-Wl,-rpath," + os.path.join(ROOT, "grtm_cuda_amd")], check=True)
    r = subprocess.run([str(exe), os.path.join(GOLDEN, "demo_input.txt"), "w.txt", "t.txt"], cwd=tmp_path,
                       capture_output=True, text=True)
    if r.returncode == 0:  # a GPU is present
        assert (tmp_path / "w.txt").exists() and (tmp_path / "t.txt").exists()
    else:
        assert r.returncode == 1 and "HIP device error" in r.stderr
    bad = subprocess.run([str(exe), str(tmp_path / "missing.txt")], cwd=tmp_path, capture_output=True, text=True)
    assert bad.returncode == 1 and "file i/o error" in bad.stderr
