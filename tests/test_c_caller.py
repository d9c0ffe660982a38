"""A compiled C program through the boundary: tests/c_caller/lj_nve.c, the
reference's own driver's call sequence (cuda/tcpu_0.c:27-60), built with
``gcc -std=c99 -I include`` and linked against seplib_amd/libsep.so.

  * CPU: it compiles and links against the headers and the library alone
    (every symbol it uses resolves);
  * GPU: run on golden G1's initial state (LJ N=1000, skin 0.5) it prints,
    per step, the sepret energies, both stress tensors and tnow of the
    compiled reference's run (first 60 steps, 1e-10 of each column's scale;
    tnow exactly), and writes the final configuration with sep_save_xyz.
"""
import os
import subprocess

import numpy as np
import pytest

from seplib_amd import workloads as W
from tests.golden import runs as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_caller", "lj_nve.c")
LIBDIR = os.path.join(ROOT, "seplib_amd")


def _build(out_dir):
    exe = os.path.join(str(out_dir), "lj_nve")
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", exe, "-L", LIBDIR, "-lsep", "-Wl,-rpath," + LIBDIR, "-lm"],
                   check=True, capture_output=True, text=True)
    return exe


def test_c_program_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libsep.so")):
        pytest.skip("libsep.so not built")
    exe = _build(tmp_path)
    assert os.path.getsize(exe) > 0


@pytest.mark.gpu
def test_c_program_matches_reference_golden(tmp_path):
    g = np.load(os.path.join(ROOT, "tests", "golden", "g1_lj_nve.npz"), allow_pickle=False)
    w = R.workload_from(g)
    start = str(tmp_path / "start.xyz")
    final = str(tmp_path / "final.xyz")
    W.write_xyz(start, w)
    exe = _build(tmp_path)
    n = 60
    p = subprocess.run([exe, start, str(n), "0.5", final], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    rows = [l.split()[2:] for l in p.stdout.splitlines() if l.startswith("step ")]
    got = np.array(rows, dtype=float)
    assert got.shape == (n, 21)
    sc = g["scal"][:n]
    ref = np.concatenate([sc[:, :2], sc[:, 3:21], sc[:, 21:22]], axis=1)
    for c in range(20):
        scale = np.abs(ref[:, c]).max()
        if scale > 0:
            assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-10 * scale, c
    assert np.array_equal(got[:, 20], ref[:, 20])
    # sep_save_xyz: "N", the box, then one "type x y z vx vy vz m q" line per atom
    with open(final) as f:
        lines = f.read().splitlines()
    assert int(lines[0]) == len(w["m"]) and len(lines) == len(w["m"]) + 2
    x = np.array([[float(v) for v in l.split()[1:4]] for l in lines[2:]])
    assert np.all((x >= 0.0) & (x <= w["L"][0]))
