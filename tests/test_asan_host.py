"""The host layer's binary checkpoint path under AddressSanitizer (CPU).

tests/asan/save_load.c is compiled together with the library's host
sources (seplib_amd/csrc/host/*.c, instrumented) and linked against
libsep.so for the device layer, which these calls never reach.  One
process saves a system with sep_save; another, with a different neighbour
capacity, loads it with sep_load, writes every new neighbour array to its
full length and closes it with sep_close.  The file holds the first
process's neighb pointers: sep_load must free the old arrays before it
reads and give every atom a new one after, so that sep_close frees only
what this process allocated.  Any invalid free, double free or overflow
aborts the loader with an AddressSanitizer report.
"""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "seplib_amd")


def _build(out_dir):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not os.path.exists(os.path.join(LIBDIR, "libsep.so")):
        pytest.skip("libsep.so not built")
    exe = os.path.join(str(out_dir), "save_load")
    host = sorted(glob.glob(os.path.join(LIBDIR, "csrc", "host", "*.c")))
    cmd = ["gcc", "-std=c99", "-g", "-O1", "-D_GNU_SOURCE", "-fsanitize=address",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(LIBDIR, "csrc"), os.path.join(ROOT, "tests", "asan", "save_load.c")
           ] + host + ["-o", exe, "-L", LIBDIR, "-lsep", "-Wl,-rpath," + LIBDIR, "-lm"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    return exe


def test_load_then_close_of_a_file_saved_by_another_process(tmp_path):
    exe = _build(tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    f = str(tmp_path / "atoms.bin")
    for n, nn in ((1, 1), (37, 5), (1000, 64)):
        p = subprocess.run([exe, "save", f, str(n)], capture_output=True, text=True, env=env,
                           timeout=120)
        assert p.returncode == 0, p.stderr[-3000:]
        p = subprocess.run([exe, "load", f, str(n), str(nn)], capture_output=True, text=True,
                           env=env, timeout=120)
        assert p.returncode == 0 and "AddressSanitizer" not in p.stderr, p.stderr[-3000:]
        assert p.stdout.strip().endswith("ok"), p.stdout
