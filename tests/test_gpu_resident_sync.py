"""The resident step against the synchronised step, in both precision
forms.  A resident graph run of [sep_force_pairs, sep_nosehoover,
sep_leapfrog] launches one fused kernel per step (pair walk, thermostat and
leapfrog) and finishes each step with the same reductions; the
synchronised run launches the pair, thermostat-sum and leapfrog kernels
separately and moves the state over PCIe around each call.  Both walk the
same lists over the same blocks, and the thermostat's sum of m v^2 is
reduced over the leapfrog's blocks in the leapfrog's order, so until the
first list rebuild the two agree to the bit (positions, velocities,
forces, energies, stress, thermostat).  A rebuild permutes the atoms: the
resident run's thermostat then uses the sum its last leapfrog formed in the
old order, the synchronised run one formed afresh in the new order, and
the two may part in the last bit.  Through rebuilds they agree to 1e-12,
for grids of every shape:

  * the large LJ workload (256000 atoms, 23^3 cells);
  * 32000 atoms on 8^3 cells (62 atoms a cell);
  * 32000 atoms on 3^3 cells (1185 atoms a cell);
  * 1000 atoms on 11^3 cells of a sparse box (8 atoms a row of cells);
  * a small box of 2 cells per axis (the stencil reaches each cell from
    both sides).
"""
import numpy as np
import pytest

from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem

pytestmark = pytest.mark.gpu

CASES = {
    "lj256k": (lambda: W.lj_fcc(40, 0.8442, 1.0, seed=2), 0.3, (23, 23, 23), 20),
    "cells8": (lambda: W.lj_fcc(20, 0.8442, 1.0, seed=5), 0.3, (8, 8, 8), 40),
    "cells3": (lambda: W.lj_fcc(20, 0.8442, 1.0, seed=6), 0.3, (3, 3, 3), 20),
    "sparse": (lambda: W.random_packed(1000, 0.03, 1.5, seed=7), 0.3, (11, 11, 11), 60),
    "two": (lambda: W.lj_sc(8, 0.8442, 1.0, seed=8), 0.5, None, 60),
}


def _run(product, w, skin, nsub, nsteps, resident, precision):
    s = SepSystem.from_workload(product, w, cf=2.5, dt=0.005, skin=skin, nsubbox=nsub)
    try:
        s.gpu_precision(precision)
        if resident:
            s.gpu_resident(True)
            ops = [s.op_pairs("AA", 2.5, "sep_lj_shift"), s.op_nosehoover(1.0, 10.0),
                   s.op_leapfrog()]
            s.gpu_run(ops, nsteps, graph=True)
            s.gpu_sync()
        else:
            for _ in range(nsteps):
                s.reset_retval()
                s.reset_force()
                s.force_pairs("AA", 2.5, "sep_lj_shift")
                s.nosehoover(1.0, 10.0)
                s.leapfrog()
        st = {k: np.array(s.view[k]) for k in ("x", "v", "f", "a")}
        st.update(epot=s.ret.epot, ekin=s.ret.ekin, potP=S.mat(s.ret.pot_P).copy(),
                  alpha=s.alpha.value, rebuilds=s.sys.nupdate_neighb)
        return st
    finally:
        s.close()


PRECISIONS = pytest.mark.parametrize("precision", [S.SEP_GPU_MIXED, S.SEP_GPU_FP64],
                                     ids=["mixed", "fp64"])


@PRECISIONS
def test_resident_step_equals_synchronised_before_rebuild(product, precision):
    make, skin, nsub, _ = CASES["lj256k"]
    w = make()
    a = _run(product, w, skin, nsub, 5, True, precision)
    b = _run(product, w, skin, nsub, 5, False, precision)
    assert a["rebuilds"] == b["rebuilds"] == 1
    for k in ("x", "v", "f", "a"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("epot", "ekin", "alpha"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert np.array_equal(a["potP"], b["potP"])


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@PRECISIONS
@pytest.mark.parametrize("case", list(CASES))
def test_resident_step_follows_synchronised(product, case, precision):
    make, skin, nsub, nsteps = CASES[case]
    w = make()
    print("case %s: %d atoms, cells %s, %d steps" % (case, len(w["m"]), nsub, nsteps),
          flush=True)
    a = _run(product, w, skin, nsub, nsteps, True, precision)
    b = _run(product, w, skin, nsub, nsteps, False, precision)
    assert a["rebuilds"] == b["rebuilds"] and a["rebuilds"] >= 2, (a["rebuilds"], b["rebuilds"])
    L = np.asarray(w["L"], dtype=float)
    assert np.abs(a["x"] - b["x"]).max() <= 1e-12 * L.max()
    for k in ("v", "f", "a"):
        assert _rel(a[k], b[k]) <= 1e-12, (k, _rel(a[k], b[k]))
    for k in ("epot", "ekin", "alpha"):
        assert abs(a[k] - b[k]) <= 1e-12 * max(abs(b[k]), 1.0), (k, a[k], b[k])
    assert _rel(a["potP"], b["potP"]) <= 1e-12
