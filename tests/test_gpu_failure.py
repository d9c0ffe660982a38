"""A device failure inside a resident run: a neighbour-list overflow, forced
from the host with sep_gpu_list_limit (the build truncates each list at the
bound and raises SEP_GPU_ENEIGHB; every access stays inside the allocated
list).

  * the failing graph-replayed run returns SEP_GPU_ENEIGHB, and the system
    stays at the step before the rebuild that overflowed: x, v, f, a, xn,
    crossings, sepret, alpha, tnow and the rebuild counters equal those of
    an unbounded run stopped there, bit for bit (for an odd and an even
    count of frozen steps, which end on either position buffer);
  * with the bound lifted, the next run continues from there and ends bit
    for bit where an uninterrupted run ends.

No kernel faults or waits here: the overflow only sets the device error
word, and every later kernel of the run reads it and leaves the state
alone.
"""
import ctypes as C

import numpy as np
import pytest

from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem

pytestmark = pytest.mark.gpu

KW = dict(cf=2.5, dt=0.005, skin=0.3)


def _ops(s):
    return [s.op_pairs("AA", 2.5), s.op_nosehoover(1.0, 10.0), s.op_leapfrog()]


def _run(s, nsteps, graph=True):
    ops = _ops(s)
    arr = (S.sep_gpu_op * len(ops))(*ops)
    return s._lib().sep_gpu_run(s.atoms, s.sysp, s.retp, arr, len(ops), nsteps,
                                S.SEP_GPU_RUN_GRAPH if graph else 0)


def _snap(s):
    s.gpu_sync()
    st = s.state()
    r = s.ret
    st.update(epot=r.epot, ekin=r.ekin, potP=S.mat(r.pot_P), kinP=S.mat(r.kin_P))
    return st


def _same(a, b):
    for k in ("x", "v", "f", "a", "xn", "cross_neighb", "crossings", "potP", "kinP"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("epot", "ekin", "alpha", "max_dist2", "nupdate_neighb", "neighb_flag", "tnow"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("precision", [S.SEP_GPU_FP64, S.SEP_GPU_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("extra", [10, 11])
def test_list_overflow_freezes_a_graph_run(product, precision, extra):
    w = W.lj_fcc(10, 0.8442, 1.0, seed=4)  # 4000 atoms
    lib = product.lib
    a = SepSystem.from_workload(product, w, **KW)
    b = SepSystem.from_workload(product, w, **KW)
    try:
        for s in (a, b):
            s.gpu_precision(precision)
            s.gpu_resident(True)
            # the first build (checked on the host) and the second, which the
            # reference's zero xn triggers after step 1
            assert _run(s, 2, graph=False) == 0
        mean, mx = C.c_double(0.0), C.c_int(0)
        assert lib.sep_gpu_list_stats(b.atoms, C.byref(mean), C.byref(mx)) == 0
        assert mean.value > 60.0
        # reference: one step at a time until the next step rebuilds
        steps = 0
        while True:
            a.gpu_sync()
            if a.sys.neighb_flag == 1:
                break
            assert _run(a, 1, graph=False) == 0
            steps += 1
            assert steps < 200
        assert steps > 2
        ref = _snap(a)
        assert ref["nupdate_neighb"] >= 1
        # bounded well below the mean list: the next rebuild overflows
        assert lib.sep_gpu_list_limit(b.atoms, b.sysp, 48) == 0
        print("case: precision %d, rebuild after %d steps, %d frozen steps" %
              (precision, steps, extra), flush=True)
        rc = _run(b, steps + extra)
        assert rc == S.SEP_GPU_ENEIGHB, rc
        assert b"overflow" in lib.sep_gpu_last_error_message() or \
            b"neighbours" in lib.sep_gpu_last_error_message()
        _same(_snap(b), ref)
        # lifted: both continue and agree
        assert lib.sep_gpu_list_limit(b.atoms, b.sysp, 0) == 0
        assert _run(b, 30) == 0
        assert _run(a, 30) == 0
        end_a, end_b = _snap(a), _snap(b)
        assert end_a["nupdate_neighb"] > ref["nupdate_neighb"]
        _same(end_b, end_a)
    finally:
        a.close()
        b.close()


def test_list_limit_in_synchronised_mode(product):
    """The synchronised API with a bound below the lists it needs reports
    SEP_GPU_ENEIGHB (the reference's "Too many neighbours") instead of
    growing the list; lifting the bound lets the same call succeed."""
    w = W.lj_fcc(8, 0.8442, 1.0, seed=5)
    lib = product.lib
    s = SepSystem.from_workload(product, w, **KW)
    try:
        assert lib.sep_gpu_list_limit(s.atoms, s.sysp, 32) == 0
        lib.sep_gpu_clear_error()
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", 2.5)
        assert lib.sep_gpu_last_error() == S.SEP_GPU_ENEIGHB
        assert lib.sep_gpu_list_limit(s.atoms, s.sysp, 0) == 0
        lib.sep_gpu_clear_error()
        s.sys.neighb_flag = 1
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", 2.5)
        assert lib.sep_gpu_last_error() == 0
        assert np.abs(np.array(s.view["f"])).max() > 0
    finally:
        s.close()
