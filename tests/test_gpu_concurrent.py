"""Concurrent force terms (SEP_GPU_RUN_CONCURRENT): the per-term kernels of
a step's bonded calls on a second stream beside its pair kernels, the GPU
form of the reference's OpenMP sections over independent terms
(source/sepomp.c, prgs/prg5.c:58-77).

Each atom still adds its terms in call order, so the two-stream run must
equal the one-stream run bit for bit; the one-stream run is held against
the oracle by test_gpu_water.py and test_gpu_golden.py.

  * water, 4000 molecules: 30 resident steps of LJ, Coulomb, bond, angle,
    Nose-Hoover and leapfrog, eager and replayed from a graph, in two call
    orders (bonded calls after both pair calls, and between them);
  * the reference's butane input (prgs/prg2.c terms): LJ, bond, harmonic
    angle, cosine-squared angle and Ryckaert-Bellemans torsion, where the
    second angle call reuses the angle family's per-term slots and so must
    wait for the first one's gather;
  * a step with no pair call before its bonded calls runs on one stream;
  * the water step's pair and Coulomb calls in one list walk (a
    sep_gpu_run step) against the same calls made one at a time on the
    resident system: the same bits after 20 steps, FP64 and mixed; and
    with SEP_BRUTE, where the Coulomb call builds its own list;
  * the Model II entry points themselves (sep_omp_make_neighb and the
    sep_omp_* terms into caller-owned arrays) against the oracle to 1e-12
    on water and butane, leaving the atoms' forces alone, and refused on a
    resident system.
"""
import os

import numpy as np
import pytest

from seplib_amd import sep as S
from seplib_amd import water as WAT
from seplib_amd.sim import SepSystem
from tests.golden import runs as R

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _final(s):
    s.gpu_sync()
    st = s.state()
    r = s.ret
    return dict(x=st["x"], v=st["v"], f=st["f"], xn=st["xn"], crossings=st["crossings"],
                ret=R.ret_vec(s), alpha=s.alpha.value, tnow=st["tnow"],
                nupdate=st["nupdate_neighb"], epot=r.epot, ecoul=r.ecoul, ekin=r.ekin)


def _run(api, w, make_ops, steps, graph, concurrent, cf, dt, nneighb):
    s = SepSystem.from_workload(api, w, cf=cf, dt=dt, nneighb=nneighb)
    try:
        s.gpu_resident(True)
        s.gpu_run(make_ops(s), steps, graph=graph, concurrent=concurrent)
        return _final(s)
    finally:
        s.close()


def _assert_same(a, b):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.fixture(scope="module")
def water():
    return WAT.system(4000)


def _water_ops_pairs_first(s):
    return WAT.step_ops(s)


def _water_ops_interleaved(s):
    o = WAT.step_ops(s)
    return [o[0], o[2], o[3], o[1], o[4], o[5]]


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("make_ops", [_water_ops_pairs_first, _water_ops_interleaved],
                         ids=["pairs_first", "interleaved"])
def test_water_two_streams_bit_identical(product, water, make_ops, graph):
    kw = dict(cf=WAT.RC_COUL, dt=WAT.DT, nneighb=WAT.NNEIGHB)
    one = _run(product, water, make_ops, 30, graph, False, **kw)
    two = _run(product, water, make_ops, 30, graph, True, **kw)
    assert one["nupdate"] > 0  # the run crossed list rebuilds
    _assert_same(one, two)


def _butane_ops(s):
    p = R.BUTANE
    return [s.op_pairs("CC", 2.5, "sep_lj_shift", S.SEP_EXCL_SAME_MOL),
            s.op_bonded(S.OP_BOND, 0, [p["lbond"], p["kspring"]]),
            s.op_bonded(S.OP_ANGLE, 0, [p["angle"], p["kangle"]]),
            s.op_bonded(S.OP_COSSQ, 0, [p["angle"], p["kangle"]]),
            s.op_bonded(S.OP_TORSION, 0, p["rb"]),
            s.op_nosehoover(4.0, 0.1),
            s.op_leapfrog()]


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_butane_two_streams_bit_identical(product, graph):
    w = R.workload_from(np.load(os.path.join(GOLD, "g3_butane.npz"), allow_pickle=False))
    kw = dict(cf=2.5, dt=0.001, nneighb=S.SEP_NEIGHB)
    one = _run(product, w, _butane_ops, 40, graph, False, **kw)
    two = _run(product, w, _butane_ops, 40, graph, True, **kw)
    _assert_same(one, two)


def test_bonded_before_pairs_stays_on_one_stream(product, water):
    def ops(s):
        o = WAT.step_ops(s)
        return [o[2], o[3], o[0], o[1], o[4], o[5]]
    kw = dict(cf=WAT.RC_COUL, dt=WAT.DT, nneighb=WAT.NNEIGHB)
    _assert_same(_run(product, water, ops, 10, True, False, **kw),
                 _run(product, water, ops, 10, True, True, **kw))


# ---- the Model II entry points themselves (sep_omp_*, include/sepomp.h) ----------

def _model2_case(name):
    if name == "water":
        return WAT.system(1500), R.WATER["cf"], R.model2_water
    w = R.workload_from(np.load(os.path.join(GOLD, "g3_butane.npz"), allow_pickle=False))
    return w, 2.5, R.model2_butane


@pytest.mark.parametrize("name", ["water", "butane"])
def test_model2_terms_match_oracle(product, oracle, name):
    """sep_omp_make_neighb, then sep_omp_pairs, _pairs_lj, _coulomb, _bond,
    _angle and _torsion into caller arrays: against the oracle's
    restatement (pinned bitwise to the compiled reference) to 1e-12, and
    the atoms' own forces left as they were."""
    from tests.parity import FP64_TOL, force_error
    w, cf, terms = _model2_case(name)
    a, f0, f1 = R.run_model2(product, w, cf, S.SEP_EXCL_SAME_MOL, terms)
    b, _, _ = R.run_model2(oracle, w, cf, S.SEP_EXCL_SAME_MOL, terms)
    assert np.array_equal(f0, f1)
    for k in range(len(a)):
        assert np.abs(b[k]).max() > 0, k
        assert force_error(a[k], b[k]) <= FP64_TOL, (name, k, force_error(a[k], b[k]))


def test_model2_refused_on_resident_system(product):
    w, cf, terms = _model2_case("butane")
    s = SepSystem.from_workload(product, w, cf=cf, dt=0.001, nneighb=S.SEP_NEIGHB)
    try:
        s.gpu_resident(True)
        product.lib.sep_gpu_clear_error()
        f = s.omp_term("bond", 0, R.BUTANE["lbond"], R.BUTANE["kspring"])
        assert product.lib.sep_gpu_last_error() == S.SEP_GPU_ESTATE
        assert not f.any()
        product.lib.sep_gpu_clear_error()
    finally:
        s.close()


@pytest.mark.parametrize("precision", [S.SEP_GPU_FP64, S.SEP_GPU_MIXED], ids=["fp64", "mixed"])
def test_water_pair_coulomb_walk_bit_identical(product, water, precision):
    kw = dict(cf=WAT.RC_COUL, dt=WAT.DT, nneighb=WAT.NNEIGHB)
    out = []
    for fused in (True, False):
        s = SepSystem.from_workload(product, water, **kw)
        try:
            s.gpu_precision(precision)
            s.gpu_resident(True)
            if fused:
                s.gpu_run(WAT.step_ops(s), 20, graph=True)
            else:
                for _ in range(20):
                    s.reset_retval()
                    s.reset_force()
                    s.force_pairs("OO", WAT.RC_LJ, "sep_lj_shift", S.SEP_EXCL_SAME_MOL)
                    s.coulomb_sf(WAT.RC_COUL, S.SEP_EXCL_SAME_MOL)
                    s.stretch_harmonic(0, WAT.LBOND, WAT.KBOND)
                    s.angle_harmonic(0, WAT.THETA0, WAT.KANGLE)
                    s.nosehoover(WAT.TEMP, WAT.TAU)
                    s.leapfrog()
            out.append(_final(s))
        finally:
            s.close()
    assert out[0]["nupdate"] > 0
    _assert_same(out[0], out[1])


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_water_brute_pair_then_coulomb_order(product, water, graph):
    """SEP_BRUTE: each pair call builds its own list, so a recorded pair call
    has to run before the Coulomb call's build replaces the list.  A
    sep_gpu_run step equals the same calls made one at a time.  The two
    calls build at different cutoffs (two grids in one step), which a graph
    capture must hold without a host check between them."""
    kw = dict(cf=WAT.RC_COUL, dt=WAT.DT, nneighb=WAT.NNEIGHB, update=S.SEP_BRUTE)
    out = []
    for run in (True, False):
        s = SepSystem.from_workload(product, water, **kw)
        try:
            s.gpu_resident(True)
            if run:
                s.gpu_run(WAT.step_ops(s), 5, graph=graph)
            else:
                for _ in range(5):
                    s.reset_retval()
                    s.reset_force()
                    s.force_pairs("OO", WAT.RC_LJ, "sep_lj_shift", S.SEP_EXCL_SAME_MOL)
                    s.coulomb_sf(WAT.RC_COUL, S.SEP_EXCL_SAME_MOL)
                    s.stretch_harmonic(0, WAT.LBOND, WAT.KBOND)
                    s.angle_harmonic(0, WAT.THETA0, WAT.KANGLE)
                    s.nosehoover(WAT.TEMP, WAT.TAU)
                    s.leapfrog()
            out.append(_final(s))
        finally:
            s.close()
    assert np.abs(out[0]["f"]).max() > 0
    _assert_same(out[0], out[1])


def _water_brute_equal_cutoffs(s):
    """LJ and Coulomb at the same cutoff, then the bonded terms: the
    Coulomb call's brute-force build reuses the pair call's grid, and the
    pair call it flushes records a fork point before that build."""
    return [s.op_pairs("OO", WAT.RC_COUL, "sep_lj_shift", S.SEP_EXCL_SAME_MOL),
            s.op_coulomb(WAT.RC_COUL, S.SEP_EXCL_SAME_MOL),
            s.op_bonded(S.OP_BOND, 0, [WAT.LBOND, WAT.KBOND]),
            s.op_bonded(S.OP_ANGLE, 0, [WAT.THETA0, WAT.KANGLE]),
            s.op_nosehoover(WAT.TEMP, WAT.TAU),
            s.op_leapfrog()]


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_water_brute_equal_cutoffs_two_streams(product, water, graph):
    """SEP_BRUTE water with the pair and Coulomb calls at one cutoff and the
    bonded calls after them on the second stream: the bonded kernels must
    wait for the Coulomb call's rebuild, which permutes the atoms, so the
    two-stream run equals the one-stream run bit for bit."""
    kw = dict(cf=WAT.RC_COUL, dt=WAT.DT, nneighb=WAT.NNEIGHB)

    def run(concurrent):
        s = SepSystem.from_workload(product, water, update=S.SEP_BRUTE, **kw)
        try:
            s.gpu_resident(True)
            s.gpu_run(_water_brute_equal_cutoffs(s), 6, graph=graph, concurrent=concurrent)
            return _final(s)
        finally:
            s.close()

    one, two = run(False), run(True)
    assert np.abs(one["f"]).max() > 0
    _assert_same(one, two)
