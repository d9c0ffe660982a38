"""The oracle against the compiled reference (oracle/_ref/libsepref.so,
built from the reference's sources by oracle/Makefile), run live on inputs
beyond the golden fixtures.  CPU only; skipped where the reference sources
were not available to build it."""
import numpy as np

from seplib_amd import sep as S
from seplib_amd import workloads as W
from tests.golden import runs as R


def test_lj_nve_other_seed_bitwise(reference, oracle):
    w = W.lj_sc(9, 0.80, 1.1, seed=21)
    a = R.run_nve(reference, oracle.lib, w, 300, (0, 299))
    b = R.run_nve(oracle, oracle.lib, w, 300, (0, 299))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_lj_nvt_other_seed_bitwise(reference, oracle):
    w = W.lj_fcc(5, 0.9, 1.3, seed=22)
    assert np.array_equal(R.run_nvt(reference, w, 150), R.run_nvt(oracle, w, 150))


def test_ka_other_seed_bitwise(reference, oracle):
    w = W.ka_mixture(864, 1.2, 0.8, seed=23)
    for x, y in zip(R.run_ka(reference, w), R.run_ka(oracle, w)):
        assert np.array_equal(x, y)


def test_water_generated_bitwise(reference, oracle):
    """A generated water box (4 cells per axis: the 27-cell stencil) with the
    linked-list and brute-force paths."""
    w = W.water(400, 3.16, 3.86, seed=24)
    for update in (S.SEP_LLIST_NEIGHBLIST, S.SEP_BRUTE):
        for x, y in zip(R.run_water(reference, w, update), R.run_water(oracle, w, update)):
            assert np.array_equal(x, y)


def _butane():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "g3_butane.npz"),
                allow_pickle=False)
    return R.workload_from(g)


def test_model2_terms_bitwise(reference, oracle):
    """The OpenMP Model II terms (source/sepomp.c) after sep_omp_make_neighb:
    a generated water box and the reference's butane input."""
    cases = [(W.water(400, 3.16, 3.86, seed=24), R.WATER["cf"], S.SEP_EXCL_SAME_MOL,
              R.model2_water),
             (_butane(), 2.5, S.SEP_EXCL_SAME_MOL, R.model2_butane)]
    for w, cf, opt, terms in cases:
        a, _, fa = R.run_model2(reference, w, cf, opt, terms)
        b, _, fb = R.run_model2(oracle, w, cf, opt, terms)
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)
        assert np.array_equal(fa, fb)
