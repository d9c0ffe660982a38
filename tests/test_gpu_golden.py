"""The product (libsep.so on the GPU) against the golden vectors written by
the compiled reference (tests/golden/make_golden.py).

Bars (plain FP64 path): forces max|dF_i| / RMS|F| <= 1e-12 per force term;
energies and the virial tensors <= 1e-12 relative (tensors relative to their
largest component); pair sets bit-exact; rebuild flags and counters exact.
Trajectories (LJ N=1000 NVE) are compared over their first 60 steps, before the
chaotic growth of summation-order differences (~1e-16 per step) can reach
the bar; the state-synchronised tests in test_gpu_lj.py cover long runs.
"""
import os

import numpy as np
import pytest

from seplib_amd import sep as S
from tests.golden import runs as R
from tests.parity import FP64_TOL, force_error, rel, tensor_error

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _check_terms(names, f, r, gf, gr, tol=FP64_TOL):
    for k, name in enumerate(names):
        assert force_error(f[k], gf[k]) <= tol, (name, force_error(f[k], gf[k]))
        # epot, ekin, ecoul, then pot_P, kin_P, pot_P_bond
        for i in range(3):
            # zero in the reference (ekin, ecoul before Coulomb) means zero
            assert abs(r[k][i] - gr[k][i]) <= tol * abs(gr[k][i]), (name, i, r[k][i], gr[k][i])
        for t in range(3):
            a = r[k][3 + 9 * t:12 + 9 * t]
            b = gr[k][3 + 9 * t:12 + 9 * t]
            if np.abs(b).max() > 0:
                assert tensor_error(a, b) <= tol, (name, t)
            else:
                assert np.abs(a).max() == 0.0, (name, t)


@pytest.mark.parametrize("tag,update", [("llist", S.SEP_LLIST_NEIGHBLIST),
                                        ("brute", S.SEP_BRUTE)])
def test_water_terms(product, tag, update):
    """Reference water input (2 cells per axis, forward/backward stencil):
    LJ O-O with same-molecule exclusion, bonds, both angle forms and
    shifted-force Coulomb on the pair call's list."""
    g = _load("g2_water.npz")
    f, r = R.run_water(product, R.workload_from(g), update)
    _check_terms(R.WATER_TERMS, f, r, g["forces_" + tag], g["rets_" + tag])


def test_butane_terms(product):
    g = _load("g3_butane.npz")
    f, r = R.run_butane(product, R.workload_from(g))
    _check_terms(R.BUTANE_TERMS, f, r, g["forces"], g["rets"])


def test_ka_force_lj(product):
    g = _load("g5_ka.npz")
    f, r = R.run_ka(product, R.workload_from(g))
    _check_terms(("AA", "AB", "BB"), f, r, g["forces"], g["rets"])


def test_lj_nve_trajectory(product, oracle):
    g = _load("g1_lj_nve.npz")
    n = 60
    scal, flags, hashes, npairs, forces = R.run_nve(product, oracle.lib, R.workload_from(g),
                                                    n, (0, 1))
    assert np.array_equal(npairs, g["npairs"][:n])
    assert np.array_equal(hashes, g["hashes"][:n])
    assert np.array_equal(flags[:, :2], g["flags"][:n, :2])
    # steps 0 and 1 start from the lattice, where every force is rounding
    # noise (~1e-13): judged in absolute terms, against a unit force scale
    for k in range(2):
        assert np.abs(forces[k] - g["forces"][k]).max() <= FP64_TOL
    gs = g["scal"][:n]
    for c in range(3):
        scale = np.abs(gs[:, c]).max()
        assert np.abs(scal[:, c] - gs[:, c]).max() <= 1e-10 * scale, c
    for t in range(2):
        a, b = scal[:, 3 + 9 * t:12 + 9 * t], gs[:, 3 + 9 * t:12 + 9 * t]
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


def test_lj_nvt_alpha_sequence(product):
    g = _load("g4_lj_nvt.npz")
    n = 60
    out = R.run_nvt(product, R.workload_from(g), n)
    ref = g["alpha_ekin_epot"][:n]
    for c in range(3):
        assert np.abs(out[:, c] - ref[:, c]).max() <= 1e-10 * np.abs(ref[:, c]).max(), c


def test_water_terms_resident(product):
    """The same five calls with the device holding the state (no per-call
    synchronisation); results are read back by the sync at the end."""
    from seplib_amd.sim import SepSystem
    g = _load("g2_water.npz")
    w = R.workload_from(g)
    s = SepSystem.from_workload(product, w, cf=R.WATER["cf"], dt=0.001)
    try:
        s.gpu_resident(True)
        s.reset_retval()
        s.reset_force()
        prev = np.zeros((s.n, 3))
        for k, c in enumerate(R.water_calls()):
            c(s)
            s.gpu_sync()
            f = np.array(s.view["f"])
            assert force_error(f - prev, g["forces_llist"][k]) <= FP64_TOL, R.WATER_TERMS[k]
            prev = f
            rv = R.ret_vec(s)
            assert rel(rv[0], g["rets_llist"][k][0]) <= FP64_TOL
    finally:
        s.close()


@pytest.mark.parametrize("resident", [False, True], ids=["sync", "resident"])
@pytest.mark.parametrize("case", R.G6_CASES, ids=[c[0] for c in R.G6_CASES])
def test_exclusions_and_list_modes(product, oracle, case, resident):
    """G6: SEP_EXCL_BONDED through the linked-cell list (bond, angle and
    dihedral partners) and the O(N^2) list and SEP_BRUTE (bonds only),
    SEP_EXCL_SAME_MOL and SEP_ALL through the O(N^2) list, and brute-force
    Coulomb with and without the bonded rule: the pair multiset bit-exact
    (host half lists when synchronised, the device list when resident),
    forces and sepret fields within 1e-12."""
    from seplib_amd.sim import SepSystem
    tag, system, update, call = case
    g = _load("g6_excl.npz")
    w = R.g6_workload(g, system)
    s = SepSystem.from_workload(product, w, cf=R.g6_cf(system), dt=0.001, update=update)
    try:
        if resident:
            s.gpu_resident(True)
        s.reset_retval()
        s.reset_force()
        call(s)
        if resident:
            s.gpu_sync()
        f = np.array(s.view["f"])
        r = R.ret_vec(s)
        if update != S.SEP_BRUTE:
            if resident:
                p = s.gpu_pairs()
                keys = np.sort(p[:, 0] * s.n + p[:, 1])
            else:
                keys = R.pair_keys(oracle.lib, s, cap_per_atom=2048)
            assert [len(keys), R.pair_hash(keys)] == list(g["pairs_" + tag])
    finally:
        s.close()
    _check_terms((tag,), f[None], r[None], g["f_" + tag][None], g["r_" + tag][None])
