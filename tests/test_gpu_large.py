"""GPU parity at the large LJ size: N = 256000 atoms on a 40^3 fcc lattice
(rho 0.8442), jittered off the lattice, shifted LJ rc = 2.5, skin 0.3,
23^3 cells (edge 2.921 >= rc + skin).

This is the size where the list build takes its per-cell image-shift path
(>= 5 cells per axis, every atom inside its own cell) and the pair kernels
run at full occupancy, so these tests pin those paths against the C oracle:
  * neighbour pair set: bit-exact, from the host lists and from the device;
  * FP64 forces max|dF_i| / RMS|F| <= 1e-12, energy and virial <= 1e-12;
  * the same after Nose-Hoover + leapfrog steps with a rebuild among them;
  * an atom outside [0, L) (binned by truncation, as the reference does)
    switches the build to the per-candidate wrap and keeps the set exact;
  * mixed form (FP32 pair arithmetic, FP64 sums): force error <= 1e-5.
"""
import numpy as np
import pytest

from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem
from tests.parity import FP64_TOL, force_error, half_pair_keys, pair_keys, rel, tensor_error
from tests.test_gpu_lj import _check_after_leapfrog, _check_forces, _pair

pytestmark = pytest.mark.gpu

NNEIGHB = 256  # host half-list room per atom (about 39 are used)
NSUB = (23, 23, 23)


def _workload(jitter=0.2, seed=11):
    w = W.lj_fcc(40, 0.8442, 1.0, seed=2)
    n = len(w["m"])
    d = (W.uniform(seed, 3 * n).reshape(n, 3) - 0.5) * jitter
    w["x"] = np.mod(w["x"] + d, w["L"])
    return w


def _systems(product, oracle, w):
    kw = dict(cf=2.5, dt=0.005, skin=0.3, nsubbox=NSUB, nneighb=NNEIGHB)
    return (SepSystem.from_workload(product, w, **kw),
            SepSystem.from_workload(oracle, w, **kw))


def _check_pairs(ga, orc, oracle):
    ko = half_pair_keys(oracle.lib, orc)
    assert np.array_equal(half_pair_keys(oracle.lib, ga), ko)
    assert np.array_equal(pair_keys(ga.gpu_pairs(), ga.n), ko)
    return len(ko)


def test_lj_256k_pairs_forces_steps(product, oracle):
    w = _workload()
    ga, orc = _systems(product, oracle, w)
    try:
        _pair(ga, orc, types="AA", cf=2.5)
        npairs = _check_pairs(ga, orc, oracle)
        assert 36 * ga.n < npairs < 42 * ga.n
        _check_forces(ga, orc)
        rebuilt = 0
        for step in range(8):
            if step:
                flag = orc.sys.neighb_flag
                _pair(ga, orc, types="AA", cf=2.5)
                if flag:
                    rebuilt += 1
                    _check_pairs(ga, orc, oracle)
                _check_forces(ga, orc)
            for s in (ga, orc):
                s.nosehoover(1.0, 10.0)
            assert rel(ga.alpha.value, orc.alpha.value, 1e-3) <= FP64_TOL
            for s in (ga, orc):
                s.leapfrog()
            _check_after_leapfrog(ga, orc)
        assert rebuilt >= 1
    finally:
        ga.close()
        orc.close()


def test_lj_256k_atom_outside_box(product, oracle):
    """x < 0 on some axis: the reference bins it by truncation into cell 0
    (x >= L has no cell there), so the build must evaluate the image per
    candidate."""
    w = _workload(seed=12)
    w["x"][1234, 0] = -0.01
    w["x"][4321, 2] = -0.3
    ga, orc = _systems(product, oracle, w)
    try:
        _pair(ga, orc, types="AA", cf=2.5)
        _check_pairs(ga, orc, oracle)
        _check_forces(ga, orc)
    finally:
        ga.close()
        orc.close()


def test_lj_256k_mixed_precision(product, oracle):
    w = _workload(seed=13)
    ga, orc = _systems(product, oracle, w)
    try:
        ga.gpu_precision(S.SEP_GPU_MIXED)
        _pair(ga, orc, types="AA", cf=2.5)
        ef = force_error(np.array(ga.view["f"]), np.array(orc.view["f"]))
        assert ef <= 1e-5, ef
        a, b = ga.ret_dict(), orc.ret_dict()
        assert rel(a["epot"], b["epot"]) <= 1e-6
        assert tensor_error(a["pot_P"], b["pot_P"]) <= 1e-6
    finally:
        ga.close()
        orc.close()


def test_position_beyond_grid_is_reported(product):
    """z >= L maps past the reference's last cell (its array would be read
    out of bounds; x or y >= L only alias into a neighbouring row or layer
    of its linear cell index): the library records SEP_GPU_ECELL instead."""
    w = _workload(seed=14)
    w["x"][777, 2] = w["L"][2] + 0.01
    kw = dict(cf=2.5, dt=0.005, skin=0.3, nsubbox=NSUB, nneighb=NNEIGHB)
    ga = SepSystem.from_workload(product, w, **kw)
    try:
        lib = ga._lib()
        lib.sep_gpu_clear_error()
        ga.reset_retval()
        ga.reset_force()
        ga.force_pairs("AA", 2.5)
        assert lib.sep_gpu_last_error() == S.SEP_GPU_ECELL
        lib.sep_gpu_clear_error()
    finally:
        ga.close()


def test_lj_1m_wide_list_slots(product, oracle):
    """N = 1048576 (64^3 fcc cells): npad >= 2^20, so the list build writes
    its entries through the 64-bit slot arithmetic instead of the 32-bit one
    (24-bit multiply) that every smaller system takes.  The pair set is
    bit-exact against the oracle, FP64 forces, energy and virial within
    1e-12, and the mixed form within 1e-5 of them."""
    w = W.lj_fcc(64, 0.8442, 1.0, seed=3)
    n = len(w["m"])
    d = (W.uniform(15, 3 * n).reshape(n, 3) - 0.5) * 0.2
    w["x"] = np.mod(w["x"] + d, w["L"])
    assert n == 1 << 20
    # host half-list room: a half-shell list gives some atoms up to the
    # full-list length (about 80 here)
    kw = dict(cf=2.5, dt=0.005, skin=0.3, nsubbox=(38, 38, 38), nneighb=128)
    ga = SepSystem.from_workload(product, w, **kw)
    orc = SepSystem.from_workload(oracle, w, **kw)
    try:
        _pair(ga, orc, types="AA", cf=2.5)
        npairs = _check_pairs(ga, orc, oracle)
        assert 36 * n < npairs < 42 * n
        _check_forces(ga, orc)
        f64 = np.array(ga.view["f"])
        ga.gpu_precision(S.SEP_GPU_MIXED)
        ga.reset_retval()
        ga.reset_force()
        ga.force_pairs("AA", 2.5)
        assert force_error(np.array(ga.view["f"]), f64) <= 1e-5
    finally:
        ga.close()
        orc.close()


def test_dense_cells_workgroup_sort(product, oracle):
    """30000 atoms at random in a box of 3^3 cells, about 1100 atoms a cell:
    the cell sort takes a workgroup per cell and ranks the members against
    the cell's keys in LDS, more than one block of keys (1024) a cell.  A
    short cutoff (rc 0.7, skin 0.3, cells of edge 3) keeps the lists near
    170 entries.  The pair set is bit-exact against the oracle's before and
    after a step that moves the atoms and rebuilds the list."""
    n = 30000
    L = 9.0
    x = W.uniform(21, 3 * n).reshape(n, 3) * L
    w = dict(x=x, v=np.zeros((n, 3)), m=np.ones(n), q=np.zeros(n),
             type=np.full(n, b"A"), L=np.array([L, L, L]), mol=np.full(n, -1))
    kw = dict(cf=0.7, dt=1e-6, skin=0.3, nsubbox=(3, 3, 3), nneighb=320)
    ga = SepSystem.from_workload(product, w, **kw)
    orc = SepSystem.from_workload(oracle, w, **kw)
    try:
        _pair(ga, orc, types="AA", cf=0.7)
        npairs = _check_pairs(ga, orc, oracle)
        assert 80 * n < npairs < 95 * n
        # move every atom by a fixed random step and rebuild on both sides
        d = (W.uniform(22, 3 * n).reshape(n, 3) - 0.5) * 0.4
        w2 = dict(w, x=np.mod(x + d, L))
        for s in (ga, orc):
            s.load(w2)
            s.sys.neighb_flag = 1
        _pair(ga, orc, types="AA", cf=0.7)
        _check_pairs(ga, orc, oracle)
    finally:
        ga.close()
        orc.close()
