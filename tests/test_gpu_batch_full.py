"""The batch workload at its stated size: the 64 replicas of
batch.replicas(64, 11) -- atom counts log-uniform in [500, 32000] (554 to
31529 here, 365498 atoms in all), densities uniform in [0.70, 0.95],
temperatures uniform in [0.8, 2.0], shifted LJ rc 2.5, skin 0.3, dt 0.004,
Nose-Hoover tau 0.2 -- advanced as one batch.

  * one FP64 evaluation: forces, energy and virial of every replica equal
    the oracle's sep_force_pairs("AA", 2.5, sep_lj_shift) over its own
    linked-cell list (prgs/prg1.c's call, SEP_LLIST_NEIGHBLIST) to the FP64
    bar;
  * 60 FP64 steps of the whole batch: the 4 largest replicas follow the
    oracle's step loop (sep_force_pairs, sep_nosehoover, sep_leapfrog on
    the linked-cell list) to 1e-10 -- positions, velocities, energies,
    kinetic stress and the thermostat state;
  * the mixed form, after 30 steps of the batch, agrees with FP64 on every
    replica to 1e-5 (forces max|dF|/RMS|F|, energy, stress).
"""
import numpy as np
import pytest

from seplib_amd import batch as B
from seplib_amd import sep as S
from seplib_amd.sim import SepSystem
from tests.parity import FP64_TOL, force_error, tensor_error

pytestmark = pytest.mark.gpu

SEED = 11
NNEIGHB = 256


@pytest.fixture(scope="module")
def reps():
    r = B.replicas(64, SEED)
    # every box has at least 3 cells per axis, so the oracle's linked-cell
    # list holds each pair once, as the batch's does
    assert all(min(B.cells_for(x["w"]["L"], B.RC + B.SKIN)) >= 3 for x in r)
    return r


def _oracle_system(oracle, r):
    w = r["w"]
    return SepSystem.from_workload(oracle, w, cf=B.RC, dt=B.DT, skin=B.SKIN, nneighb=NNEIGHB,
                                   nsubbox=B.cells_for(w["L"], B.RC + B.SKIN))


def _close(b):
    b.close()
    for s in b.systems:
        s.close()


def test_full_batch_eval_matches_oracle(product, oracle, reps):
    systems = B.make_systems(product, reps)
    b = B.Batch(product, systems, [r["T"] for r in reps], precision=S.SEP_GPU_FP64)
    try:
        b.eval(S.SEP_GPU_FP64)
        rets, _ = b.sync()
        for k, r in enumerate(reps):
            s = _oracle_system(oracle, r)
            try:
                s.reset_retval()
                s.reset_force()
                s.force_pairs("AA", B.RC, "sep_lj_shift")
                f, e, P = np.array(s.view["f"]), s.ret.epot, S.mat(s.ret.pot_P)
            finally:
                s.close()
            fb = np.array(b.systems[k].view["f"])
            assert force_error(fb, f) <= FP64_TOL, (k, r["n"], force_error(fb, f))
            assert abs(rets[k].epot - e) <= FP64_TOL * abs(e), (k, rets[k].epot, e)
            assert tensor_error(S.mat(rets[k].pot_P).ravel(), P.ravel()) <= FP64_TOL, k
    finally:
        _close(b)


def test_full_batch_steps_follow_oracle_on_largest(product, oracle, reps):
    nsteps = 60
    systems = B.make_systems(product, reps)
    b = B.Batch(product, systems, [r["T"] for r in reps], precision=S.SEP_GPU_FP64)
    try:
        b.run(nsteps)
        rets, alphas = b.sync()
        largest = sorted(range(len(reps)), key=lambda k: reps[k]["n"])[-4:]
        for k in largest:
            r = reps[k]
            s = _oracle_system(oracle, r)
            try:
                for _ in range(nsteps):
                    s.reset_retval()
                    s.reset_force()
                    s.force_pairs("AA", B.RC, "sep_lj_shift")
                    s.nosehoover(r["T"], B.TAU)
                    s.leapfrog()
                x, v = np.array(s.view["x"]), np.array(s.view["v"])
                e, ek, kP = s.ret.epot, s.ret.ekin, S.mat(s.ret.kin_P)
                al, tnow = s.alpha.value, s.sys.tnow
            finally:
                s.close()
            sb = b.systems[k]
            xb, vb = np.array(sb.view["x"]), np.array(sb.view["v"])
            assert np.abs(xb - x).max() <= 1e-10 * r["w"]["L"][0], (k, r["n"])
            assert np.abs(vb - v).max() <= 1e-10 * np.abs(v).max(), (k, r["n"])
            assert abs(rets[k].epot - e) <= 1e-10 * abs(e), k
            assert abs(rets[k].ekin - ek) <= 1e-10 * abs(ek), k
            assert tensor_error(S.mat(rets[k].kin_P).ravel(), kP.ravel()) <= 1e-10, k
            assert abs(alphas[k] - al) <= 1e-10 * max(1.0, abs(al)), (k, alphas[k], al)
            assert abs(sb.sys.tnow - tnow) <= 1e-12
            assert sb.sys.nupdate_neighb >= 1
    finally:
        _close(b)


def test_full_batch_mixed_agrees_with_fp64(product, reps):
    systems = B.make_systems(product, reps)
    b = B.Batch(product, systems, [r["T"] for r in reps], precision=S.SEP_GPU_MIXED)
    try:
        b.tune()
        b.run(30)
        res = B.agreement(b)
        assert len(res) == 64
        for k, a in enumerate(res):
            assert a["force"] <= 1e-5 and a["epot"] <= 1e-5 and a["stress"] <= 1e-5, (k, a)
    finally:
        _close(b)
