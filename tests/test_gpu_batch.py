"""Batches of independent LJ systems (sep_gpu_batch_*), through the C ABI.

  * FP64 forces, energy and stress of every system equal the oracle's
    sep_force_pairs("AA", 2.5, sep_lj_shift) with the O(N^2) list
    (SEP_BRUTE, exact for any box) to the FP64 bar, including boxes with
    two cells per axis;
  * 60 steps of the batch (list rebuilds, Nose-Hoover, leapfrog) follow the
    oracle's step loop per system: positions, velocities, energies and the
    thermostat state to 1e-10;
  * the graph replay equals the eager launches bit for bit;
  * the mixed form agrees with FP64 to 1e-5 (forces, energy);
  * the autotuner's shapes leave the results unchanged (1e-12);
  * invalid batches are refused.
"""
import ctypes as C

import numpy as np
import pytest

from seplib_amd import batch as B
from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem
from tests.parity import FP64_TOL, force_error, tensor_error

pytestmark = pytest.mark.gpu

# (N, rho, T): two boxes with 2 cells per axis, the rest 3-4
CASES = [(200, 0.80, 1.2), (520, 0.95, 0.9), (700, 0.72, 1.9), (1500, 0.85, 1.0),
         (2600, 0.93, 1.5), (1100, 0.70, 0.8)]


def _reps(cases=CASES, seed=3):
    out = []
    for k, (n, rho, T) in enumerate(cases):
        out.append(dict(n=n, rho=rho, T=T, w=W.random_packed(n, rho, T, seed=seed + 10 * k)))
    return out


def _batch(api, reps, precision=S.SEP_GPU_FP64):
    systems = B.make_systems(api, reps)
    return B.Batch(api, systems, [r["T"] for r in reps], precision=precision)


def _oracle_forces(oracle, w):
    s = SepSystem.from_workload(oracle, w, cf=B.RC, dt=B.DT, skin=B.SKIN,
                                update=S.SEP_BRUTE, nneighb=256)
    try:
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", B.RC, "sep_lj_shift")
        return np.array(s.view["f"]), s.ret.epot, S.mat(s.ret.pot_P)
    finally:
        s.close()


def test_batch_eval_matches_oracle(product, oracle):
    reps = _reps()
    b = _batch(product, reps)
    try:
        b.eval(S.SEP_GPU_FP64)
        rets, _ = b.sync()
        for k, r in enumerate(reps):
            f, e, P = _oracle_forces(oracle, r["w"])
            fb = np.array(b.systems[k].view["f"])
            assert force_error(fb, f) <= FP64_TOL, (k, force_error(fb, f))
            assert abs(rets[k].epot - e) <= FP64_TOL * abs(e), (k, rets[k].epot, e)
            assert tensor_error(S.mat(rets[k].pot_P).ravel(), P.ravel()) <= FP64_TOL, k
    finally:
        b.close()
        for s in b.systems:
            s.close()


def _oracle_run(oracle, r, nsteps):
    s = SepSystem.from_workload(oracle, r["w"], cf=B.RC, dt=B.DT, skin=B.SKIN,
                                update=S.SEP_BRUTE, nneighb=256)
    try:
        for _ in range(nsteps):
            s.reset_retval()
            s.reset_force()
            s.force_pairs("AA", B.RC, "sep_lj_shift")
            s.nosehoover(r["T"], B.TAU)
            s.leapfrog()
        return (np.array(s.view["x"]), np.array(s.view["v"]), s.ret.epot, s.ret.ekin,
                S.mat(s.ret.kin_P), s.alpha.value, s.sys.tnow)
    finally:
        s.close()


def test_batch_steps_follow_oracle(product, oracle):
    reps = _reps(CASES[:4])
    nsteps = 60
    b = _batch(product, reps)
    try:
        b.run(nsteps)
        rets, alphas = b.sync()
        for k, r in enumerate(reps):
            x, v, e, ek, kP, al, tnow = _oracle_run(oracle, r, nsteps)
            s = b.systems[k]
            xb, vb = np.array(s.view["x"]), np.array(s.view["v"])
            assert np.abs(xb - x).max() <= 1e-10 * r["w"]["L"][0], k
            assert np.abs(vb - v).max() <= 1e-10 * np.abs(v).max(), k
            assert abs(rets[k].epot - e) <= 1e-10 * abs(e), k
            assert abs(rets[k].ekin - ek) <= 1e-10 * abs(ek), k
            assert tensor_error(S.mat(rets[k].kin_P).ravel(), kP.ravel()) <= 1e-10, k
            assert abs(alphas[k] - al) <= 1e-10 * max(1.0, abs(al)), (k, alphas[k], al)
            assert abs(s.sys.tnow - tnow) <= 1e-12
            assert s.sys.nupdate_neighb >= 1
    finally:
        b.close()
        for s in b.systems:
            s.close()


def test_graph_replay_equals_eager(product):
    reps = _reps(CASES[2:5], seed=9)
    out = []
    for graph in (True, False):
        b = _batch(product, reps, precision=S.SEP_GPU_MIXED)
        try:
            b.run(40, graph=graph)
            rets, al = b.sync()
            out.append(([np.array(s.view["x"]) for s in b.systems],
                        [r.epot for r in rets], al))
        finally:
            b.close()
            for s in b.systems:
                s.close()
    (xa, ea, aa), (xb, eb, ab) = out
    for k in range(len(reps)):
        assert np.array_equal(xa[k], xb[k]), k
    assert ea == eb and aa == ab


def test_mixed_agrees_with_fp64(product):
    reps = _reps(seed=21)
    b = _batch(product, reps, precision=S.SEP_GPU_MIXED)
    try:
        b.run(30)
        for a in B.agreement(b):
            assert a["force"] <= 1e-5 and a["epot"] <= 1e-5 and a["stress"] <= 1e-5, a
    finally:
        b.close()
        for s in b.systems:
            s.close()


def test_tuned_shapes_leave_results_unchanged(product):
    reps = _reps(seed=31)
    b = _batch(product, reps)
    try:
        b.eval(S.SEP_GPU_FP64)
        r0, _ = b.sync()
        f0 = [np.array(s.view["f"]) for s in b.systems]
        lanes, grid = b.tune(reps=3)
        assert all(l in (1, 2, 4, 8, 16) for l in lanes) and grid > 0
        assert b.shape() == (lanes, grid)
        for trial in ([1] * len(reps), [16] * len(reps), lanes):
            b.set_shape(trial, 0)
            b.eval(S.SEP_GPU_FP64)
            r1, _ = b.sync()
            for k, s in enumerate(b.systems):
                assert force_error(np.array(s.view["f"]), f0[k]) <= 1e-12, (trial, k)
                assert abs(r1[k].epot - r0[k].epot) <= 1e-12 * abs(r0[k].epot)
    finally:
        b.close()
        for s in b.systems:
            s.close()


def test_invalid_batches_are_refused(product):
    reps = _reps(CASES[:2])
    systems = B.make_systems(product, reps)
    try:
        systems[1].view["type"][5] = b"B"
        with pytest.raises(RuntimeError, match="type"):
            B.Batch(product, systems, [1.0, 1.0])
        systems[1].view["type"][5] = b"A"
        # a box whose half is below the list cutoff
        small = SepSystem.from_workload(product, W.random_packed(100, 0.9, 1.0, seed=2),
                                        cf=B.RC, dt=B.DT, skin=B.SKIN, nneighb=64)
        try:
            with pytest.raises(RuntimeError, match="too small"):
                B.Batch(product, [systems[0], small], [1.0, 1.0])
        finally:
            small.close()
    finally:
        for s in systems:
            s.close()
