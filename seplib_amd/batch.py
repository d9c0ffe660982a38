"""Batches of independent Lennard-Jones systems on one GPU.

A batch advances many seppart/sepsys pairs together (sep_gpu_batch_* in
include/sepgpu.h): each system runs the reference's step loop
(prgs/prg1.c:55-65: sep_force_pairs "AA" with sep_lj_shift, sep_nosehoover,
sep_leapfrog) and all of them share one persistent launch per step.

    python -m seplib_amd.batch [--systems 64] [--steps 5000] [--seed 11]

runs the replica workload -- atom counts log-uniform in [500, 32000],
densities uniform in [0.70, 0.95], temperatures uniform in [0.8, 2.0], each
a random-packed cubic box (workloads.random_packed), shifted LJ rc 2.5,
skin 0.3, dt 0.004, Nose-Hoover tau 0.2, FP32 pair arithmetic with FP64
accumulation -- after the autotuner has picked the tile shape per size
bucket, and prints one JSON line: aggregate atom-steps/s of the batch, the
same replicas launched one by one through the single-system path, and the
per-replica energy and force agreement of the mixed form with FP64 on the
final state.
"""
import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np

from . import sep as S
from . import workloads as W
from .sim import ProductAPI, SepSystem

RC = 2.5
SKIN = 0.3
DT = 0.004
TAU = 0.2
NNEIGHB = 128


def replicas(nsys=64, seed=11, nmin=500, nmax=32000, rho=(0.70, 0.95), temp=(0.8, 2.0)):
    """Seeded replica parameters and their starting configurations."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nsys):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        r = float(rng.uniform(*rho))
        T = float(rng.uniform(*temp))
        w = W.random_packed(n, r, T, seed=int(rng.integers(1, 2 ** 30)))
        out.append(dict(n=n, rho=r, T=T, w=w))
    return out


def cells_for(L, rlist):
    return tuple(max(1, int(L[k] / rlist)) for k in range(3))


def make_systems(api, reps, nneighb=NNEIGHB):
    """One SepSystem per replica (host state the batch is created from)."""
    out = []
    for r in reps:
        w = r["w"]
        s = SepSystem.from_workload(api, w, cf=RC, dt=DT, skin=SKIN, nneighb=nneighb,
                                    nsubbox=cells_for(w["L"], RC + SKIN))
        out.append(s)
    return out


class Batch:
    """A sep_gpu_batch over the given SepSystems."""

    def __init__(self, api, systems, temps, precision=S.SEP_GPU_MIXED, cap=0, tau=TAU):
        self.api = api
        self.lib = api.lib
        self.systems = systems
        ns = len(systems)
        self.ns = ns
        self._atoms = (S.P_ATOMS * ns)(*[s.atoms for s in systems])
        self._sys = (S.P_SYS * ns)(*[s.sysp for s in systems])
        Td = (C.c_double * ns)(*temps)
        ta = (C.c_double * ns)(*([tau] * ns))
        al = (C.c_double * ns)(*([0.0] * ns))
        h = self.lib.sep_gpu_batch_create(ns, self._atoms, self._sys, b"AA", RC, Td, ta, al,
                                          precision, cap)
        if not h:
            raise RuntimeError("sep_gpu_batch_create: %s"
                               % self.lib.sep_gpu_last_error_message().decode())
        self.h = C.c_void_p(h)

    def check(self, rc):
        if rc:
            raise RuntimeError("sepgpu error %d: %s"
                               % (rc, self.lib.sep_gpu_last_error_message().decode()))

    def run(self, nsteps, graph=True):
        fn = self.lib.sep_gpu_batch_run if graph else self.lib.sep_gpu_batch_run_eager
        self.check(fn(self.h, nsteps))

    def eval(self, precision):
        self.check(self.lib.sep_gpu_batch_eval(self.h, precision))

    def synchronize(self):
        self.check(self.lib.sep_gpu_batch_synchronize(self.h))

    def sync(self):
        """Host state of every system updated; returns (rets, alphas)."""
        rets = (S.sepret * self.ns)()
        al = (C.c_double * self.ns)()
        self.check(self.lib.sep_gpu_batch_sync(self.h, rets, al))
        return rets, list(al)

    def tune(self, reps=20):
        lanes = (C.c_int * self.ns)()
        grid = C.c_int(0)
        self.check(self.lib.sep_gpu_batch_tune(self.h, reps, lanes, C.byref(grid)))
        return list(lanes), grid.value

    def shape(self):
        lanes = (C.c_int * self.ns)()
        grid = C.c_int(0)
        self.check(self.lib.sep_gpu_batch_get_shape(self.h, lanes, C.byref(grid)))
        return list(lanes), grid.value

    def set_shape(self, lanes=None, grid=0):
        arr = (C.c_int * self.ns)(*lanes) if lanes is not None else None
        self.check(self.lib.sep_gpu_batch_set_shape(self.h, arr, grid))

    def close(self):
        if self.h:
            self.lib.sep_gpu_batch_free(self.h)
            self.h = None


def agreement(batch):
    """Mixed vs FP64 on the current state of every system: relative epot
    difference, force error max|dF|/RMS|F| and stress error per system."""
    out = {}
    for prec in (S.SEP_GPU_FP64, S.SEP_GPU_MIXED):
        batch.eval(prec)
        rets, _ = batch.sync()
        out[prec] = ([np.array(s.view["f"]) for s in batch.systems],
                     [(r.epot, S.mat(r.pot_P)) for r in rets])
    res = []
    for k in range(batch.ns):
        f64, f32 = out[S.SEP_GPU_FP64][0][k], out[S.SEP_GPU_MIXED][0][k]
        (e64, p64), (e32, p32) = out[S.SEP_GPU_FP64][1][k], out[S.SEP_GPU_MIXED][1][k]
        rms = np.sqrt((f64 ** 2).sum(axis=1).mean())
        res.append(dict(epot=abs(e32 - e64) / abs(e64),
                        force=float(np.sqrt(((f32 - f64) ** 2).sum(axis=1)).max() / rms),
                        stress=float(np.abs(p32 - p64).max() / np.abs(p64).max())))
    return res


def one_by_one(api, reps, steps, precision=S.SEP_GPU_MIXED):
    """The same replicas through the single-system path, one after another
    (resident, each step replayed from its own graph): seconds and
    atom-steps."""
    total, atom_steps = 0.0, 0
    for r in reps:
        w = r["w"]
        s = SepSystem.from_workload(api, w, cf=RC, dt=DT, skin=SKIN, nneighb=NNEIGHB,
                                    nsubbox=cells_for(w["L"], RC + SKIN))
        s.gpu_precision(precision)
        s.gpu_resident(True)
        ops = [s.op_pairs("AA", RC, "sep_lj_shift"), s.op_nosehoover(r["T"], TAU),
               s.op_leapfrog()]
        s.gpu_run(ops, 20, graph=True)
        s.check(api.lib.sep_gpu_synchronize(s.atoms))
        t0 = time.perf_counter()
        s.gpu_run(ops, steps, graph=True)
        s.check(api.lib.sep_gpu_synchronize(s.atoms))
        total += time.perf_counter() - t0
        atom_steps += len(w["m"]) * steps
        s.close()
    return total, atom_steps


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--single-steps", type=int, default=500,
                    help="steps per replica of the one-by-one comparison")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the one-by-one comparison")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    api = ProductAPI()
    reps = replicas(args.systems, args.seed)
    natoms = sum(r["n"] for r in reps)
    log("%d replicas, %d atoms (N %d..%d)" % (len(reps), natoms, min(r["n"] for r in reps),
                                               max(r["n"] for r in reps)))
    systems = make_systems(api, reps)
    b = Batch(api, systems, [r["T"] for r in reps])
    t0 = time.perf_counter()
    shape = b.tune() if not args.no_tune else b.shape()
    t_tune = time.perf_counter() - t0
    lanes, grid = shape
    log("tuned shape: grid %d, lanes per bucket %s (%.1f s)" % (
        grid, sorted({(("<2048", "<8192", ">=8192")[0 if r["n"] < 2048 else (1 if r["n"] < 8192 else 2)], l)
                      for r, l in zip(reps, lanes)}), t_tune))
    b.run(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(args.steps)
    b.synchronize()
    secs = time.perf_counter() - t0
    value = natoms * args.steps / secs
    log("batch: %d steps in %.3f s, %.4g atom-steps/s" % (args.steps, secs, value))
    rets, _ = b.sync()
    temps = [2.0 / 3.0 * r.ekin / s.n for r, s in zip(rets, systems)]
    agree = agreement(b)
    b.close()
    for s in systems:
        s.close()
    single = None
    if not args.no_single:
        t_single, as_single = one_by_one(api, reps, args.single_steps)
        single = as_single / t_single
        log("one by one: %.4g atom-steps/s" % single)
    print(json.dumps({
        "workload": "%d LJ NVT replicas, N log-uniform 500..32000 (total %d), rho U[0.70,0.95], "
                    "T U[0.8,2.0], rc 2.5, skin 0.3, dt 0.004, NH tau 0.2, FP32 pair arithmetic "
                    "+ FP64 accumulation" % (len(reps), natoms),
        "steps": args.steps, "atom_steps_per_s": value, "ms_per_step": 1e3 * secs / args.steps,
        "one_by_one_atom_steps_per_s": single,
        "speedup_vs_one_by_one": value / single if single else None,
        "grid": grid, "lanes": lanes,
        "max_epot_rel_vs_fp64": max(a["epot"] for a in agree),
        "max_force_err_vs_fp64": max(a["force"] for a in agree),
        "max_stress_rel_vs_fp64": max(a["stress"] for a in agree),
        "final_T_over_target": [t / r["T"] for t, r in zip(temps, reps)][:8],
        "seconds_tune": round(t_tune, 2)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
