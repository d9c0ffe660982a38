"""Step timing of the flexible water system, with the bonded terms on the
library stream or on a second stream beside the pair terms.

The system: 32000 3-site molecules (96000 atoms) at 3.16 atoms per unit
volume, LJ O-O (eps = sigma = 1, rc 2.5) and shifted-force Coulomb (rc 2.9)
over the same-molecule-excluded list, harmonic O-H bonds (k 68000, l0 0.316)
and the harmonic H-O-H angle (k 490, theta0 1.91), Nose-Hoover at T 3.86,
dt 0.0005, skin 0.25 (the reference default).  The step is the reference's
call sequence (source/sepprfrc.c, source/sepcoulomb.c, source/sepmol.c,
source/sepintgr.c):

    sep_force_pairs("OO", 2.5, sep_lj_shift, SEP_EXCL_SAME_MOL)
    sep_coulomb_sf(2.9, SEP_EXCL_SAME_MOL)
    sep_stretch_harmonic(0, 0.316, 68000)
    sep_angle_harmonic(0, 1.91, 490)
    sep_nosehoover(3.86, 0.2); sep_leapfrog()

run resident, replayed from a captured graph.  With SEP_GPU_RUN_CONCURRENT
the bond and angle kernels run on a second stream while the pair kernels
run (the GPU form of the reference's OpenMP sections, source/sepomp.c and
prgs/prg5.c:58-77); their per-atom gathers stay on the library stream in
call order, so both runs give the same bits.

    python -m seplib_amd.water [--molecules 32000] [--steps 2000]

prints one JSON line: atom-steps/s for FP64 and the mixed form, each with
the terms on one stream and on two, whether the two-stream run's final
state equals the one-stream run's bit for bit, and the mixed form's force
and energy differences from FP64 on the starting state.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import sep as S
from . import workloads as W
from .sim import ProductAPI, SepSystem

LBOND, KBOND, THETA0, KANGLE = 0.316, 68000.0, 1.91, 490.0
RC_LJ, RC_COUL = 2.5, 2.9
TEMP, TAU, DT = 3.86, 0.2, 0.0005
NNEIGHB = 600


def system(nmol, seed=5):
    w = W.water(nmol, 3.16, TEMP, seed=seed, l0=LBOND, theta0=THETA0)
    # off the ideal geometry, so that bonds and angles carry force
    n = len(w["m"])
    w["x"] = np.mod(w["x"] + (W.uniform(41, 3 * n).reshape(n, 3) - 0.5) * 0.01, w["L"])
    return w


def step_ops(s):
    return [s.op_pairs("OO", RC_LJ, "sep_lj_shift", S.SEP_EXCL_SAME_MOL),
            s.op_coulomb(RC_COUL, S.SEP_EXCL_SAME_MOL),
            s.op_bonded(S.OP_BOND, 0, [LBOND, KBOND]),
            s.op_bonded(S.OP_ANGLE, 0, [THETA0, KANGLE]),
            s.op_nosehoover(TEMP, TAU),
            s.op_leapfrog()]


def timed(api, w, steps, warmup, precision, concurrent):
    """Seconds for `steps` graph-replayed steps after `warmup`, the final
    positions, velocities, forces and sepret words, and the list rebuilds
    of the whole run."""
    s = SepSystem.from_workload(api, w, cf=RC_COUL, dt=DT, nneighb=NNEIGHB)
    try:
        s.gpu_precision(precision)
        s.gpu_resident(True)
        ops = step_ops(s)
        s.gpu_run(ops, warmup, graph=True, concurrent=concurrent)
        s.check(api.lib.sep_gpu_synchronize(s.atoms))
        t0 = time.perf_counter()
        s.gpu_run(ops, steps, graph=True, concurrent=concurrent)
        s.check(api.lib.sep_gpu_synchronize(s.atoms))
        secs = time.perf_counter() - t0
        s.gpu_sync()
        st = s.state()
        r = s.ret
        final = (st["x"].copy(), st["v"].copy(), np.array(s.view["f"]),
                 np.array([r.epot, r.ecoul, r.ekin]), s.alpha.value)
        return secs, final, st["nupdate_neighb"]
    finally:
        s.close()


def agreement(api, w):
    """Mixed against FP64 on the starting state: max |dF_i| / RMS|F| of the
    total force, and the relative differences of epot and ecoul."""
    out = {}
    for prec in (S.SEP_GPU_FP64, S.SEP_GPU_MIXED):
        s = SepSystem.from_workload(api, w, cf=RC_COUL, dt=DT, nneighb=NNEIGHB)
        try:
            s.gpu_precision(prec)
            s.reset_retval()
            s.reset_force()
            s.force_pairs("OO", RC_LJ, "sep_lj_shift", S.SEP_EXCL_SAME_MOL)
            s.coulomb_sf(RC_COUL, S.SEP_EXCL_SAME_MOL)
            s.stretch_harmonic(0, LBOND, KBOND)
            s.angle_harmonic(0, THETA0, KANGLE)
            out[prec] = (np.array(s.view["f"]), s.ret.epot, s.ret.ecoul)
        finally:
            s.close()
    f64, e64, c64 = out[S.SEP_GPU_FP64]
    f32, e32, c32 = out[S.SEP_GPU_MIXED]
    rms = np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))
    return {"force_err_max": float(np.abs(f32 - f64).max() / rms),
            "epot_rel": float(abs(e32 - e64) / abs(e64)),
            "ecoul_rel": float(abs(c32 - c64) / abs(c64))}


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--molecules", type=int, default=32000)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    args = ap.parse_args(argv)
    api = ProductAPI()
    w = system(args.molecules)
    n = len(w["m"])
    out = {"workload": "flexible 3-site water, %d molecules (%d atoms), rho 3.16, LJ O-O "
                       "rc 2.5 + SF Coulomb rc 2.9 (same-molecule exclusion), harmonic "
                       "bonds and angles, NH T 3.86, dt 0.0005, skin 0.25; resident, graph "
                       "replay" % (args.molecules, n),
           "steps": args.steps}
    for name, prec in (("fp64", S.SEP_GPU_FP64), ("mixed", S.SEP_GPU_MIXED)):
        secs1, fin1, nup = timed(api, w, args.steps, args.warmup, prec, False)
        secs2, fin2, _ = timed(api, w, args.steps, args.warmup, prec, True)
        out[name] = {"one_stream_atom_steps_per_s": n * args.steps / secs1,
                     "two_streams_atom_steps_per_s": n * args.steps / secs2,
                     "one_stream_us_per_step": 1e6 * secs1 / args.steps,
                     "two_streams_us_per_step": 1e6 * secs2 / args.steps,
                     "bit_identical": bool(same(fin1, fin2)),
                     "rebuilds": int(nup)}
        print("%s: one stream %.1f us/step, two streams %.1f us/step, identical %s" % (
            name, 1e6 * secs1 / args.steps, 1e6 * secs2 / args.steps,
            out[name]["bit_identical"]), file=sys.stderr, flush=True)
    out["mixed_vs_fp64"] = agreement(api, w)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
