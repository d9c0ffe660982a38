"""Throughput benchmark of the seplib per-timestep hot path on one MI355X.

Workload (the headline configuration): LJ NVT, N = 256000 atoms on a
40x40x40 fcc lattice (4 atoms per cell) at density 0.8442 (L = 67.1838),
seeded Maxwell-Boltzmann velocities at T = 1.0, shifted LJ (eps = sigma = 1,
rc = 2.5, sep_lj_shift), Verlet list with skin 0.3 on a 23^3 cell grid
(cell edge 2.921 >= rc + skin), rebuild when the largest displacement
exceeds skin/2, Nose-Hoover at T0 = 1.0 with Q = 10, dt = 0.005.

One step is the reference's call sequence (prgs/prg1.c:55-65):
sep_reset_retval, sep_reset_force, sep_force_pairs (+ list rebuild when
flagged), sep_nosehoover, sep_leapfrog -- all on the device in resident
mode, replayed from a captured HIP graph, energies and stress reduced on
the device every step with no host copy.

The timed region is K steps of the mixed form (FP32 pair arithmetic on
positions quantized to 2^-32 of the box, FP64 accumulation, FP64 positions
and integrator) and K steps of the plain FP64 form; ``value`` is the mixed
form's atom-steps/s.  The mixed form's force
error against FP64 is measured on the final state.  The step's dominant
kernel -- pair forces, the Nose-Hoover term and the leapfrog fused in one
launch -- gives the roofline: its average launch duration (``achieved``)
comes from HIP events on the library stream around 200 back-to-back
launches on the final state of the timed run, with the integrator's outputs
sent to scratch so that every launch sees the same state.
``traffic`` is the HBM bytes per launch of that kernel from the rocprofv3
FETCH_SIZE / WRITE_SIZE counter summary committed under profiles/, scaled by
the counters' reading for a plain device copy of known size.

cpu_baseline: the reference's own C (oracle/_ref/libsepref_fast.so, built
from its sources with the shipped -Ofast flags) with OpenMP on the host,
over a bounded 200-step sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--copy-calib]

--copy-calib first runs a plain 256 MiB device-to-device copy kernel (for
the counter calibration of a profiled run).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_SIDE = 40
RHO = 0.8442
TEMP = 1.0
RC = 2.5
SKIN = 0.3
DT = 0.005
NSUB = 23
Q_NH = 10.0
KERNEL_REPS = 200


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(api, w, precision=None):
    from seplib_amd.sim import SepSystem
    s = SepSystem.from_workload(api, w, cf=RC, dt=DT, skin=SKIN, nsubbox=(NSUB,) * 3)
    if precision is not None:
        s.gpu_precision(precision)
    return s


def step_ops(s):
    return [s.op_pairs("AA", RC, "sep_lj_shift"), s.op_nosehoover(TEMP, Q_NH), s.op_leapfrog()]


def run_gpu(api, w, precision, steps, warmup, time_kernel=True):
    """Returns (seconds for `steps` graph-replayed steps, pair-kernel ms per
    launch, mean full-list length, system)."""
    s = make_system(api, w, precision)
    lib = api.lib
    s.gpu_resident(True)
    ops = step_ops(s)
    s.gpu_run(ops, warmup, graph=True)
    s.check(lib.sep_gpu_synchronize(s.atoms))
    t0 = time.perf_counter()
    s.gpu_run(ops, steps, graph=True)
    s.check(lib.sep_gpu_synchronize(s.atoms))
    secs = time.perf_counter() - t0
    kern_ms = None
    if time_kernel:
        import ctypes as C
        # HIP events on the library stream around KERNEL_REPS back-to-back
        # launches of the step's fused kernel on the final state
        ms = C.c_double(0)
        s.check(lib.sep_gpu_time_pair_kernel(s.atoms, KERNEL_REPS, C.byref(ms)))
        kern_ms = ms.value
    import ctypes as C
    mean, mx = C.c_double(0), C.c_int(0)
    s.check(lib.sep_gpu_list_stats(s.atoms, C.byref(mean), C.byref(mx)))
    s.gpu_sync()
    return secs, kern_ms, mean.value, s


def force_error(api, state_src):
    """Mixed vs FP64 forces on one state: max|dF|/RMS|F| and RMS|dF|/RMS|F|."""
    import numpy as np
    from seplib_amd import sep as S
    st = state_src.state()
    out = {}
    for prec in (S.SEP_GPU_FP64, S.SEP_GPU_MIXED):
        s = make_system(api, state_src.workload, prec)
        s.set_state(st)
        s.sys.neighb_flag = 1
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", RC, "sep_lj_shift", S.SEP_ALL)
        out[prec] = (np.array(s.view["f"]), s.ret.epot)
        s.close()
    f64, f32 = out[S.SEP_GPU_FP64][0], out[S.SEP_GPU_MIXED][0]
    rms = np.sqrt((f64 ** 2).sum(axis=1).mean())
    d = np.sqrt(((f32 - f64) ** 2).sum(axis=1))
    e64, e32 = out[S.SEP_GPU_FP64][1], out[S.SEP_GPU_MIXED][1]
    return float(d.max() / rms), float(np.sqrt((d ** 2).mean()) / rms), abs(e32 - e64) / abs(e64)


def cpu_baseline(w, steps=200, warm=20):
    """The compiled reference with OpenMP on the host, bounded sample."""
    from seplib_amd import sep as S
    from oracle.checkers import REFERENCE_FAST_LIB, ReferenceAPI
    if not ReferenceAPI.available(REFERENCE_FAST_LIB):
        return None
    api = ReferenceAPI(REFERENCE_FAST_LIB)
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    s = make_system(api, w)
    api.set_omp(threads, s.sysp)

    def one():
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", RC, "sep_lj_shift", S.SEP_ALL)
        s.nosehoover(TEMP, Q_NH)
        s.leapfrog()
    for _ in range(warm):
        one()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    secs = time.perf_counter() - t0
    s.close()
    return dict(value=len(w["m"]) * steps / secs, unit="atom-steps/s", cores=threads,
                kind="reference",
                sample="%d NVT steps of the same N=256000 workload after %d warm-up steps, "
                       "reference C (shipped -Ofast flags) with %d OpenMP threads, "
                       "%.1f s" % (steps, warm, threads, secs))


def energy_drift(api, w, precision, nvt_steps, nve_steps, every=100):
    """NVT equilibration, then NVE (no thermostat) with the total energy
    epot + ekin read every `every` steps: relative drift |E_end - E_0| / |E_0|
    and the largest excursion max_t |E(t) - E_0| / |E_0|."""
    s = make_system(api, w, precision)
    s.gpu_resident(True)
    s.gpu_run(step_ops(s), nvt_steps, graph=True)
    nve = [s.op_pairs("AA", RC, "sep_lj_shift"), s.op_leapfrog()]
    energies = []
    for k in range(nve_steps // every):
        s.gpu_run(nve, every, graph=True)
        s.gpu_sync()
        energies.append(s.ret.epot + s.ret.ekin)
    s.close()
    e0 = energies[0]
    drift = abs(energies[-1] - e0) / abs(e0)
    excursion = max(abs(e - e0) for e in energies) / abs(e0)
    return drift, excursion, e0 / len(w["m"])


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_lj_lj256k_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary,
    or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(nbytes=256 << 20):
    """A plain device copy of known size, for counter calibration."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copy calibration: 3 x %d bytes device-to-device" % nbytes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--copy-calib", action="store_true")
    ap.add_argument("--drift", action="store_true",
                    help="also measure the NVE energy drift (1e4 NVT + 1e4 NVE steps per form)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from seplib_amd import sep as S
    from seplib_amd import workloads as W
    from seplib_amd.sim import ProductAPI

    if args.copy_calib:
        copy_calibration()
    w = W.lj_fcc(N_SIDE, RHO, TEMP, seed=2)
    n = len(w["m"])
    api = ProductAPI()
    log("N=%d L=%.6f, %d timed steps after %d warm-up" % (n, w["L"][0], args.steps, args.warmup))

    res = {}
    for name, prec in (("fp64", S.SEP_GPU_FP64), ("mixed", S.SEP_GPU_MIXED)):
        secs, kms, nbar, s = run_gpu(api, w, prec, args.steps, args.warmup)
        s.workload = w
        res[name] = dict(secs=secs, kms=kms, nbar=nbar, sys=s,
                         rebuilds=s.sys.nupdate_neighb, epot=s.ret.epot / n)
        log("%s: %.3f s, %.4g atom-steps/s, pair kernel %.2f us, nbar %.2f, rebuilds %d, "
            "epot/N %.6f" % (name, secs, n * args.steps / secs, kms * 1e3, nbar,
                             s.sys.nupdate_neighb, s.ret.epot / n))
    ferr_max, ferr_rms, eerr = force_error(api, res["mixed"]["sys"])
    log("mixed vs fp64 forces: max %.3e rms %.3e (relative to RMS|F|), epot %.3e" %
        (ferr_max, ferr_rms, eerr))

    m = res["mixed"]
    value = n * args.steps / m["secs"]
    fp64_value = n * args.steps / res["fp64"]["secs"]
    # The step's dominant kernel is the fused pair + Nose-Hoover + leapfrog
    # kernel.  Algorithmic bytes per atom per launch: int32 full-list
    # entries and the list length (4 nbar + 4); the pair part's own 16-byte
    # quantized record; the integrator's reads of x (32), v (24), m (8),
    # xn (24), cross_neighb and crossings (12 + 12); its writes of f, a, v
    # (3 x 24), the new position and its quantized record (32 + 16),
    # cross_neighb and crossings (12 + 12).
    bytes_per_atom = 4.0 * m["nbar"] + 4 + 16 + (32 + 24 + 8 + 24 + 12 + 12) + \
        (3 * 24 + 32 + 16 + 12 + 12)
    achieved = bytes_per_atom * n / (m["kms"] * 1e-3) / 1e9
    kname = "k_lj_step<true, true>"
    traffic = measured_traffic(kname)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=None if traffic is None else round(traffic),
                    kernel=kname + " (mixed pair forces, Nose-Hoover and leapfrog)",
                    kernel_us=round(m["kms"] * 1e3, 3),
                    algorithmic_bytes=round(bytes_per_atom * n),
                    bytes_per_atom=round(bytes_per_atom, 2))
    drift = None
    if args.drift:
        drift = {}
        for name, prec in (("fp64", S.SEP_GPU_FP64), ("mixed", S.SEP_GPU_MIXED)):
            d, x, e0 = energy_drift(api, w, prec, 10000, 10000)
            drift[name] = dict(drift=d, max_excursion=x, e0_per_atom=e0)
            log("%s NVE over 1e4 steps after 1e4 NVT: drift %.3e, max excursion %.3e "
                "(E0/N %.6f)" % (name, d, x, e0))
    cpu = None if args.no_cpu else cpu_baseline(w, steps=args.cpu_steps)
    if cpu:
        log("cpu baseline: %.4g atom-steps/s on %d threads" % (cpu["value"], cpu["cores"]))
    out = {
        "metric": "atom-steps/s LJ N=256k, FP32+FP64-acc vs FP64 GPU at ≤1e-5 force rel err",
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["secs"] / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-pair/fp64-acc",
        "data": "synthetic (seeded fcc lattice + Maxwell-Boltzmann velocities)",
        "config": {"workload": "LJ NVT N=256000 fcc rho=0.8442 T=1.0 rc=2.5 skin=0.3 "
                               "cells=23^3 NH Q=10 dt=0.005",
                   "parallelism": "single GPU"},
        "fp64_value": round(fp64_value, 1),
        "mixed_over_fp64": round(value / fp64_value, 3),
        "force_rel_err_max": ferr_max,
        "force_rel_err_rms": ferr_rms,
        "fp64_pair_kernel_us": round(res["fp64"]["kms"] * 1e3, 3),
        "rebuilds": m["rebuilds"],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if drift is not None:
        out["nve_energy_drift"] = drift
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
