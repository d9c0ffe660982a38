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
form's atom-steps/s.  After the timed runs, the precision study of the
same workload (seplib_amd/precision.py): 10^4 NVT steps of the mixed form
with its pair forces checked against FP64 forces on the same positions
every 100 steps, then 10^4 NVE steps per form for the energy drift.  The
step's dominant
kernel -- pair forces, the Nose-Hoover term and the leapfrog fused in one
launch -- gives the roofline: its average launch duration (``achieved``)
comes from HIP events on the library stream around 200 back-to-back
launches on the final state of the timed run, with the integrator's outputs
sent to scratch so that every launch sees the same state.
``traffic`` is the HBM bytes per launch of that kernel from the rocprofv3
FETCH_SIZE / WRITE_SIZE counter summary committed under profiles/, scaled by
the counters' reading for a plain device copy of known size.  ``valu`` gives
the same kernel's FP64 and FP32 arithmetic rates against the chip's vector
peaks, from algorithmic operation counts: per full-list entry 8 FP64
operations (three scalings of the integer differences, r^2); per entry
inside the cutoff (counted on the final state with a periodic k-d tree) 7
more in FP64 (three force FMAs, the energy sum) and 12 in FP32 (the
reciprocal and the LJ terms); per atom 45 FP64 for the thermostat term, the
leapfrog and its sums.

cpu_baseline: the reference's own C (oracle/_ref/libsepref_fast.so, built
from its sources with the shipped -Ofast flags) with OpenMP on the host's
cores (at most 16, the host share of one GPU here; OMP_PROC_BIND=close,
OMP_PLACES=cores) over a bounded 200-step sample of the same workload, and
on one thread over 20 steps; the CPU model and core counts are recorded.
Where the compiled reference is absent, the repository's serial C
restatement of the same path (oracle/liboracle.so) on one thread, kind
"port".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-study]
                    [--copy-calib]

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
FP64_PEAK_TFLOPS = 78.6   # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (packed)
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


def pair_call(s):
    from seplib_amd import sep as S
    s.force_pairs("AA", RC, "sep_lj_shift", S.SEP_ALL)


def nve_ops(s):
    return [s.op_pairs("AA", RC, "sep_lj_shift"), s.op_leapfrog()]


def final_state_force_error(api, state_src):
    """Mixed vs FP64 pair forces on one state: max|dF|/RMS|F|, RMS|dF|/RMS|F|
    and the relative energy difference."""
    from seplib_amd import precision as P
    sampler = P.ForceSampler(lambda prec: make_system(api, state_src.workload, prec),
                             pair_call)
    try:
        emax, erms = sampler.sample(state_src.state())
        return emax, erms, sampler.worst_epot
    finally:
        sampler.close()


def cpu_info():
    """CPU model, physical cores and logical CPUs of the host (/proc/cpuinfo),
    and the CPUs this process may run on."""
    model, cores, logical = None, set(), 0
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return dict(model=model, physical_cores=len(cores) or None, logical_cpus=logical or None,
                allowed_cpus=allowed)


def _cpu_rate(api, w, threads, steps, warm):
    from seplib_amd import sep as S
    s = make_system(api, w)
    if threads is not None:
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
    return len(w["m"]) * steps / secs, secs


def cpu_baseline(w, steps=200, warm=20, steps1=20, warm1=2):
    """The compiled reference with OpenMP on the host's cores (bounded
    sample) and on one thread; the repository's C restatement on one thread
    where the compiled reference is absent."""
    from oracle.checkers import REFERENCE_FAST_LIB, OracleAPI, ReferenceAPI
    info = cpu_info()
    n = len(w["m"])
    if ReferenceAPI.available(REFERENCE_FAST_LIB):
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")
        api = ReferenceAPI(REFERENCE_FAST_LIB)
        threads = max(1, min(16, info["allowed_cpus"]))
        value, secs = _cpu_rate(api, w, threads, steps, warm)
        v1, secs1 = _cpu_rate(api, w, 1, steps1, warm1)
        return dict(value=value, unit="atom-steps/s", cores=threads, kind="reference",
                    sample="%d NVT steps of the same N=%d workload after %d warm-up steps, "
                           "reference C (shipped -Ofast flags) with %d OpenMP threads "
                           "(OMP_PROC_BIND=close, OMP_PLACES=cores), %.1f s; one thread: "
                           "%d steps, %.1f s" % (steps, n, warm, threads, secs, steps1, secs1),
                    single_thread=dict(value=v1, steps=steps1, secs=round(secs1, 2)),
                    cpu=info)
    api = OracleAPI()
    v1, secs1 = _cpu_rate(api, w, None, steps1, warm1)
    return dict(value=v1, unit="atom-steps/s", cores=1, kind="port",
                sample="%d NVT steps of the same N=%d workload after %d warm-up steps, the "
                       "repository's serial C restatement of the path (oracle/liboracle.so), "
                       "%.1f s (the compiled reference is not built here)"
                       % (steps1, n, warm1, secs1),
                single_thread=dict(value=v1, steps=steps1, secs=round(secs1, 2)), cpu=info)


def precision_study(api, w, nvt_steps, nve_steps, every=100):
    from seplib_amd import precision as P
    return P.study(lambda prec: make_system(api, w, prec), pair_call, step_ops, nve_ops,
                   nvt_steps, nve_steps, every)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_lj_lj256k_traffic.json")


def pairs_inside(x, L, rc):
    """Ordered pairs (i, j != i) closer than rc under the minimum image."""
    import numpy as np
    from scipy.spatial import cKDTree
    x = np.mod(np.asarray(x, dtype=float), L)
    t = cKDTree(x, boxsize=L)
    return int(t.count_neighbors(t, rc)) - len(x)


def valu_rates(n, nbar, inside, kernel_ms):
    """FP64 / FP32 operation rates of the fused mixed step kernel."""
    entries = n * nbar
    fp64 = 8.0 * entries + 7.0 * inside + 45.0 * n
    fp32 = 12.0 * inside
    sec = kernel_ms * 1e-3
    r64, r32 = fp64 / sec / 1e12, fp32 / sec / 1e12
    return dict(fp64_tflops=round(r64, 3), fp64_peak=FP64_PEAK_TFLOPS,
                fp64_frac=round(r64 / FP64_PEAK_TFLOPS, 4),
                fp32_tflops=round(r32, 3), fp32_peak=FP32_PEAK_TFLOPS,
                fp32_frac=round(r32 / FP32_PEAK_TFLOPS, 4),
                fp64_ops=round(fp64), fp32_ops=round(fp32), pairs_inside=inside)


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
    ap.add_argument("--no-study", action="store_true",
                    help="skip the precision study (1e4 NVT steps of the mixed form with force "
                         "checks every 100 steps, then 1e4 NVE steps per form for the drift)")
    ap.add_argument("--study-steps", type=int, default=10000)
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
    ferr_max, ferr_rms, eerr = final_state_force_error(api, res["mixed"]["sys"])
    inside = pairs_inside(res["mixed"]["sys"].view["x"], w["L"][0], RC)
    log("mixed vs fp64 forces on the final state: max %.3e rms %.3e (relative to RMS|F|), "
        "epot %.3e" % (ferr_max, ferr_rms, eerr))
    for r in res.values():
        r["sys"].close()

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
    valu = valu_rates(n, m["nbar"], inside, m["kms"])
    study = None
    if not args.no_study:
        k = args.study_steps
        t0 = time.perf_counter()
        st = precision_study(api, w, k, k)
        study = {"nvt_steps": k, "nve_steps": k, "every": 100,
                 "force_samples": st["force_samples"],
                 "force_rel_err_max": st["force_err_max"],
                 "force_rel_err_rms": st["force_err_rms"],
                 "epot_rel_err_max": st["epot_err_max"],
                 "nve_drift": {f: {"drift": st[f]["drift"],
                                   "max_excursion": st[f]["max_excursion"],
                                   "e0_per_atom": st[f]["e0"] / n,
                                   "rebuilds": st[f]["rebuilds"]} for f in ("mixed", "fp64")}}
        log("precision study (%.1f s): mixed trajectory, %d samples: force error max %.3e "
            "rms %.3e; NVE drift over %d steps: mixed %.3e, fp64 %.3e"
            % (time.perf_counter() - t0, st["force_samples"], st["force_err_max"],
               st["force_err_rms"], k, st["mixed"]["drift"], st["fp64"]["drift"]))
    cpu = None if args.no_cpu else cpu_baseline(w, steps=args.cpu_steps)
    if cpu:
        log("cpu baseline (%s): %.4g atom-steps/s on %d threads, %.4g on one thread; %s"
            % (cpu["kind"], cpu["value"], cpu["cores"], cpu["single_thread"]["value"],
               cpu["cpu"]["model"]))
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
        "valu": valu,
        "cpu_baseline": cpu,
    }
    if study is not None:
        out["precision_study"] = study
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
