"""The precision study of an LJ workload: the mixed form (FP32 pair
arithmetic on box-quantized positions, FP64 sums, FP64 positions and
integrator) against the plain FP64 form, both on the device.

  * force error along the mixed trajectory: every `every` steps of a
    resident NVT run in the mixed form, its positions go to two resident
    systems that evaluate the FP64 and the mixed pair forces on them (each
    building its list afresh); the error of a sample is max_i |dF_i| /
    RMS|F| and RMS|dF| / RMS|F| with F the FP64 forces, and the study keeps
    the worst sample;
  * NVE drift, per form: after the NVT segment the thermostat is dropped
    and the total energy epot + ekin (the leapfrog's half-step kinetic
    energy, as sep_leapfrog reports it) is read every `every` steps;
    drift = |E_end - E_0| / |E_0| and the largest excursion
    max_t |E(t) - E_0| / |E_0|.

Used by bench.py (the 256000-atom workload) and tests/test_gpu_precision.py.
"""
import numpy as np

from . import sep as S


def _forces(s, st, pair_call):
    """Pair forces of a resident system on the positions of state st."""
    s.set_state({"x": st["x"], "v": st["v"], "neighb_flag": 1})
    s.check(s._lib().sep_gpu_upload(s.atoms, s.sysp))
    s.reset_retval()
    s.reset_force()
    pair_call(s)
    s.gpu_sync()
    return np.array(s.view["f"]), s.ret.epot


def force_errors(f, fref):
    rms = np.sqrt((fref ** 2).sum(axis=1).mean())
    d = np.sqrt(((f - fref) ** 2).sum(axis=1))
    return float(d.max() / rms), float(np.sqrt((d ** 2).mean()) / rms)


class ForceSampler:
    """FP64 and mixed pair forces on given positions (two resident systems
    made by factory(precision))."""

    def __init__(self, factory, pair_call):
        self.pair_call = pair_call
        self.sys = {p: factory(p) for p in (S.SEP_GPU_FP64, S.SEP_GPU_MIXED)}
        for s in self.sys.values():
            s.gpu_resident(True)
        self.worst_max = self.worst_rms = self.worst_epot = 0.0
        self.samples = 0

    def sample(self, st):
        f64, e64 = _forces(self.sys[S.SEP_GPU_FP64], st, self.pair_call)
        f32, e32 = _forces(self.sys[S.SEP_GPU_MIXED], st, self.pair_call)
        emax, erms = force_errors(f32, f64)
        self.worst_max = max(self.worst_max, emax)
        self.worst_rms = max(self.worst_rms, erms)
        self.worst_epot = max(self.worst_epot, abs(e32 - e64) / abs(e64))
        self.samples += 1
        return emax, erms

    def close(self):
        for s in self.sys.values():
            s.close()


def trajectory(factory, precision, nvt_ops, nve_ops, nvt_steps, nve_steps, every,
               sampler=None):
    """NVT then NVE on the device in `every`-step graph runs; the sampler
    (if any) sees the state after every NVT chunk.  Returns the NVE drift,
    the largest excursion, E_0 and the number of list rebuilds."""
    s = factory(precision)
    try:
        s.gpu_resident(True)
        for _ in range(nvt_steps // every):
            s.gpu_run(nvt_ops(s), every, graph=True)
            if sampler is not None:
                s.gpu_sync()
                sampler.sample(s.state())
        energies = []
        for _ in range(nve_steps // every):
            s.gpu_run(nve_ops(s), every, graph=True)
            s.gpu_sync()
            energies.append(s.ret.epot + s.ret.ekin)
        e0 = energies[0]
        return dict(drift=abs(energies[-1] - e0) / abs(e0),
                    max_excursion=max(abs(e - e0) for e in energies) / abs(e0),
                    e0=e0, rebuilds=int(s.sys.nupdate_neighb))
    finally:
        s.close()


def study(factory, pair_call, nvt_ops, nve_ops, nvt_steps=10000, nve_steps=10000,
          every=100):
    """Both forms' NVE drift, and the force error sampled along the mixed
    form's NVT trajectory."""
    sampler = ForceSampler(factory, pair_call)
    try:
        out = {"mixed": trajectory(factory, S.SEP_GPU_MIXED, nvt_ops, nve_ops, nvt_steps,
                                   nve_steps, every, sampler)}
        out["fp64"] = trajectory(factory, S.SEP_GPU_FP64, nvt_ops, nve_ops, nvt_steps,
                                 nve_steps, every)
        out["force_err_max"] = sampler.worst_max
        out["force_err_rms"] = sampler.worst_rms
        out["epot_err_max"] = sampler.worst_epot
        out["force_samples"] = sampler.samples
        return out
    finally:
        sampler.close()
