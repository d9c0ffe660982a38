"""sep_save / sep_load (source/sepmisc.c:319-354) on a resident system.

  * sep_save after resident graph steps writes the device state: loaded
    back into a host system, x, v, f, a, xn, cross_neighb and crossings
    equal the synchronised AoS of the resident run bit for bit;
  * sep_load into a resident system replaces its device state and rebuilds
    its list: pair forces, energy and virial then equal those of a
    synchronised system loaded from the same file to 1e-12;
  * a load with smaller neighbour arrays than sep_init gave is respected
    by the host half lists written afterwards (no overflow).
"""
import numpy as np
import pytest

from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem
from tests.parity import FP64_TOL, force_error, rel, tensor_error

pytestmark = pytest.mark.gpu

KW = dict(cf=2.5, dt=0.005, skin=0.4)


def test_resident_save_and_load(product, tmp_path):
    w = W.lj_sc(10, 0.8442, 1.0, seed=6)
    path = str(tmp_path / "ckpt.bin").encode()
    lib = product.lib
    a = SepSystem.from_workload(product, w, **KW)
    b = SepSystem.from_workload(product, w, **KW)
    c = SepSystem.from_workload(product, w, **KW)
    try:
        a.gpu_resident(True)
        ops = [a.op_pairs("AA", 2.5), a.op_nosehoover(1.0, 10.0), a.op_leapfrog()]
        a.gpu_run(ops, 30, graph=True)
        lib.sep_save(a.atoms, a.n, path)
        a.gpu_sync()
        sa = a.state()
        # the file holds the device state
        lib.sep_load(c.atoms, S.SEP_NEIGHB, c.n, path)
        sc = c.state()
        for k in ("x", "v", "f", "a", "xn", "cross_neighb", "crossings"):
            assert np.array_equal(sa[k], sc[k]), k
        # a resident system loaded from it: forces on the loaded positions
        b.gpu_resident(True)
        b.force_pairs("AA", 2.5)  # a list for the original positions
        lib.sep_load(b.atoms, S.SEP_NEIGHB, b.n, path)
        b.reset_retval()
        b.reset_force()
        b.force_pairs("AA", 2.5)
        b.gpu_sync()
        c.sys.neighb_flag = 1
        c.reset_retval()
        c.reset_force()
        c.force_pairs("AA", 2.5)
        assert force_error(np.array(b.view["f"]), np.array(c.view["f"])) <= FP64_TOL
        rb, rc = b.ret_dict(), c.ret_dict()
        assert rel(rb["epot"], rc["epot"]) <= FP64_TOL
        assert tensor_error(rb["pot_P"], rc["pot_P"]) <= FP64_TOL
    finally:
        a.close()
        b.close()
        c.close()


def test_load_with_smaller_neighbour_arrays(product, tmp_path):
    w = W.lj_sc(8, 0.8442, 1.0, seed=7)
    path = str(tmp_path / "small.bin").encode()
    s = SepSystem.from_workload(product, w, **KW)
    try:
        s.force_pairs("AA", 2.5)
        product.lib.sep_save(s.atoms, s.n, path)
        product.lib.sep_load(s.atoms, 160, s.n, path)
        s.sys.neighb_flag = 1
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", 2.5)
        half = s.half_pairs()
        assert len(half) > 0
    finally:
        s.close()


def test_resident_load_replaces_bonded_partners(product, tmp_path):
    """sep_load reads each atom's bond, angle and dihedral partners from the
    file: a resident system that excluded bonded pairs before the load
    excludes by the loaded partners after it (here none), as a synchronised
    system loaded from the same file does."""
    import os
    from tests.golden import runs as R
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "g6_excl.npz"),
                allow_pickle=False)
    w = R.g6_workload(g, "butane")
    plain = {k: v for k, v in w.items() if k not in ("bonds", "angles", "dihedrals")}
    path = str(tmp_path / "nobonds.bin").encode()
    kw = dict(cf=2.5, dt=0.001)
    lib = product.lib
    src = SepSystem.from_workload(product, plain, **kw)
    b = SepSystem.from_workload(product, w, **kw)
    c = SepSystem.from_workload(product, plain, **kw)
    try:
        lib.sep_save(src.atoms, src.n, path)
        b.gpu_resident(True)
        b.reset_retval()
        b.reset_force()
        b.force_pairs("CC", 2.5, "sep_lj_shift", S.SEP_EXCL_BONDED)  # partners uploaded
        lib.sep_load(b.atoms, S.SEP_NEIGHB, b.n, path)
        for s in (b, c):
            s.reset_retval()
            s.reset_force()
            s.force_pairs("CC", 2.5, "sep_lj_shift", S.SEP_EXCL_BONDED)
        b.gpu_sync()
        fb, fc = np.array(b.view["f"]), np.array(c.view["f"])
        assert force_error(fb, fc) <= FP64_TOL
        assert rel(b.ret.epot, c.ret.epot) <= FP64_TOL
        # without exclusions the bonded neighbours dominate the forces
        assert np.abs(fc).max() > 1e3
    finally:
        src.close()
        b.close()
        c.close()


def _snap(s):
    s.gpu_sync()
    st = s.state()
    r = s.ret
    st.update(epot=r.epot, ekin=r.ekin, potP=S.mat(r.pot_P), kinP=S.mat(r.kin_P))
    return st


@pytest.mark.parametrize("precision", [S.SEP_GPU_FP64, S.SEP_GPU_MIXED], ids=["fp64", "mixed"])
def test_device_checkpoint_resumes_bit_for_bit(product, tmp_path, precision):
    """sep_gpu_checkpoint after 37 graph steps, then 53 more: a second
    system (its device state first taken elsewhere by 5 steps of its own)
    restored from the file and run 53 steps ends bit for bit where the first
    one does -- positions, velocities, forces, xn, crossings, the sepret,
    the thermostat alpha and the sepsys counters.  Writing the checkpoint
    leaves the host copy of the atoms alone (no AoS round trip)."""
    import ctypes as C
    w = W.lj_fcc(10, 0.8442, 1.0, seed=8)
    path = str(tmp_path / "run.ckpt").encode()
    lib = product.lib
    kw = dict(cf=2.5, dt=0.005, skin=0.3)
    a = SepSystem.from_workload(product, w, **kw)
    b = SepSystem.from_workload(product, w, **kw)

    def ops(s):
        return [s.op_pairs("AA", 2.5), s.op_nosehoover(1.0, 10.0), s.op_leapfrog()]
    try:
        for s in (a, b):
            s.gpu_precision(precision)
            s.gpu_resident(True)
        a.gpu_run(ops(a), 37, graph=True)
        x_host = np.array(a.view["x"])
        assert lib.sep_gpu_checkpoint(a.atoms, a.sysp, a.retp, C.byref(a.alpha), path) == 0
        assert np.array_equal(np.array(a.view["x"]), x_host)
        a.gpu_sync()
        at_ckpt = (a.sys.tnow, a.sys.nupdate_neighb, a.alpha.value, a.ret.epot)
        a.gpu_run(ops(a), 53, graph=True)
        end_a = _snap(a)
        b.gpu_run(ops(b), 5, graph=False)
        assert lib.sep_gpu_restore(b.atoms, b.sysp, b.retp, C.byref(b.alpha), path) == 0
        assert (b.sys.tnow, b.sys.nupdate_neighb, b.alpha.value, b.ret.epot) == at_ckpt
        b.gpu_run(ops(b), 53, graph=True)
        end_b = _snap(b)
        assert end_a["nupdate_neighb"] > at_ckpt[1]  # rebuilds after the restore
        for k in ("x", "v", "f", "a", "xn", "cross_neighb", "crossings", "potP", "kinP"):
            assert np.array_equal(end_a[k], end_b[k]), k
        for k in ("epot", "ekin", "alpha", "max_dist2", "nupdate_neighb", "neighb_flag", "tnow"):
            assert end_a[k] == end_b[k], k
        # a file that is not a checkpoint, and a system of another size
        bad = str(tmp_path / "bad.ckpt").encode()
        with open(bad, "wb") as f:
            f.write(b"not a checkpoint")
        lib.sep_gpu_clear_error()
        assert lib.sep_gpu_restore(b.atoms, b.sysp, None, None, bad) == S.SEP_GPU_EARG
        small = SepSystem.from_workload(product, W.lj_fcc(6, 0.8442, 1.0, seed=8), **kw)
        try:
            small.gpu_resident(True)
            lib.sep_gpu_clear_error()
            assert lib.sep_gpu_restore(small.atoms, small.sysp, None, None, path) == \
                S.SEP_GPU_EARG
        finally:
            small.close()
        lib.sep_gpu_clear_error()
    finally:
        a.close()
        b.close()
