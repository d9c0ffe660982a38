"""Python driver over the seplib C99 API.

``SepSystem`` owns one atom array, its ``sepsys`` and a ``sepret`` in the
library's memory and exposes the reference's per-step calls under their C
names, so a test or benchmark reads like the reference's example programs
(prgs/prg1.c:52-83):

    s = SepSystem.from_workload(ProductAPI(), w, cf=2.5, dt=0.005)
    s.reset_retval(); s.reset_force()
    s.force_pairs("AA", 2.5, "sep_lj_shift", SEP_ALL)
    s.nosehoover(1.0, alpha, 10.0); s.leapfrog()

The ``api`` object hides the only differences between the libraries that
share this ABI (allocation entry points and name prefixes); the product's
is ``ProductAPI``.  Checkers provide their own in oracle/checkers.py.
"""
import ctypes as C
import os
import tempfile

import numpy as np

from . import sep as S


class ProductAPI:
    """libsep.so, the MI355X library."""

    name = "product"
    prefix = "sep_"

    def __init__(self, lib=None):
        self.lib = lib or S.load_product()

    def alloc(self, n, nneighb):
        return self.lib.sep_init(n, nneighb)

    def free(self, ptr, n):
        self.lib.sep_close(ptr, n)

    def sys_setup(self, L, cf, dt, n, update):
        return self.lib.sep_sys_setup(L[0], L[1], L[2], cf, dt, n, update)

    def free_sys(self, sysp):
        self.lib.sep_free_sys(sysp)

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def pot(self, name):
        """Address of a potential function for the fun argument."""
        short = name[4:] if name.startswith("sep_") else name
        return S.fn_address(self.lib, self.prefix + short)

    def force_pairs(self, ptr, types, cf, fun, sysp, retp, opt, nneighb):
        return self.lib.sep_force_pairs(ptr, types, cf, fun, sysp, retp, opt)

    def force_lj(self, ptr, types, prm, sysp, retp, opt, nneighb):
        return self.lib.sep_force_lj(ptr, types, prm, sysp, retp, opt)

    def coulomb_sf(self, ptr, cf, sysp, retp, opt):
        self.lib.sep_coulomb_sf(ptr, cf, sysp, retp, opt)

    def torsion(self, ptr, type_, g, sysp, retp):
        self.lib.sep_torsion_Ryckaert(ptr, type_, g, sysp, retp)

    def omp_make_neighb(self, ptr, cf, sysp, opt, nneighb):
        self.lib.sep_omp_make_neighb(ptr, cf, sysp, opt)

    def read_topology(self, ptr, path, sysp):
        self.lib.sep_read_topology_file(ptr, path.encode(), sysp, b"q")


class SepSystem:
    def __init__(self, api, n, L, cf, dt, update=S.SEP_LLIST_NEIGHBLIST,
                 nneighb=S.SEP_NEIGHB):
        self.api = api
        self.n = n
        self.nneighb = nneighb
        self.atoms = api.alloc(n, nneighb)
        if not self.atoms:
            raise RuntimeError("allocation failed")
        self.view = S.atoms_view(self.atoms, n)
        # start from a fully defined record (the reference leaves xn and
        # the velocity scratch fields uninitialised)
        for f in ("x", "v", "f", "a", "xn", "x0", "xtrue", "xp", "px", "pv", "pa",
                  "randn", "prevf"):
            self.view[f] = 0.0
        self.view["cross_neighb"] = 0
        self.view["crossings"] = 0
        self.sys = api.sys_setup(L, cf, dt, n, update)
        self.ret = S.sepret()
        self.sysp = C.pointer(self.sys)
        self.retp = C.pointer(self.ret)
        self.alpha = C.c_double(0.0)
        self._fns = {}

    @classmethod
    def from_workload(cls, api, w, cf, dt, update=S.SEP_LLIST_NEIGHBLIST,
                      nneighb=S.SEP_NEIGHB, skin=None, nsubbox=None):
        s = cls(api, len(w["m"]), w["L"], cf, dt, update, nneighb)
        s.load(w)
        if skin is not None:
            s.sys.skin = skin
        if nsubbox is not None:
            s.set_grid(nsubbox)
        if "bonds" in w or "angles" in w or "dihedrals" in w:
            s.load_topology(w)
        return s

    # ---- state ----------------------------------------------------------
    def load(self, w):
        v = self.view
        v["x"] = w["x"]
        v["v"] = w["v"]
        v["m"] = w["m"]
        v["z"] = w["q"]
        v["type"] = w["type"]
        v["molindex"] = w.get("mol", np.full(self.n, -1))

    def set_grid(self, nsub):
        for k in range(3):
            self.sys.nsubbox[k] = int(nsub[k])
            self.sys.lsubbox[k] = self.sys.length[k] / int(nsub[k])

    def load_topology(self, w):
        from .workloads import write_top
        fd, path = tempfile.mkstemp(suffix=".top")
        os.close(fd)
        try:
            write_top(path, w)
            self.api.read_topology(self.atoms, path, self.sysp)
        finally:
            os.unlink(path)

    def state(self):
        v = self.view
        out = {k: np.array(v[k]) for k in ("x", "v", "f", "a", "xn", "cross_neighb",
                                           "crossings")}
        out.update(max_dist2=self.sys.max_dist2, neighb_flag=self.sys.neighb_flag,
                   nupdate_neighb=self.sys.nupdate_neighb, tnow=self.sys.tnow,
                   alpha=self.alpha.value)
        return out

    def set_state(self, st):
        v = self.view
        for k in ("x", "v", "f", "a", "xn", "cross_neighb", "crossings"):
            if k in st:
                v[k] = st[k]
        for k in ("max_dist2", "neighb_flag", "nupdate_neighb", "tnow"):
            if k in st:
                setattr(self.sys, k, type(getattr(self.sys, k))(st[k]))
        if "alpha" in st:
            self.alpha.value = st["alpha"]

    def ret_dict(self):
        r = self.ret
        return dict(epot=r.epot, ekin=r.ekin, ecoul=r.ecoul, pot_P=S.mat(r.pot_P),
                    kin_P=S.mat(r.kin_P), pot_P_bond=S.mat(r.pot_P_bond))

    def half_pairs(self):
        """Sorted (i, j) pairs of the host half lists atoms[i].neighb."""
        out = []
        for i in range(self.n):
            p = self.atoms[i].neighb
            k = 0
            while p[k] != -1:
                j = p[k]
                out.append((min(i, j), max(i, j)))
                k += 1
        a = np.array(out, dtype=np.int64).reshape(-1, 2)
        if len(a):
            a = a[np.lexsort((a[:, 1], a[:, 0]))]
        return a

    # ---- API calls --------------------------------------------------------
    def reset_retval(self):
        self.api.fn("reset_retval")(self.retp)

    def reset_force(self):
        self.api.fn("reset_force")(self.atoms, self.sysp)

    def force_pairs(self, types, cf, fun="sep_lj_shift", opt=S.SEP_ALL):
        return self.api.force_pairs(self.atoms, types.encode(), cf, self.api.pot(fun),
                                    self.sysp, self.retp, opt, self.nneighb)

    def force_lj(self, types, param, opt=S.SEP_ALL):
        prm = (C.c_double * 4)(*param)
        return self.api.force_lj(self.atoms, types.encode(), prm, self.sysp, self.retp,
                                 opt, self.nneighb)

    def coulomb_sf(self, cf, opt=S.SEP_ALL):
        self.api.coulomb_sf(self.atoms, cf, self.sysp, self.retp, opt)

    def stretch_harmonic(self, type_, lbond, ks):
        self.api.fn("stretch_harmonic")(self.atoms, type_, lbond, ks, self.sysp, self.retp)

    def angle_harmonic(self, type_, angle0, k):
        self.api.fn("angle_harmonic")(self.atoms, type_, angle0, k, self.sysp, self.retp)

    def angle_cossq(self, type_, angle0, k):
        self.api.fn("angle_cossq")(self.atoms, type_, angle0, k, self.sysp, self.retp)

    def omp_make_neighb(self, cf, opt=S.SEP_ALL):
        return self.api.omp_make_neighb(self.atoms, cf, self.sysp, opt, self.nneighb)

    def omp_term(self, name, *args):
        """One OpenMP Model II term (sep_omp_<name>) into a fresh N x 3
        array, returned."""
        f = np.zeros((self.n, 3))
        mat, _rows = S.row_pointers(f)
        self.api.fn("omp_" + name)(mat, self.atoms, *args, self.sysp)
        return f

    def torsion(self, type_, g):
        self.api.torsion(self.atoms, type_, (C.c_double * 6)(*g), self.sysp, self.retp)

    def nosehoover(self, Td, tau, alpha=None):
        a = self.alpha if alpha is None else alpha
        self.api.fn("nosehoover")(self.atoms, Td, C.byref(a), tau, self.sysp)

    def leapfrog(self):
        self.api.fn("leapfrog")(self.atoms, self.sysp, self.retp)

    def pressure_tensor(self):
        self.api.fn("pressure_tensor")(self.retp, self.sysp)

    def close(self):
        if self.atoms:
            self.api.free_sys(self.sysp)
            self.api.free(self.atoms, self.n)
            self.atoms = None

    # ---- device extension (product only) ------------------------------------
    def _lib(self):
        return self.api.lib

    def check(self, rc):
        if rc:
            msg = self._lib().sep_gpu_last_error_message().decode()
            raise RuntimeError("sepgpu error %d: %s" % (rc, msg))

    def gpu_resident(self, on=True):
        self.check(self._lib().sep_gpu_resident(self.atoms, self.sysp, 1 if on else 0))

    def gpu_sync(self):
        self.check(self._lib().sep_gpu_sync(self.atoms, self.sysp))

    def gpu_precision(self, mode):
        self.check(self._lib().sep_gpu_set_precision(self.atoms, self.sysp, mode))

    def gpu_run(self, ops, nsteps, graph=False, concurrent=False):
        arr = (S.sep_gpu_op * len(ops))(*ops)
        flags = (S.SEP_GPU_RUN_GRAPH if graph else 0) | (S.SEP_GPU_RUN_CONCURRENT if concurrent else 0)
        self.check(self._lib().sep_gpu_run(self.atoms, self.sysp, self.retp, arr, len(ops),
                                           nsteps, flags))

    def gpu_pairs(self):
        lib = self._lib()
        cnt = C.c_long(0)
        self.check(lib.sep_gpu_neighbour_pairs(self.atoms, self.sysp, None, 0, C.byref(cnt)))
        buf = (C.c_int * (2 * cnt.value + 2))()
        self.check(lib.sep_gpu_neighbour_pairs(self.atoms, self.sysp, buf, cnt.value,
                                               C.byref(cnt)))
        a = np.frombuffer(buf, dtype=np.int32, count=2 * cnt.value).reshape(-1, 2)
        a = a.astype(np.int64)
        return a[np.lexsort((a[:, 1], a[:, 0]))]

    def op_pairs(self, types, cf, fun="sep_lj_shift", opt=S.SEP_ALL):
        o = S.sep_gpu_op(op=S.OP_PAIRS, types=types.encode(), opt=opt)
        o.p[0] = cf
        o.fun = self.api.pot(fun)
        return o

    def op_lj(self, types, param, opt=S.SEP_ALL):
        o = S.sep_gpu_op(op=S.OP_LJ, types=types.encode(), opt=opt)
        for i, v in enumerate(param):
            o.p[i] = v
        return o

    def op_coulomb(self, cf, opt=S.SEP_ALL):
        o = S.sep_gpu_op(op=S.OP_COULOMB, opt=opt)
        o.p[0] = cf
        return o

    def op_bonded(self, kind, type_, params):
        o = S.sep_gpu_op(op=kind, type=type_)
        for i, v in enumerate(params):
            o.p[i] = v
        return o

    def op_nosehoover(self, Td, tau):
        o = S.sep_gpu_op(op=S.OP_NOSEHOOVER)
        o.p[0] = Td
        o.p[1] = tau
        o.alpha = C.pointer(self.alpha)
        return o

    def op_leapfrog(self):
        return S.sep_gpu_op(op=S.OP_LEAPFROG)
