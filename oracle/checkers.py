"""Checker libraries for tests/, smoke() and bench.py's cpu_baseline leg.

TEST INFRASTRUCTURE ONLY.  Loads, through the ABI of include/sepstrct.h:
  * the C oracle (oracle/liboracle.so, orc_* names), the repository's
    restatement of the reference hot path;
  * the compiled reference (oracle/_ref/libsepref.so, strict flags, or
    libsepref_fast.so with the shipped -Ofast flags), built from the
    reference's own sources by oracle/Makefile where they are available.
Both plug into seplib_amd.sim.SepSystem as its ``api``.
"""
import ctypes as C
import os

from seplib_amd import sep as S

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REFERENCE_LIB = os.path.join(HERE, "_ref", "libsepref.so")
REFERENCE_FAST_LIB = os.path.join(HERE, "_ref", "libsepref_fast.so")


class ReferenceAPI:
    """The reference compiled from its sources (same names as libsep.so)."""

    name = "reference"
    prefix = "sep_"

    def __init__(self, path=REFERENCE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        # RTLD_LOCAL: the reference's symbols must not interpose libsep.so's
        self.lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        S.bind_api(self.lib, "sep_")
        S.bind_omp(self.lib, "sep_")
        self.lib.sep_omp_make_neighb.restype = None
        self.lib.sep_omp_make_neighb.argtypes = [S.P_ATOMS, C.c_double, S.P_SYS, C.c_uint]
        S.bind_sep_names(self.lib)
        self.lib.sep_set_omp.argtypes = [C.c_uint, S.P_SYS]

    @staticmethod
    def available(path=REFERENCE_LIB):
        return os.path.exists(path)

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

    def read_topology(self, ptr, path, sysp):
        self.lib.sep_read_topology_file(ptr, path.encode(), sysp, b"q")

    def omp_make_neighb(self, ptr, cf, sysp, opt, nneighb):
        self.lib.sep_omp_make_neighb(ptr, cf, sysp, opt)

    def set_omp(self, nthreads, sysp):
        self.lib.sep_set_omp(nthreads, sysp)


class OracleAPI:
    """The repository's serial C restatement (orc_* names)."""

    name = "oracle"
    prefix = "orc_"

    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run make -C oracle)")
        self.lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        S.bind_api(self.lib, "orc_")
        S.bind_omp(self.lib, "orc_")
        self.lib.orc_omp_make_neighb.restype = C.c_int
        self.lib.orc_omp_make_neighb.argtypes = [S.P_ATOMS, C.c_double, S.P_SYS, C.c_uint,
                                                 C.c_int]
        L = self.lib

        def f(name, restype, *args):
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = list(args)
        f("orc_alloc", S.P_ATOMS, C.c_long, C.c_int)
        f("orc_free", None, S.P_ATOMS, C.c_long)
        f("orc_sys_setup", S.sepsys, C.c_double, C.c_double, C.c_double, C.c_double,
          C.c_double, C.c_long, C.c_uint)
        f("orc_free_sys", None, S.P_SYS)
        f("orc_build_list", C.c_int, S.P_ATOMS, S.P_SYS, C.c_uint, C.c_int)
        f("orc_force_pairs", C.c_int, S.P_ATOMS, C.c_char_p, C.c_double, C.c_void_p,
          S.P_SYS, S.P_RET, C.c_uint, C.c_int)
        f("orc_force_lj", C.c_int, S.P_ATOMS, C.c_char_p, S.P_D, S.P_SYS, S.P_RET,
          C.c_uint, C.c_int)
        f("orc_coulomb_sf", None, S.P_ATOMS, C.c_double, S.P_SYS, S.P_RET, C.c_uint)
        f("orc_torsion_ryckaert", None, S.P_ATOMS, C.c_int, S.P_D, S.P_SYS, S.P_RET)
        f("orc_read_topology", C.c_int, S.P_ATOMS, C.c_char_p, S.P_SYS)
        f("orc_list_len", C.c_int, S.P_ATOMS, C.c_long)
        f("orc_half_pairs", C.c_long, S.P_ATOMS, C.c_long, C.c_void_p, C.c_long)

    def alloc(self, n, nneighb):
        return self.lib.orc_alloc(n, nneighb)

    def free(self, ptr, n):
        self.lib.orc_free(ptr, n)

    def sys_setup(self, L, cf, dt, n, update):
        return self.lib.orc_sys_setup(L[0], L[1], L[2], cf, dt, n, update)

    def free_sys(self, sysp):
        self.lib.orc_free_sys(sysp)

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def pot(self, name):
        short = name[4:] if name.startswith("sep_") else name
        return S.fn_address(self.lib, self.prefix + short)

    def force_pairs(self, ptr, types, cf, fun, sysp, retp, opt, nneighb):
        return self.lib.orc_force_pairs(ptr, types, cf, fun, sysp, retp, opt, nneighb)

    def force_lj(self, ptr, types, prm, sysp, retp, opt, nneighb):
        return self.lib.orc_force_lj(ptr, types, prm, sysp, retp, opt, nneighb)

    def coulomb_sf(self, ptr, cf, sysp, retp, opt):
        self.lib.orc_coulomb_sf(ptr, cf, sysp, retp, opt)

    def torsion(self, ptr, type_, g, sysp, retp):
        self.lib.orc_torsion_ryckaert(ptr, type_, g, sysp, retp)

    def read_topology(self, ptr, path, sysp):
        rc = self.lib.orc_read_topology(ptr, path.encode(), sysp)
        if rc:
            raise RuntimeError("oracle could not read " + path)

    def omp_make_neighb(self, ptr, cf, sysp, opt, nneighb):
        return self.lib.orc_omp_make_neighb(ptr, cf, sysp, opt, nneighb)
