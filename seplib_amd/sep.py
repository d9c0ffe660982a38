"""ctypes mirror of the seplib C99 API, for the MI355X library and checkers.

The structures below repeat include/sepstrct.h field by field (the
reference's include/sepstrct.h:23-204 layout), so one set of definitions
serves three shared libraries with the same ABI:

  * ``load_product()`` -> seplib_amd/libsep.so, the MI355X library;
  * the checkers (the compiled reference and the C oracle), loaded by
    oracle/checkers.py for tests and benchmarks only.

``atoms_view`` exposes an atom array as a numpy structured array over the
same memory, so tests read and write x, v, f ... without copies.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PRODUCT_LIB = os.path.join(HERE, "libsep.so")

# constants of include/sepdef.h
SEP_BOND, SEP_ANGLE, SEP_DIHED = 10, 10, 20
SEP_NEIGHB = 3000
SEP_ALL, SEP_EXCL_BONDED, SEP_EXCL_SAME_MOL = 1, 2, 3
SEP_BRUTE, SEP_NEIGHBLIST, SEP_LLIST_NEIGHBLIST = 0, 1, 2
SEP_LJCF2 = 0.016316891136

# sepgpu.h
SEP_GPU_OK = 0
SEP_GPU_EDEVICE, SEP_GPU_EUNSUPPORTED, SEP_GPU_ENEIGHB = 1, 2, 3
SEP_GPU_ECELL, SEP_GPU_EARG, SEP_GPU_EHOST, SEP_GPU_ESTATE = 4, 5, 6, 7
SEP_GPU_FP64, SEP_GPU_MIXED = 0, 1
(OP_PAIRS, OP_LJ, OP_COULOMB, OP_BOND, OP_ANGLE, OP_COSSQ, OP_TORSION,
 OP_NOSEHOOVER, OP_LEAPFROG) = range(1, 10)
SEP_GPU_RUN_GRAPH = 1
SEP_GPU_RUN_CONCURRENT = 2

d3 = C.c_double * 3
i3 = C.c_int * 3
m33 = (C.c_double * 3) * 3


class seppart(C.Structure):
    _fields_ = [
        ("x", d3), ("v", d3), ("f", d3), ("a", d3),
        ("m", C.c_double), ("type", C.c_char), ("z", C.c_double),
        ("neighb", C.POINTER(C.c_int)),
        ("cross_neighb", i3), ("crossings", i3),
        ("molindex", C.c_int),
        ("bond", C.c_int * SEP_BOND), ("angle", C.c_int * SEP_ANGLE),
        ("dihed", C.c_int * SEP_DIHED),
        ("sigma", C.c_double), ("collid", C.POINTER(C.c_int)),
        ("colltime", C.POINTER(C.c_double)), ("ldiff", C.c_double),
        ("xtrue", d3), ("x0", d3), ("xn", d3), ("xp", d3), ("px", d3),
        ("pv", d3), ("pa", d3), ("randn", d3), ("prevf", d3),
    ]


class sepmolinfo(C.Structure):
    _fields_ = [
        ("num_mols", C.c_uint), ("max_nuau", C.c_uint),
        ("flag_bonds", C.c_int), ("flag_angles", C.c_int), ("flag_dihedrals", C.c_int),
        ("num_bonds", C.c_uint), ("blist", C.POINTER(C.c_uint)), ("num_btypes", C.c_uint),
        ("num_angles", C.c_uint), ("alist", C.POINTER(C.c_uint)), ("num_atypes", C.c_uint),
        ("num_dihedrals", C.c_uint), ("dlist", C.POINTER(C.c_uint)), ("num_dtypes", C.c_uint),
        ("blengths", C.POINTER(C.c_double)), ("angles", C.POINTER(C.c_double)),
        ("dihedrals", C.POINTER(C.c_double)),
        ("flag_Fij", C.c_uint), ("Fij", C.c_void_p), ("Fiajb", C.c_void_p),
    ]


class sepsys(C.Structure):
    _fields_ = [
        ("npart", C.c_long), ("length", d3), ("volume", C.c_double),
        ("intgr_type", C.c_int), ("dt", C.c_double), ("tnow", C.c_double),
        ("ndof", C.c_uint), ("max_dist2", C.c_double),
        ("cf", C.c_double), ("lsubbox", d3), ("nsubbox", i3), ("skin", C.c_double),
        ("neighb_update", C.c_uint), ("neighb_flag", C.c_uint),
        ("nupdate_neighb", C.c_uint),
        ("omp_flag", C.c_bool), ("nthreads", C.c_uint), ("fun_cstate", C.c_int),
        ("molptr", C.POINTER(sepmolinfo)),
    ]


class sepret(C.Structure):
    _fields_ = [
        ("etot", C.c_double), ("ekin", C.c_double), ("epot", C.c_double),
        ("ecoul", C.c_double), ("sumv2", C.c_double),
        ("P", m33), ("kin_P", m33), ("pot_P", m33), ("p", C.c_double),
        ("P_mol", m33), ("kin_P_mol", m33), ("pot_P_mol", m33), ("p_mol", C.c_double),
        ("pot_P_conservative", m33), ("pot_P_random", m33),
        ("pot_P_dissipative", m33), ("pot_P_bond", m33),
        ("pot_T_mol", m33), ("kin_T_mol", m33), ("T_mol", m33), ("t_mol", C.c_double),
    ]


class sep_gpu_op(C.Structure):
    _fields_ = [
        ("op", C.c_int), ("types", C.c_char * 2), ("type", C.c_int), ("opt", C.c_uint),
        ("p", C.c_double * 6), ("fun", C.c_void_p), ("alpha", C.POINTER(C.c_double)),
    ]


POTFN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_char)
P_ATOMS = C.POINTER(seppart)
P_SYS = C.POINTER(sepsys)
P_RET = C.POINTER(sepret)
P_D = C.POINTER(C.c_double)


def _dtype_of(struct):
    """numpy structured dtype with the exact offsets of a ctypes Structure."""
    names, formats, offsets = [], [], []
    for name, ctype in struct._fields_:
        off = getattr(struct, name).offset
        if issubclass(ctype, C.Array):
            base, shape = ctype, []
            while issubclass(base, C.Array):
                shape.append(base._length_)
                base = base._type_
            fmt = (np.dtype(base), tuple(shape))
        elif ctype is C.c_char:
            fmt = np.dtype("S1")
        elif ctype in (C.c_double,):
            fmt = np.float64
        elif ctype in (C.c_int, C.c_uint, C.c_long, C.c_bool):
            fmt = np.dtype(ctype)
        else:
            fmt = np.uintp
        names.append(name)
        formats.append(fmt)
        offsets.append(off)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": C.sizeof(struct)})


SEPPART_DTYPE = _dtype_of(seppart)


def atoms_view(ptr, n):
    """Structured numpy view of n atoms at a seppart* (no copy)."""
    addr = C.cast(ptr, C.c_void_p).value
    buf = (C.c_char * (n * C.sizeof(seppart))).from_address(addr)
    return np.frombuffer(buf, dtype=SEPPART_DTYPE, count=n)


def mat(m33_value):
    return np.array([[m33_value[i][j] for j in range(3)] for i in range(3)])


def bind_api(lib, p):
    """Declares the common API signatures with name prefix p."""
    def f(name, restype, *args):
        fn = getattr(lib, p + name)
        fn.restype = restype
        fn.argtypes = list(args)
        return fn
    f("reset_retval", None, P_RET)
    f("pressure_tensor", None, P_RET, P_SYS)
    f("reset_force", None, P_ATOMS, P_SYS)
    f("nosehoover", None, P_ATOMS, C.c_double, P_D, C.c_double, P_SYS)
    f("leapfrog", None, P_ATOMS, P_SYS, P_RET)
    f("stretch_harmonic", None, P_ATOMS, C.c_int, C.c_double, C.c_double, P_SYS, P_RET)
    f("angle_harmonic", None, P_ATOMS, C.c_int, C.c_double, C.c_double, P_SYS, P_RET)
    f("angle_cossq", None, P_ATOMS, C.c_int, C.c_double, C.c_double, P_SYS, P_RET)
    for nm in ("lj", "lj_shift", "wca"):
        f(nm, C.c_double, C.c_double, C.c_char)


P_MAT = C.POINTER(C.POINTER(C.c_double))


def bind_omp(lib, p):
    """The OpenMP Model II term signatures (include/sepomp.h) with prefix p."""
    def f(name, *args):
        fn = getattr(lib, p + name)
        fn.restype = None
        fn.argtypes = list(args)
    f("omp_pairs", P_MAT, P_ATOMS, C.c_char_p, C.c_double, C.c_void_p, P_SYS)
    f("omp_pairs_lj", P_MAT, P_ATOMS, C.c_char_p, P_D, P_SYS)
    f("omp_coulomb", P_MAT, P_ATOMS, C.c_double, P_SYS)
    f("omp_bond", P_MAT, P_ATOMS, C.c_int, C.c_double, C.c_double, P_SYS)
    f("omp_angle", P_MAT, P_ATOMS, C.c_int, C.c_double, C.c_double, P_SYS)
    f("omp_torsion", P_MAT, P_ATOMS, C.c_int, P_D, P_SYS)


def row_pointers(a):
    """A double ** over the rows of a C-contiguous float64 (n, 3) array."""
    base = a.ctypes.data
    rows = (C.POINTER(C.c_double) * len(a))(
        *[C.cast(base + i * a.strides[0], C.POINTER(C.c_double)) for i in range(len(a))])
    return C.cast(rows, P_MAT), rows


def load_product(path=PRODUCT_LIB):
    """The MI355X library; raises if it has not been built."""
    if not os.path.exists(path):
        raise RuntimeError("libsep.so not built: run python -m seplib_amd.build")
    lib = C.CDLL(path)
    bind_api(lib, "sep_")
    bind_sep_names(lib)
    bind_omp(lib, "sep_")
    lib.sep_omp_make_neighb.restype = None
    lib.sep_omp_make_neighb.argtypes = [P_ATOMS, C.c_double, P_SYS, C.c_uint]
    f = lambda n, r, *a: (setattr(getattr(lib, n), "restype", r),
                          setattr(getattr(lib, n), "argtypes", list(a)))
    f("sep_gpu_last_error", C.c_int)
    f("sep_gpu_last_error_message", C.c_char_p)
    f("sep_gpu_clear_error", None)
    f("sep_gpu_set_precision", C.c_int, P_ATOMS, P_SYS, C.c_int)
    f("sep_gpu_resident", C.c_int, P_ATOMS, P_SYS, C.c_int)
    f("sep_gpu_sync", C.c_int, P_ATOMS, P_SYS)
    f("sep_gpu_upload", C.c_int, P_ATOMS, P_SYS)
    f("sep_gpu_run", C.c_int, P_ATOMS, P_SYS, P_RET, C.POINTER(sep_gpu_op), C.c_int,
      C.c_long, C.c_int)
    f("sep_gpu_neighbour_pairs", C.c_int, P_ATOMS, P_SYS, C.POINTER(C.c_int), C.c_long,
      C.POINTER(C.c_long))
    f("sep_gpu_timing_enable", C.c_int, P_ATOMS, C.c_int)
    f("sep_gpu_timing_reset", C.c_int, P_ATOMS)
    f("sep_gpu_timing_get", C.c_int, P_ATOMS, C.c_char_p, P_D, C.POINTER(C.c_long))
    f("sep_gpu_list_stats", C.c_int, P_ATOMS, P_D, C.POINTER(C.c_int))
    f("sep_gpu_time_pair_kernel", C.c_int, P_ATOMS, C.c_int, P_D)
    f("sep_gpu_force_lj_table", C.c_int, P_ATOMS, C.c_int, C.c_char_p, P_D, P_SYS, P_RET,
      C.c_uint)
    f("sep_gpu_lj_adjoint", C.c_int, P_ATOMS, P_SYS, C.c_int, C.c_char_p, P_D, P_D, P_D, P_D)
    f("sep_gpu_lj_adjoint_noalloc", C.c_int, P_ATOMS, P_SYS, C.c_int, C.c_char_p, P_D, P_D,
      P_D, P_D)
    f("sep_gpu_synchronize", C.c_int, P_ATOMS)
    f("sep_gpu_list_limit", C.c_int, P_ATOMS, P_SYS, C.c_int)
    f("sep_gpu_checkpoint", C.c_int, P_ATOMS, P_SYS, P_RET, P_D, C.c_char_p)
    f("sep_gpu_restore", C.c_int, P_ATOMS, P_SYS, P_RET, P_D, C.c_char_p)
    f("sep_gpu_record_steps", C.c_int, P_ATOMS, C.c_long)
    f("sep_gpu_read_steps", C.c_int, P_ATOMS, C.c_void_p, P_D, C.c_long,
      C.POINTER(C.c_long), C.POINTER(C.c_long))
    P_I = C.POINTER(C.c_int)
    f("sep_gpu_batch_create", C.c_void_p, C.c_int, C.POINTER(P_ATOMS), C.POINTER(P_SYS),
      C.c_char_p, C.c_double, P_D, P_D, P_D, C.c_int, C.c_int)
    f("sep_gpu_batch_free", None, C.c_void_p)
    f("sep_gpu_batch_run", C.c_int, C.c_void_p, C.c_long)
    f("sep_gpu_batch_run_eager", C.c_int, C.c_void_p, C.c_long)
    f("sep_gpu_batch_eval", C.c_int, C.c_void_p, C.c_int)
    f("sep_gpu_batch_sync", C.c_int, C.c_void_p, P_RET, P_D)
    f("sep_gpu_batch_synchronize", C.c_int, C.c_void_p)
    f("sep_gpu_batch_set_shape", C.c_int, C.c_void_p, P_I, C.c_int)
    f("sep_gpu_batch_get_shape", C.c_int, C.c_void_p, P_I, P_I)
    f("sep_gpu_batch_time_eval", C.c_int, C.c_void_p, P_I, C.c_int, P_D)
    f("sep_gpu_batch_tune", C.c_int, C.c_void_p, C.c_int, P_I, P_I)
    return lib


def bind_sep_names(lib):
    def f(name, restype, *args):
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = list(args)
    f("sep_init", P_ATOMS, C.c_size_t, C.c_size_t)
    f("sep_close", None, P_ATOMS, C.c_size_t)
    f("sep_init_xyz", P_ATOMS, P_D, C.POINTER(C.c_int), C.c_char_p, C.c_char)
    f("sep_sys_setup", sepsys, C.c_double, C.c_double, C.c_double, C.c_double,
      C.c_double, C.c_size_t, C.c_size_t)
    f("sep_free_sys", None, P_SYS)
    f("sep_set_skin", None, P_SYS, C.c_double)
    f("sep_set_omp", None, C.c_uint, P_SYS)
    f("sep_force_pairs", C.c_int, P_ATOMS, C.c_char_p, C.c_double, C.c_void_p, P_SYS,
      P_RET, C.c_uint)
    f("sep_force_lj", None, P_ATOMS, C.c_char_p, P_D, P_SYS, P_RET, C.c_uint)
    f("sep_coulomb_sf", None, P_ATOMS, C.c_double, P_SYS, P_RET, C.c_uint)
    f("sep_torsion_Ryckaert", None, P_ATOMS, C.c_int, P_D, P_SYS, P_RET)
    f("sep_read_topology_file", None, P_ATOMS, C.c_char_p, P_SYS, C.c_char)
    f("sep_get_pressure", C.c_double, P_RET, P_SYS)
    f("sep_get_temperature", C.c_double, P_RET, P_SYS)
    f("sep_save_xyz", None, P_ATOMS, C.c_char_p, C.c_char_p, C.c_char_p, P_SYS)
    f("sep_save", None, P_ATOMS, C.c_size_t, C.c_char_p)
    f("sep_load", None, P_ATOMS, C.c_int, C.c_size_t, C.c_char_p)


def fn_address(lib, name):
    """Address of an exported function, for sep_force_pairs' fun argument."""
    return C.cast(getattr(lib, name), C.c_void_p).value
