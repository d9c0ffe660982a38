"""The drop-in boundary and the host logic of libsep.so, on the CPU (no
call here reaches the device: sep_init registers the array and the device
state is created by the first force call).

  * struct layouts of include/sepstrct.h equal the reference's (sizes and
    field offsets recorded from the reference header by make_golden.py);
  * libsep.so loads and exports every function include/*.h declares;
  * sep_sys_setup, sep_set_skin, the pressure/temperature helpers and the
    topology reader agree with the oracle and with their closed forms.
"""
import ctypes as C
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from seplib_amd import sep as S
from tests.golden import runs as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
INCLUDE = os.path.join(ROOT, "include")


def test_struct_layout_matches_reference():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "probe")
        subprocess.check_call(["gcc", "-std=c99", "-I", INCLUDE,
                               os.path.join(GOLD, "abi_probe.c"), "-o", exe])
        own = json.loads(subprocess.check_output([exe]).decode())
    with open(os.path.join(GOLD, "abi_layout.json")) as f:
        ref = json.load(f)
    assert own == ref
    assert ref["sizeof.seppart"] == 568


def _declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(sep_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    names = _declared_functions()
    assert {"sep_force_pairs", "sep_force_lj", "sep_coulomb_sf", "sep_leapfrog",
            "sep_nosehoover", "sep_stretch_harmonic", "sep_angle_harmonic",
            "sep_angle_cossq", "sep_torsion_Ryckaert", "sep_gpu_run"} <= names
    lib = C.CDLL(S.PRODUCT_LIB)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", S.PRODUCT_LIB]).decode()
    defined = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert names <= defined


@pytest.mark.parametrize("L,cf", [(10.580793, 2.5), (67.183848, 2.5), (6.888070, 2.9)])
def test_sys_setup_matches_oracle_and_rule(product, oracle, L, cf):
    n = 1000
    a = product.sys_setup([L, L, L], cf, 0.005, n, S.SEP_LLIST_NEIGHBLIST)
    b = oracle.sys_setup([L, L, L], cf, 0.005, n, S.SEP_LLIST_NEIGHBLIST)
    for f, _ in S.sepsys._fields_:
        if f in ("molptr",):
            continue
        va, vb = getattr(a, f), getattr(b, f)
        if hasattr(va, "__len__"):
            va, vb = list(va), list(vb)
        assert va == vb, f
    # source/sepinit.c:246-305: nsubbox = (int)(L / (cf + 0.25)), skin 0.25
    assert a.nsubbox[0] == int(L / (cf + 0.25))
    assert a.lsubbox[0] == L / a.nsubbox[0]
    assert a.skin == 0.25 and a.neighb_flag == 1 and a.ndof == 3 * n - 3
    product.lib.sep_set_skin(C.byref(a), C.c_double(0.3))
    assert a.skin == 0.3 and a.nsubbox[0] == int(L / (cf + 0.25))
    product.free_sys(C.byref(a))
    oracle.free_sys(C.byref(b))


def test_pressure_and_temperature_helpers(product):
    lib = product.lib
    n = 512
    sysv = product.sys_setup([9.0, 9.5, 10.0], 2.5, 0.005, n, S.SEP_LLIST_NEIGHBLIST)
    ret = S.sepret()
    rng = np.random.default_rng(3)
    kin = rng.normal(size=(3, 3))
    pot = rng.normal(size=(3, 3))
    for k in range(3):
        for kk in range(3):
            ret.kin_P[k][kk] = kin[k, kk]
            ret.pot_P[k][kk] = pot[k, kk]
    ret.ekin = 700.0
    lib.sep_pressure_tensor(C.byref(ret), C.byref(sysv))
    V = 9.0 * 9.5 * 10.0
    P = (kin + pot) / V
    assert np.allclose(S.mat(ret.P), P, rtol=1e-15, atol=0)
    assert abs(ret.p - np.trace(P) / 3.0) <= 1e-15 * abs(np.trace(P))
    lib.sep_get_pressure.restype = C.c_double
    lib.sep_get_temperature.restype = C.c_double
    p = lib.sep_get_pressure(C.byref(ret), C.byref(sysv))
    assert abs(p - np.trace(P) / 3.0) <= 1e-15 * abs(p)
    T = lib.sep_get_temperature(C.byref(ret), C.byref(sysv))
    assert abs(T - 2.0 / (3.0 * (3 * n - 3)) * 700.0) <= 1e-15 * T
    product.free_sys(C.byref(sysv))


def _topology_arrays(api, w):
    from seplib_amd.sim import SepSystem
    s = SepSystem.from_workload(api, w, cf=2.9, dt=0.001)
    try:
        v = s.view
        out = {k: np.array(v[k]) for k in ("bond", "angle", "dihed", "molindex")}
        mp = s.sys.molptr.contents
        out["num"] = (mp.num_mols, mp.num_bonds, mp.num_angles, mp.num_dihedrals)
        out["blist"] = np.ctypeslib.as_array(mp.blist, (3 * mp.num_bonds,)).copy()
        out["alist"] = np.ctypeslib.as_array(mp.alist, (4 * mp.num_angles,)).copy()
        if mp.num_dihedrals:
            out["dlist"] = np.ctypeslib.as_array(mp.dlist, (5 * mp.num_dihedrals,)).copy()
    finally:
        s.close()
    return out


@pytest.mark.parametrize("fixture", ["g2_water.npz", "g3_butane.npz"])
def test_topology_reader_matches_oracle(product, oracle, fixture):
    g = np.load(os.path.join(GOLD, fixture), allow_pickle=False)
    w = R.workload_from(g)
    a, b = _topology_arrays(product, w), _topology_arrays(oracle, w)
    assert a["num"] == b["num"] and a["num"][1] == len(w["bonds"])
    for k in b:
        if k != "num":
            assert np.array_equal(a[k], b[k]), k


def _saved_text(api, w, types, tmp_path, name):
    from seplib_amd.sim import SepSystem
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.005)
    try:
        path = str(tmp_path / name)
        api.lib.sep_save_xyz(s.atoms, types, path.encode(), b"w", s.sysp)
        with open(path) as f:
            return f.read()
    finally:
        s.close()


def test_save_xyz_matches_reference_format(product, tmp_path):
    """sep_save_xyz writes the sep_init_xyz format (source/sepmisc.c:536-573);
    reading it back gives the same configuration to 15 decimals."""
    g = np.load(os.path.join(GOLD, "g2_water.npz"), allow_pickle=False)
    w = R.workload_from(g)
    text = _saved_text(product, w, b"OH", tmp_path, "a.xyz")
    back = R.read_xyz(str(tmp_path / "a.xyz"))
    assert len(back["m"]) == len(w["m"])
    assert np.abs(back["x"] - w["x"]).max() <= 1e-15 * np.abs(w["x"]).max() + 1e-15
    assert np.abs(back["v"] - w["v"]).max() <= 1e-12 * max(1.0, np.abs(w["v"]).max())
    assert np.array_equal(back["type"], w["type"])
    assert text.splitlines()[0] == str(len(w["m"]))
    only_o = _saved_text(product, w, b"O", tmp_path, "o.xyz")
    assert only_o.splitlines()[0] == str(int((w["type"] == b"O").sum()))


def test_save_xyz_byte_identical_to_reference(product, reference, tmp_path):
    g = np.load(os.path.join(GOLD, "g3_butane.npz"), allow_pickle=False)
    w = R.workload_from(g)
    assert (_saved_text(product, w, b"C", tmp_path, "p.xyz") ==
            _saved_text(reference, w, b"C", tmp_path, "r.xyz"))


def _pointer_mask(n):
    """Byte mask of n seppart records: the named fields, without the
    pointer fields (addresses of the process that wrote the file), sigma
    (which neither library's sep_init sets) and the struct padding (heap
    bytes)."""
    from seplib_amd import sep as S
    rec = C.sizeof(S.seppart)
    keep = np.zeros(rec, dtype=bool)
    for name, _ in S.seppart._fields_:
        if name in ("neighb", "collid", "colltime", "sigma"):
            continue
        fld = getattr(S.seppart, name)
        keep[fld.offset:fld.offset + fld.size] = True
    return np.tile(keep, n)


def _saved_binary(api, w, tmp_path, name, crossings=True):
    from seplib_amd.sim import SepSystem
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.005)
    try:
        if crossings:
            for i in range(0, s.n, 7):
                s.atoms[i].crossings[1] = i % 5 - 2
                s.atoms[i].xn[2] = 0.25 * i
        path = str(tmp_path / name)
        api.lib.sep_save(s.atoms, s.n, path.encode())
        with open(path, "rb") as f:
            return np.frombuffer(f.read(), dtype=np.uint8), s.n
    finally:
        s.close()


def test_save_byte_identical_to_reference(product, reference, tmp_path):
    """sep_save writes the raw seppart records (source/sepmisc.c:319-330):
    the same bytes as the compiled reference except the pointer fields."""
    g = np.load(os.path.join(GOLD, "g3_butane.npz"), allow_pickle=False)
    w = R.workload_from(g)
    a, n = _saved_binary(product, w, tmp_path, "p.bin")
    b, _ = _saved_binary(reference, w, tmp_path, "r.bin")
    assert len(a) == len(b) == n * C.sizeof(__import__("seplib_amd.sep", fromlist=["x"]).seppart)
    m = _pointer_mask(n)
    assert np.array_equal(a[m], b[m])


def test_load_round_trip(product, tmp_path):
    """sep_load restores every non-pointer field sep_save wrote and gives
    each atom a fresh neighbour array (source/sepmisc.c:333-354)."""
    from seplib_amd.sim import SepSystem
    g = np.load(os.path.join(GOLD, "g2_water.npz"), allow_pickle=False)
    w = R.workload_from(g)
    saved, n = _saved_binary(product, w, tmp_path, "w.bin")
    w2 = dict(w)
    w2["x"] = np.mod(w["x"] + 0.37, w["L"])
    s = SepSystem.from_workload(product, w2, cf=2.5, dt=0.005)
    try:
        old = [C.cast(s.atoms[i].neighb, C.c_void_p).value for i in range(s.n)]
        product.lib.sep_load(s.atoms, 64, s.n, str(tmp_path / "w.bin").encode())
        back = np.frombuffer(C.string_at(C.addressof(s.atoms[0]), len(saved)), dtype=np.uint8)
        m = _pointer_mask(n)
        assert np.array_equal(back[m], saved[m])
        for i in range(s.n):
            assert s.atoms[i].neighb[0] == -1
        assert any(C.cast(s.atoms[i].neighb, C.c_void_p).value != old[i] for i in range(s.n))
    finally:
        s.close()


def test_device_atom_bound_is_refused_before_any_device_call():
    """sepdev_create refuses more than 2^27 - 64 atoms (per-atom records are
    reached by 32-bit byte offsets), and none, with SEP_GPU_EARG before it
    touches the device."""
    lib = C.CDLL(S.PRODUCT_LIB)
    lib.sepdev_create.restype = C.c_void_p
    lib.sepdev_create.argtypes = [C.c_long, C.POINTER(C.c_int)]
    err = C.c_int(0)
    assert lib.sepdev_create((1 << 27) - 63, C.byref(err)) is None
    assert err.value == S.SEP_GPU_EARG
    err.value = 0
    assert lib.sepdev_create(0, C.byref(err)) is None
    assert err.value == S.SEP_GPU_EARG
