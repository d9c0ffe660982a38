"""Drivers shared by the golden-vector writer (make_golden.py, which runs
them on the compiled reference) and the tests (which run them on the oracle
and on the product and compare with the stored vectors).  Each driver takes
any library with the seppart ABI as ``api`` and returns plain arrays; the
call sequences follow the reference's own programs prgs/prg0.c (LJ),
prgs/prg2.c (butane) and prgs/prg3.c (water)."""
import ctypes as C
import hashlib

import numpy as np

from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem


def pair_keys(orc_lib, system, cap_per_atom=256):
    """Sorted keys i * N + j (i < j) of a system's host half lists, read by
    the oracle's orc_half_pairs."""
    n = system.n
    cap = n * cap_per_atom
    buf = np.empty(2 * cap, dtype=np.int32)
    c = orc_lib.orc_half_pairs(system.atoms, n, buf.ctypes.data, cap)
    assert c >= 0
    a = buf[:2 * c].reshape(-1, 2).astype(np.int64)
    k = a[:, 0] * n + a[:, 1]
    k.sort()
    return k


def pair_hash(keys):
    return int.from_bytes(hashlib.sha256(keys.tobytes()).digest()[:8], "little") >> 1


def ret_vec(s):
    """epot, ekin, ecoul, pot_P (9), kin_P (9), pot_P_bond (9)."""
    r = s.ret_dict()
    return np.concatenate([[r["epot"], r["ekin"], r["ecoul"]], r["pot_P"].ravel(),
                           r["kin_P"].ravel(), r["pot_P_bond"].ravel()])


def read_xyz(path):
    """The sep_init_xyz text format: N, box, then type x y z vx vy vz m q."""
    with open(path) as f:
        n = int(f.readline())
        L = np.array([float(v) for v in f.readline().split()[:3]])
        t, x, v, m, q = [], [], [], [], []
        for _ in range(n):
            p = f.readline().split()
            t.append(p[0].encode())
            x.append([float(u) for u in p[1:4]])
            v.append([float(u) for u in p[4:7]])
            m.append(float(p[7]))
            q.append(float(p[8]))
    return dict(x=np.array(x), v=np.array(v), m=np.array(m), q=np.array(q),
                type=np.array(t), L=L, mol=np.full(n, -1))


def read_top(path):
    """[ bonds ] / [ angles ] / [ dihedrals ] rows (mol, atoms..., type)."""
    out = {"bonds": [], "angles": [], "dihedrals": []}
    key = None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith(";"):
                continue
            if s.startswith("["):
                key = s.strip("[] ").split()[0]
                continue
            out[key].append([int(v) for v in s.split()])
    return {k: np.array(v, dtype=np.int64) for k, v in out.items() if v}


def subset(w, top, nmol, per):
    """The first nmol molecules of per atoms each."""
    n = nmol * per
    s = {k: (v[:n] if isinstance(v, np.ndarray) and len(v) == len(w["m"]) else v)
         for k, v in w.items()}
    for k, rows in top.items():
        s[k] = rows[rows[:, 0] < nmol]
    return s


def workload_from(fx, keys=("x", "v", "m", "q", "type", "L", "bonds", "angles", "dihedrals")):
    """Rebuilds a workload dict from a loaded fixture."""
    w = {k: np.array(fx[k]) for k in keys if k in fx}
    n = len(w["m"])
    w.setdefault("q", np.zeros(n))
    w.setdefault("type", np.full(n, b"A"))
    w["mol"] = np.full(n, -1)
    return w


def term_run(api, w, cf, update, calls, skin=None, prepare=None):
    """One sepret/force reset, then each call in turn: the force each call
    added and the sepret vector after it."""
    s = SepSystem.from_workload(api, w, cf=cf, dt=0.001, update=update, skin=skin)
    try:
        if prepare:
            prepare(s)
        s.reset_retval()
        s.reset_force()
        forces, rets = [], []
        prev = np.zeros((s.n, 3))
        for c in calls:
            c(s)
            f = np.array(s.view["f"])
            forces.append(f - prev)
            prev = f
            rets.append(ret_vec(s))
    finally:
        s.close()
    return np.array(forces), np.array(rets)


# ---- LJ NVE, 1000 atoms on a simple-cubic lattice (prgs/prg0.c) -------------

def lj_sc1000():
    return W.lj_sc(10, 0.8442, 0.728, seed=1)


def run_nve(api, orc_lib, w, nsteps, fsteps=(), prepare=None):
    """sep_force_pairs("AA", 2.5, sep_lj_shift) + sep_leapfrog per step.
    Returns per-step scalars (epot, ekin, max_dist2, pot_P, kin_P, tnow),
    flags (neighb_flag, nupdate_neighb, rebuilt), half-list sizes, pair-set
    hashes at rebuild steps (-1 elsewhere) and forces at fsteps."""
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.005, skin=0.5)
    scal = np.zeros((nsteps, 22))
    flags = np.zeros((nsteps, 3), dtype=np.int64)
    hashes = np.full(nsteps, -1, dtype=np.int64)
    npairs = np.zeros(nsteps, dtype=np.int64)
    fsnap = []
    try:
        if prepare:
            prepare(s)
        for step in range(nsteps):
            rebuild = s.sys.neighb_flag == 1
            s.reset_retval()
            s.reset_force()
            s.force_pairs("AA", 2.5, "sep_lj_shift")
            keys = pair_keys(orc_lib, s)
            npairs[step] = len(keys)
            if rebuild:
                hashes[step] = pair_hash(keys)
            if step in fsteps:
                fsnap.append(np.array(s.view["f"]))
            s.leapfrog()
            r = s.ret_dict()
            scal[step] = np.concatenate([[r["epot"], r["ekin"], s.sys.max_dist2],
                                         r["pot_P"].ravel(), r["kin_P"].ravel(),
                                         [s.sys.tnow]])
            flags[step] = [s.sys.neighb_flag, s.sys.nupdate_neighb, int(rebuild)]
    finally:
        s.close()
    return scal, flags, hashes, npairs, np.array(fsnap)


def run_nvt(api, w, nsteps, prepare=None):
    """As run_nve with sep_nosehoover(T0 = 0.728, Q = 10) before each
    leapfrog: per step alpha, ekin, epot."""
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.005, skin=0.5)
    out = np.zeros((nsteps, 3))
    try:
        if prepare:
            prepare(s)
        for step in range(nsteps):
            s.reset_retval()
            s.reset_force()
            s.force_pairs("AA", 2.5, "sep_lj_shift")
            s.nosehoover(0.728, 10.0)
            s.leapfrog()
            out[step] = [s.alpha.value, s.ret.ekin, s.ret.epot]
    finally:
        s.close()
    return out


# ---- water (prgs/prg3.c parameters) ------------------------------------------

WATER = dict(lbond=0.316, kspring=68421.0, angle=1.97, kangle=490.0, cf=2.9)
WATER_TERMS = ("lj_OO", "bond", "angle_harmonic", "angle_cossq", "coulomb_sf")


def water_calls():
    p = WATER
    return [
        lambda s: s.force_pairs("OO", 2.5, "sep_lj_shift", S.SEP_EXCL_SAME_MOL),
        lambda s: s.stretch_harmonic(0, p["lbond"], p["kspring"]),
        lambda s: s.angle_harmonic(0, p["angle"], p["kangle"]),
        lambda s: s.angle_cossq(0, p["angle"], p["kangle"]),
        lambda s: s.coulomb_sf(p["cf"], S.SEP_EXCL_SAME_MOL),
    ]


def run_water(api, w, update, prepare=None):
    return term_run(api, w, WATER["cf"], update, water_calls(), prepare=prepare)


# ---- butane (prgs/prg2.c parameters) -------------------------------------------

BUTANE = dict(lbond=0.407, kspring=2074.0, angle=1.90, kangle=400.0,
              rb=[15.5000, 20.3050, -21.9170, -5.1150, 43.8340, -52.6070])
BUTANE_TERMS = ("lj_CC", "bond", "angle_harmonic", "angle_cossq", "torsion_rb")


def butane_calls():
    p = BUTANE
    return [
        lambda s: s.force_pairs("CC", 2.5, "sep_lj_shift", S.SEP_EXCL_SAME_MOL),
        lambda s: s.stretch_harmonic(0, p["lbond"], p["kspring"]),
        lambda s: s.angle_harmonic(0, p["angle"], p["kangle"]),
        lambda s: s.angle_cossq(0, p["angle"], p["kangle"]),
        lambda s: s.torsion(0, p["rb"]),
    ]


def run_butane(api, w, prepare=None):
    return term_run(api, w, 2.5, S.SEP_LLIST_NEIGHBLIST, butane_calls(), prepare=prepare)


# ---- Kob-Andersen mixture: sep_force_lj per type pair ---------------------------

KA = {"AA": [2.5, 1.0, 1.0, 1.0], "AB": [2.0, 1.5, 0.8, 1.0], "BB": [2.2, 0.5, 0.88, 1.0]}


def ka_workload():
    return W.ka_mixture(500, 1.2, 0.8, seed=5)


def run_ka(api, w, prepare=None):
    calls = [(lambda t: (lambda s: s.force_lj(t, KA[t])))(t) for t in ("AA", "AB", "BB")]
    return term_run(api, w, 2.5, S.SEP_LLIST_NEIGHBLIST, calls, skin=0.3, prepare=prepare)


# ---- OpenMP Model II terms (source/sepomp.c, prgs/prg5.c) -----------------------

def model2_water(s):
    p = WATER
    return [("pairs", b"OO", 2.5, s.api.pot("sep_lj_shift")),
            ("pairs_lj", b"OO", (C.c_double * 3)(2.5, 1.0, 1.0)),
            ("coulomb", p["cf"]),
            ("bond", 0, p["lbond"], p["kspring"]),
            ("angle", 0, p["angle"], p["kangle"])]


def model2_butane(s):
    p = BUTANE
    return [("pairs", b"CC", 2.5, s.api.pot("sep_lj_shift")),
            ("pairs_lj", b"CC", (C.c_double * 3)(2.5, 1.1, 0.9)),
            ("bond", 0, p["lbond"], p["kspring"]),
            ("angle", 0, p["angle"], p["kangle"]),
            ("torsion", 0, (C.c_double * 6)(*p["rb"]))]


def run_model2(api, w, cf, opt, terms, nneighb=S.SEP_NEIGHB):
    """sep_omp_make_neighb(cf, opt), then each Model II term into its own
    zeroed array: the arrays, and the atoms' f afterwards (which the terms
    must leave alone)."""
    s = SepSystem.from_workload(api, w, cf=cf, dt=0.001, nneighb=nneighb)
    try:
        f0 = np.array(s.view["f"])
        s.omp_make_neighb(cf, opt)
        out = [s.omp_term(t[0], *t[1:]) for t in terms(s)]
        return np.array(out), f0, np.array(s.view["f"])
    finally:
        s.close()


# ---- exclusion variants and list modes --------------------------------------
# source/sepprfrc.c:517-603 with :703-740 (linked cells, SEP_EXCL_BONDED:
# bond, angle or dihedral partners, each scan stopping at the first -1 in
# either atom), :306-343 (the O(N^2) list of SEP_NEIGHBLIST, whose bonded
# test is sep_bond_share alone), :19-92 (SEP_BRUTE, bonds alone) and
# source/sepcoulomb.c:21-93 (brute-force Coulomb, which drops a pair only
# when sep_bonded is exactly 1).

G6_BUTANE_MOL = 250  # 1000 atoms of the reference's butane input


def g6_butane(w_full, top):
    return subset(w_full, top, G6_BUTANE_MOL, 4)


def _pairs_call(types, opt):
    return lambda s: s.force_pairs(types, 2.5, "sep_lj_shift", opt)


def _coulomb_call(opt):
    return lambda s: s.coulomb_sf(WATER["cf"], opt)


# tag, system ("butane" or "water"), neighbour update mode, call
G6_CASES = (
    ("butane_llist_bonded", "butane", S.SEP_LLIST_NEIGHBLIST, _pairs_call("CC", S.SEP_EXCL_BONDED)),
    ("butane_llist_samemol", "butane", S.SEP_LLIST_NEIGHBLIST, _pairs_call("CC", S.SEP_EXCL_SAME_MOL)),
    ("butane_nlist_all", "butane", S.SEP_NEIGHBLIST, _pairs_call("CC", S.SEP_ALL)),
    ("butane_nlist_bonded", "butane", S.SEP_NEIGHBLIST, _pairs_call("CC", S.SEP_EXCL_BONDED)),
    ("butane_nlist_samemol", "butane", S.SEP_NEIGHBLIST, _pairs_call("CC", S.SEP_EXCL_SAME_MOL)),
    ("butane_brute_bonded", "butane", S.SEP_BRUTE, _pairs_call("CC", S.SEP_EXCL_BONDED)),
    ("water_llist_bonded", "water", S.SEP_LLIST_NEIGHBLIST, _pairs_call("OH", S.SEP_EXCL_BONDED)),
    ("water_nlist_bonded", "water", S.SEP_NEIGHBLIST, _pairs_call("OH", S.SEP_EXCL_BONDED)),
    ("water_nlist_all", "water", S.SEP_NEIGHBLIST, _pairs_call("OO", S.SEP_ALL)),
    ("water_brute_bonded", "water", S.SEP_BRUTE, _pairs_call("OH", S.SEP_EXCL_BONDED)),
    ("water_brute_coulomb_bonded", "water", S.SEP_BRUTE, _coulomb_call(S.SEP_EXCL_BONDED)),
    ("water_brute_coulomb_all", "water", S.SEP_BRUTE, _coulomb_call(S.SEP_ALL)),
)


def g6_workload(g, system):
    """The G6 input of one system (butane or water) with its topology, from
    the loaded fixture."""
    w = {k[len(system) + 1:]: np.array(g[k]) for k in g.files if k.startswith(system + "_")}
    w["mol"] = np.full(len(w["m"]), -1)
    return w


def g6_cf(system):
    return 2.5 if system == "butane" else WATER["cf"]


def run_g6_case(api, orc_lib, w, system, update, call, prepare=None):
    """One force call after a reset: the force, the sepret vector and, for
    the list modes, the size and hash of the host half-list multiset
    (read by the oracle's orc_half_pairs)."""
    s = SepSystem.from_workload(api, w, cf=g6_cf(system), dt=0.001, update=update)
    try:
        if prepare:
            prepare(s)
        s.reset_retval()
        s.reset_force()
        call(s)
        f = np.array(s.view["f"])
        r = ret_vec(s)
        if update == S.SEP_BRUTE:
            return f, r, -1, -1, None
        keys = pair_keys(orc_lib, s, cap_per_atom=2048)
        return f, r, len(keys), pair_hash(keys), keys
    finally:
        s.close()
