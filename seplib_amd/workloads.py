"""Seeded synthetic inputs for the seplib hot-path workloads.

All randomness comes from a SplitMix64 stream (Steele, Lea, Flood 2014)
implemented here with numpy integer arithmetic, so a (seed, size) pair
gives the same inputs on every machine.

Workloads (reduced LJ units):
  lj_sc      simple-cubic LJ liquid start (N = k^3), e.g. N=1000 at 0.8442
  lj_fcc     fcc LJ start, 4 k^3 atoms, e.g. N=256000 at 0.8442
  water      3-site flexible water molecules on a cubic lattice with random
             orientations (O-H 0.316, H-O-H 1.91 rad)
  ka_mixture Kob-Andersen 80:20 A:B on a jittered lattice
"""
import math

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(seed, count):
    """count uint64 values of the SplitMix64 stream started at seed."""
    out = np.empty(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        out[:] = z ^ (z >> np.uint64(31))
    return out


def uniform(seed, count):
    """Doubles in [0, 1) from the top 53 bits of the stream."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed, count):
    """Standard normals by the Box-Muller transform."""
    m = (count + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[:m]  # (0, 1]
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:count]


def box_length(n, rho):
    return (n / rho) ** (1.0 / 3.0)


def maxwell_boltzmann(n, temp, mass, seed):
    """Velocities with zero total momentum, scaled so sum m v^2 = 3 N T."""
    if temp <= 0.0:
        return np.zeros((n, 3))
    v = normal(seed, 3 * n).reshape(n, 3) * np.sqrt(temp / mass)[:, None]
    p = (mass[:, None] * v).sum(axis=0)
    v -= p / mass.sum()
    s = (mass[:, None] * v * v).sum()
    v *= math.sqrt(3.0 * n * temp / s)
    return v


def lj_sc(k, rho, temp, seed=1):
    """Simple-cubic start of N = k^3 atoms: dict with x, v, m, q, type, L."""
    n = k ** 3
    L = box_length(n, rho)
    a = L / k
    g = np.arange(k)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pos = (np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) + 0.5) * a
    m = np.ones(n)
    return dict(x=pos, v=maxwell_boltzmann(n, temp, m, seed), m=m, q=np.zeros(n),
                type=np.full(n, b"A"), L=np.array([L, L, L]), mol=np.full(n, -1))


def lj_fcc(k, rho, temp, seed=2):
    """fcc start of N = 4 k^3 atoms."""
    n = 4 * k ** 3
    L = box_length(n, rho)
    a = L / k
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    g = np.arange(k)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    pos = ((cells[:, None, :] + basis[None, :, :]) * a).reshape(n, 3)
    m = np.ones(n)
    return dict(x=pos, v=maxwell_boltzmann(n, temp, m, seed), m=m, q=np.zeros(n),
                type=np.full(n, b"A"), L=np.array([L, L, L]), mol=np.full(n, -1))


def _random_rotations(count, seed):
    """Uniform random rotation matrices from unit quaternions."""
    q = normal(seed, 4 * count).reshape(count, 4)
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


def water(nmol, rho_atoms, temp, seed=3, l0=0.316, theta0=1.91,
          qO=-0.8476, qH=0.4238, mO=1.0, mH=0.0625):
    """3-site water: O, H, H per molecule (atoms 3i, 3i+1, 3i+2).

    Returns the atom dict plus bonds (mol, a, b, type) and angles
    (mol, a, b, c, type) with the H-O-H angle centred on O."""
    n = 3 * nmol
    L = box_length(n, rho_atoms)
    k = int(math.ceil(nmol ** (1.0 / 3.0)))
    a = L / k
    g = np.arange(k)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    sites = (np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) + 0.5) * a
    if len(sites) > nmol:
        keep = np.argsort(uniform(seed + 17, len(sites)))[:nmol]
        sites = sites[np.sort(keep)]
    h = theta0 / 2.0
    local = np.array([[0.0, 0.0, 0.0],
                      [l0 * math.sin(h), 0.0, l0 * math.cos(h)],
                      [-l0 * math.sin(h), 0.0, l0 * math.cos(h)]])
    R = _random_rotations(nmol, seed + 29)
    pos = (sites[:, None, :] + np.einsum("mij,sj->msi", R, local)).reshape(n, 3)
    pos = np.mod(pos, L)
    m = np.tile([mO, mH, mH], nmol).astype(np.float64)
    q = np.tile([qO, qH, qH], nmol).astype(np.float64)
    types = np.tile(np.array([b"O", b"H", b"H"]), nmol)
    mol = np.repeat(np.arange(nmol), 3)
    base = 3 * np.arange(nmol)
    bonds = np.concatenate([np.stack([np.arange(nmol), base, base + 1, 0 * base], 1),
                            np.stack([np.arange(nmol), base, base + 2, 0 * base], 1)])
    bonds = bonds[np.lexsort((bonds[:, 2], bonds[:, 0]))]
    angles = np.stack([np.arange(nmol), base + 1, base, base + 2, 0 * base], 1)
    return dict(x=pos, v=maxwell_boltzmann(n, temp, m, seed), m=m, q=q, type=types,
                L=np.array([L, L, L]), mol=mol, bonds=bonds, angles=angles)


def ka_mixture(n, rho, temp, seed=5, frac_a=0.8):
    """Kob-Andersen mixture start on a jittered simple-cubic lattice."""
    L = box_length(n, rho)
    k = int(math.ceil(n ** (1.0 / 3.0)))
    a = L / k
    g = np.arange(k)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    sites = (np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) + 0.5) * a
    order = np.argsort(uniform(seed, len(sites)))[:n]
    pos = sites[np.sort(order)] + (uniform(seed + 1, 3 * n).reshape(n, 3) - 0.5) * 0.1 * a
    pos = np.mod(pos, L)
    nb = n - int(round(frac_a * n))
    types = np.full(n, b"A")
    types[np.sort(np.argsort(uniform(seed + 2, n))[:nb])] = b"B"
    m = np.ones(n)
    return dict(x=pos, v=maxwell_boltzmann(n, temp, m, seed + 3), m=m, q=np.zeros(n),
                type=types, L=np.array([L, L, L]), mol=np.full(n, -1))


def random_packed(n, rho, temp, seed=7):
    """One-type LJ start: n of the sites of the smallest simple-cubic lattice
    with at least n sites, chosen at random and jittered by up to 5 % of
    the spacing (no two atoms closer than 0.9 spacings)."""
    L = box_length(n, rho)
    k = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
    a = L / k
    g = np.arange(k)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    sites = (np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) + 0.5) * a
    order = np.argsort(uniform(seed, len(sites)))[:n]
    pos = sites[np.sort(order)] + (uniform(seed + 1, 3 * n).reshape(n, 3) - 0.5) * 0.1 * a
    pos = np.mod(pos, L)
    m = np.ones(n)
    return dict(x=pos, v=maxwell_boltzmann(n, temp, m, seed + 3), m=m, q=np.zeros(n),
                type=np.full(n, b"A"), L=np.array([L, L, L]), mol=np.full(n, -1))


def write_xyz(path, w):
    """Writes the sep_init_xyz text format: N, box, type x y z vx vy vz m q."""
    n = len(w["m"])
    with open(path, "w") as f:
        f.write("%d\n%.15f %.15f %.15f\n" % (n, *w["L"]))
        for i in range(n):
            t = w["type"][i].decode()
            f.write("%s %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                t, *w["x"][i], *w["v"][i], w["m"][i], w["q"][i]))


def write_top(path, w):
    """Writes bonds/angles/dihedrals in the .top format read by
    sep_read_topology_file."""
    with open(path, "w") as f:
        for key, head in (("bonds", "[ bonds ]"), ("angles", "[ angles ]"),
                          ("dihedrals", "[ dihedrals ]")):
            rows = w.get(key)
            if rows is None:
                continue
            f.write("%s\n; generated by seplib_amd.workloads\n" % head)
            for r in rows:
                f.write(" ".join(str(int(v)) for v in r) + "\n")
            f.write("\n")
