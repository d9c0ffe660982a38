"""Force matching on the device.

Multi-type Lennard-Jones forces F(x; theta) from a type-pair table (one
list pass, `sep_gpu_force_lj_table`) and the adjoint gradient of a loss
L(F) (`sep_gpu_lj_adjoint`), with a thin ctypes-backed
torch.autograd.Function for fitting and checking harnesses.

theta = (eps_p..., sigma_p...) over the unordered type pairs p = (a, b),
a <= b, in the order (0,0), (0,1), ..., (1,1), ...  The force is the one of
the reference's sep_force_lj (source/sepprfrc.c:785-840): with q = sigma^2/r^2,
f_ij = 48 eps (q^7 - aw/2 q^4) r_ij; cutoffs are fixed, not fitted.

    python -m seplib_amd.forcematch [--natoms 32000] [--frames 256] [--spacing 200]

runs the Kob-Andersen force-matching protocol on one GPU (frames from an
FP64 NVT run with the true parameters, a fit from a perturbed start) and
prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import sep as S
from . import workloads as W
from .sim import ProductAPI, SepSystem

# Kob-Andersen 80:20 (types A, B): eps_AA, eps_AB, eps_BB, sigma_AA, sigma_AB,
# sigma_BB; cutoffs 2.5 sigma_ab of these values
KA_TYPES = "AB"
KA_THETA = np.array([1.0, 1.5, 0.5, 1.0, 0.8, 0.88])
NNEIGHB = 192  # host half-list room per atom (about 55 are used at rho 1.2)


def pairs(nt):
    return [(a, b) for a in range(nt) for b in range(a, nt)]


class LJTable:
    """Type labels, fixed cutoffs per unordered pair and aw."""

    def __init__(self, types, rc, aw=1.0):
        self.types = types
        self.nt = len(types)
        self.rc = np.asarray(rc, dtype=np.float64)
        self.aw = aw
        self.pairs = pairs(self.nt)
        assert len(self.rc) == len(self.pairs)

    def params(self, theta):
        """params[(i*nt + k)*4 + {0,1,2,3}] = rc, eps, sigma, aw (symmetric)."""
        nt, npair = self.nt, len(self.pairs)
        out = np.zeros((nt, nt, 4))
        for p, (a, b) in enumerate(self.pairs):
            row = [self.rc[p], theta[p], theta[npair + p], self.aw]
            out[a, b] = row
            out[b, a] = row
        return np.ascontiguousarray(out.ravel())

    def column_params(self, theta, k):
        """A table whose forces are dF/dtheta_k exactly.  F is linear in
        eps_p, so dF/deps_p is the force of pair p alone at eps = 1; and
        dft/dsigma = (48 eps/sigma)(14 q^7 - 4 aw q^4) = 48 eps' (q^7 - aw'/2 q^4)
        with eps' = 14 eps/sigma and aw' = 4 aw/7, again an LJ force of pair
        p alone.  Every other pair gets eps = 0."""
        nt, npair = self.nt, len(self.pairs)
        p = k % npair
        out = np.zeros((nt, nt, 4))
        for q, (a, b) in enumerate(self.pairs):
            eps, aw = 0.0, self.aw
            if q == p:
                if k < npair:
                    eps = 1.0
                else:
                    eps, aw = 14.0 * theta[p] / theta[npair + p], 4.0 * self.aw / 7.0
            row = [self.rc[q], eps, theta[npair + q], aw]
            out[a, b] = row
            out[b, a] = row
        return np.ascontiguousarray(out.ravel())

    def fold(self, full):
        """dLdtheta[(i*nt+k)*2 + {0,1}] (both orders) -> theta layout."""
        full = np.asarray(full).reshape(self.nt, self.nt, 2)
        npair = len(self.pairs)
        out = np.zeros(2 * npair)
        for p, (a, b) in enumerate(self.pairs):
            out[p] = full[a, b, 0]
            out[npair + p] = full[a, b, 1]
        return out


class DeviceForces:
    """F(x; theta) and its adjoint for one system held by libsep.so."""

    def __init__(self, system, table):
        self.s = system
        self.t = table
        self.lib = system.api.lib
        self.types = table.types.encode()

    def _table(self, prm):
        s = self.s
        s.reset_retval()
        s.reset_force()
        rc = self.lib.sep_gpu_force_lj_table(s.atoms, self.t.nt, self.types,
                                             prm.ctypes.data_as(S.P_D), s.sysp, s.retp,
                                             S.SEP_ALL)
        s.check(rc)
        return np.array(s.view["f"])

    def forces(self, x, theta):
        """F(x; theta), N x 3; builds the neighbour list at x."""
        self.s.view["x"] = x
        self.s.sys.neighb_flag = 1  # new positions: rebuild the list
        return self._table(self.t.params(theta))

    def jacobian(self, theta):
        """dF/dtheta (N x 3 x 2P) at the positions of the last forces()
        call: one table pass per parameter (forward tangents, exact)."""
        ncol = 2 * len(self.t.pairs)
        return np.stack([self._table(self.t.column_params(theta, k)) for k in range(ncol)],
                        axis=-1)

    def adjoint(self, theta, g):
        """dL/dx (N x 3) and dL/dtheta for g = dL/dF at the positions of the
        last forces() call."""
        s = self.s
        prm = self.t.params(theta)
        g = np.ascontiguousarray(g, dtype=np.float64)
        dx = np.zeros((s.n, 3))
        full = np.zeros(self.t.nt * self.t.nt * 2)
        rc = self.lib.sep_gpu_lj_adjoint(s.atoms, s.sysp, self.t.nt, self.types,
                                         prm.ctypes.data_as(S.P_D), g.ctypes.data_as(S.P_D),
                                         dx.ctypes.data_as(S.P_D), full.ctypes.data_as(S.P_D))
        s.check(rc)
        return dx, self.t.fold(full)


def autograd_function():
    """torch.autograd.Function computing F(x; theta) on the device."""
    import torch

    class LJForces(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, theta, dev):
            f = dev.forces(x.detach().cpu().numpy(), theta.detach().cpu().numpy())
            ctx.dev = dev
            ctx.save_for_backward(x, theta)
            return torch.from_numpy(f)

        @staticmethod
        def backward(ctx, gF):
            x, theta = ctx.saved_tensors
            dev = ctx.dev
            # the device must hold this x and its list (another forward
            # may have run in between)
            dev.forces(x.detach().cpu().numpy(), theta.detach().cpu().numpy())
            dx, dth = dev.adjoint(theta.detach().cpu().numpy(), gF.detach().cpu().numpy())
            return torch.from_numpy(dx), torch.from_numpy(dth), None

    return LJForces


def reference_forces_torch(x, tidx, theta, table, L, rows=None):
    """FP64 O(N^2) torch restatement of the same forces (check harness):
    single-image wrap, fixed cutoffs, f_ij = 48 eps (q^7 - aw/2 q^4) r_ij.
    `rows` = (i0, i1) gives the forces on atoms i0..i1-1 only."""
    import torch
    n = x.shape[0]
    i0, i1 = rows if rows is not None else (0, n)
    npair = len(table.pairs)
    pidx = np.zeros((table.nt, table.nt), dtype=np.int64)
    for p, (a, b) in enumerate(table.pairs):
        pidx[a, b] = pidx[b, a] = p
    ti = np.asarray(tidx)
    P = torch.from_numpy(pidx[ti[i0:i1, None], ti[None, :]]).to(x.device)
    Lt = torch.as_tensor(np.asarray(L, dtype=np.float64), device=x.device)
    d = x[i0:i1, None, :] - x[None, :, :]
    d = torch.where(d > 0.5 * Lt, d - Lt, torch.where(d < -0.5 * Lt, d + Lt, d))
    r2 = (d * d).sum(-1)
    rc2 = torch.from_numpy(table.rc ** 2).to(x.device)[P]
    self_pair = torch.arange(i0, i1, device=x.device)[:, None] == \
        torch.arange(n, device=x.device)[None, :]
    mask = (r2 < rc2) & ~self_pair
    r2s = torch.where(mask, r2, torch.ones_like(r2))
    eps = theta[:npair][P]
    sig = theta[npair:][P]
    q = sig * sig / r2s
    q4 = q ** 4
    ft = 48.0 * eps * (q4 * q ** 3 - 0.5 * table.aw * q4)
    ft = torch.where(mask, ft, torch.zeros_like(ft))
    return (ft[:, :, None] * d).sum(1)


def reference_table_forces(x, tidx, prm, nt, L):
    """FP64 numpy forces straight from a flat parameter table
    prm[(a*nt + b)*4 + {0,1,2,3}] = rc, eps, sigma, aw (check harness)."""
    prm = np.asarray(prm).reshape(nt, nt, 4)
    ti = np.asarray(tidx)
    F = np.zeros_like(x)
    for i in range(x.shape[0]):
        d = x[i] - x
        d = np.where(d > 0.5 * L, d - L, np.where(d < -0.5 * L, d + L, d))
        r2 = (d * d).sum(1)
        pr = prm[ti[i], ti]
        m = (r2 < pr[:, 0] ** 2)
        m[i] = False
        q = pr[m, 2] ** 2 / r2[m]
        ft = 48.0 * pr[m, 1] * (q ** 7 - 0.5 * pr[m, 3] * q ** 4)
        F[i] = (ft[:, None] * d[m]).sum(0)
    return F


def reference_gradient_torch(x, tidx, theta, table, L, g, device="cpu", chunk=1024):
    """d(sum_i g_i . F_i)/dx and /dtheta of the FP64 restatement, by row
    chunks (O(N^2) work, O(chunk N) memory)."""
    import torch
    xt = torch.tensor(x, dtype=torch.float64, device=device, requires_grad=True)
    th = torch.tensor(theta, dtype=torch.float64, device=device, requires_grad=True)
    gt = torch.tensor(g, dtype=torch.float64, device=device)
    gx = torch.zeros_like(xt)
    gth = torch.zeros_like(th)
    n = x.shape[0]
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        F = reference_forces_torch(xt, tidx, th, table, L, rows=(i0, i1))
        a, b = torch.autograd.grad((F * gt[i0:i1]).sum(), (xt, th))
        gx += a
        gth += b
    return gx.detach().cpu().numpy(), gth.detach().cpu().numpy()


# ---- the Kob-Andersen force-matching protocol ------------------------------------

def ka_table():
    return LJTable(KA_TYPES, 2.5 * KA_THETA[3:])


def make_frames(api, natoms, rho, temp, nframes, spacing, seed=5, equil=2000, dt=0.004,
                log=None):
    """Positions every `spacing` steps of an FP64 NVT run with the true
    parameters (table forces + Nose-Hoover + leapfrog, resident on the
    device)."""
    w = W.ka_mixture(natoms, rho, temp, seed=seed)
    table = ka_table()
    s = SepSystem.from_workload(api, w, cf=2.5, dt=dt, skin=0.3, nneighb=NNEIGHB)
    s.gpu_resident(True)
    prm = table.params(KA_THETA)
    lib = api.lib

    def step():
        s.reset_retval()
        s.reset_force()
        s.check(lib.sep_gpu_force_lj_table(s.atoms, 2, KA_TYPES.encode(),
                                           prm.ctypes.data_as(S.P_D), s.sysp, s.retp,
                                           S.SEP_ALL))
        s.nosehoover(temp, 0.2)
        s.leapfrog()
    for _ in range(equil):
        step()
    frames = []
    for _ in range(nframes):
        for _ in range(spacing):
            step()
        s.gpu_sync()
        frames.append(np.array(s.view["x"]))
        if log and len(frames) % 32 == 0:
            log("frame %d of %d" % (len(frames), nframes))
    tidx = np.array([KA_TYPES.index(t.decode()) for t in w["type"]])
    s.close()
    return w, tidx, frames


def fit(api, w, frames, targets, theta0, iters=40, log=None):
    """Gauss-Newton / Levenberg-Marquardt on
    L(theta) = sum_frames ||F(x; theta) - F*||^2 / sum_frames ||F*||^2:
    the gradient from the adjoint kernel (dL/dtheta = J^T dL/dF), the
    curvature 2 J^T J from the exact forward tangents.  Returns theta and
    the loss history of the accepted steps."""
    table = ka_table()
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.004, skin=0.3, nneighb=NNEIGHB)
    dev = DeviceForces(s, table)
    norm = sum(float((t ** 2).sum()) for t in targets)
    ncol = 2 * len(table.pairs)

    def evaluate(theta):
        loss, grad, H = 0.0, np.zeros(ncol), np.zeros((ncol, ncol))
        for x, ft in zip(frames, targets):
            r = dev.forces(x, theta) - ft
            _, g = dev.adjoint(theta, 2.0 * r)
            J = dev.jacobian(theta).reshape(-1, ncol)
            loss += float((r * r).sum())
            grad += g
            H += 2.0 * (J.T @ J)
        return loss / norm, grad / norm, H / norm

    theta = np.array(theta0, dtype=np.float64)
    lam = 1e-3
    loss, grad, H = evaluate(theta)
    hist = [loss]
    for _ in range(iters):
        step = np.linalg.solve(H + lam * np.diag(np.diag(H)), -grad)
        trial = theta + step
        if np.all(trial > 0.0):
            lt, gt, Ht = evaluate(trial)
        else:
            lt = np.inf
        if lt < loss:
            theta, loss, grad, H = trial, lt, gt, Ht
            hist.append(loss)
            if log:
                log("fit step %d: loss %.3e" % (len(hist) - 1, loss))
            lam = max(lam / 10.0, 1e-12)
            if loss < 1e-26 or np.abs(step).max() < 1e-15 * np.abs(theta).max():
                break
        else:
            lam *= 10.0
            if lam > 1e12:
                break
    s.close()
    return theta, hist


def gradient_checks(dev, tidx, L, x0, f0, theta, ref_device="cpu", nx=6, seed=7):
    """The adjoint gradient of L = ||F(x0; theta) - f0||^2 against central
    finite differences of the device forward (h = 1e-6, all of theta and
    `nx` position components) and against the FP64 torch restatement's
    autograd gradient.  Returns the four maximum relative errors."""
    F = dev.forces(x0, theta)
    g = 2.0 * (F - f0)
    dx, dth = dev.adjoint(theta, g)
    h = 1e-6

    def loss_of(x, th):
        return float(((dev.forces(x, th) - f0) ** 2).sum())
    fd_th = np.zeros_like(theta)
    for k in range(theta.size):
        e = np.zeros_like(theta)
        e[k] = h
        fd_th[k] = (loss_of(x0, theta + e) - loss_of(x0, theta - e)) / (2 * h)
    rng = np.random.default_rng(seed)
    idx = [(int(rng.integers(x0.shape[0])), int(rng.integers(3))) for _ in range(nx)]
    fd_x = np.zeros(nx)
    for k, (i, c) in enumerate(idx):
        e = np.zeros_like(x0)
        e[i, c] = h
        fd_x[k] = (loss_of(x0 + e, theta) - loss_of(x0 - e, theta)) / (2 * h)
    ad_x = np.array([dx[i, c] for i, c in idx])
    rx, rth = reference_gradient_torch(x0, tidx, theta, dev.t, L, g, device=ref_device)
    dev.forces(x0, theta)  # leave the device at x0
    return {
        "theta_vs_fd": float(np.abs(dth - fd_th).max() / np.abs(fd_th).max()),
        "x_vs_fd": float(np.abs(ad_x - fd_x).max() / np.abs(dx).max()),
        "theta_vs_fp64_ref": float(np.abs(dth - rth).max() / np.abs(rth).max()),
        "x_vs_fp64_ref": float(np.abs(dx - rx).max() / np.abs(rx).max()),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--natoms", type=int, default=32000)
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--spacing", type=int, default=200)
    ap.add_argument("--fit-frames", type=int, default=0,
                    help="frames used in the fit (0: all)")
    ap.add_argument("--seed", type=int, default=7, help="seed of the +-10%% start")
    args = ap.parse_args(argv)
    import torch
    api = ProductAPI()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    w, tidx, frames = make_frames(api, args.natoms, 1.2, 0.8, args.frames, args.spacing,
                                  log=log)
    t_gen = time.perf_counter() - t0
    table = ka_table()
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.004, skin=0.3, nneighb=NNEIGHB)
    dev = DeviceForces(s, table)
    targets = [dev.forces(x, KA_THETA) for x in frames]
    rng = np.random.default_rng(args.seed)
    theta0 = KA_THETA * (1.0 + rng.uniform(-0.1, 0.1, size=KA_THETA.size))
    t1 = time.perf_counter()
    checks = gradient_checks(dev, tidx, w["L"], frames[0], targets[0], theta0,
                             ref_device="cuda" if torch.cuda.is_available() else "cpu")
    t_chk = time.perf_counter() - t1
    log("gradient checks %s (%.1f s)" % (json.dumps(checks), t_chk))
    s.close()
    nf = args.fit_frames or len(frames)
    t2 = time.perf_counter()
    theta, hist = fit(api, w, frames[:nf], targets[:nf], theta0, log=log)
    t_fit = time.perf_counter() - t2
    rel = np.abs(theta - KA_THETA) / np.abs(KA_THETA)
    print(json.dumps({
        "workload": "KA 80:20 N=%d rho 1.2 T 0.8, %d frames every %d steps, fit on %d"
                    % (args.natoms, len(frames), args.spacing, nf),
        "theta_true": KA_THETA.tolist(), "theta_start": theta0.tolist(),
        "theta_fit": theta.tolist(), "max_rel_err": float(rel.max()),
        "gradient_checks": checks, "loss_first": hist[0], "loss_last": hist[-1],
        "accepted_steps": len(hist) - 1, "seconds_generate": round(t_gen, 2),
        "seconds_checks": round(t_chk, 2), "seconds_fit": round(t_fit, 2)}), flush=True)
    ok = rel.max() <= 1e-3 and checks["theta_vs_fd"] <= 1e-4 and checks["x_vs_fd"] <= 1e-4
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
