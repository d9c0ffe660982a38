"""Force matching on the device: the type-pair-table forward
(sep_gpu_force_lj_table) and the adjoint (sep_gpu_lj_adjoint), through the
C ABI.

  * the table forward in one list pass equals the reference's three
    sep_force_lj calls (AA, AB, BB) on the KA golden vector and, at 3 cells
    per axis, the oracle's, to the FP64 bar (forces, epot, virial);
  * dL/dx and dL/dtheta equal torch autograd through the FP64 torch
    restatement (itself pinned to the oracle and to finite differences on
    the CPU) to 1e-10, and central finite differences of the product's own
    forward to 1e-5;
  * the Jacobian columns (one table pass per parameter) equal forward-mode
    autodiff of the restatement;
  * a Levenberg-Marquardt fit (adjoint gradient, Gauss-Newton curvature)
    from a 10% perturbed start recovers the parameters to 1e-3.
"""
import os

import numpy as np
import pytest

from seplib_amd import forcematch as FM
from seplib_amd import sep as S
from seplib_amd import workloads as W
from seplib_amd.sim import SepSystem
from tests.golden import runs as R
from tests.parity import FP64_TOL, force_error, tensor_error

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TABLE = FM.LJTable("AB", [R.KA["AA"][0], R.KA["AB"][0], R.KA["BB"][0]])
THETA = np.array([R.KA[p][1] for p in ("AA", "AB", "BB")] +
                 [R.KA[p][2] for p in ("AA", "AB", "BB")])


def _table_forward(api, w, theta, resident=False):
    s = SepSystem.from_workload(api, w, cf=2.5, dt=0.001, skin=0.3)
    try:
        if resident:
            s.gpu_resident(True)
        s.reset_retval()
        s.reset_force()
        prm = TABLE.params(theta)
        s.check(api.lib.sep_gpu_force_lj_table(s.atoms, 2, b"AB", prm.ctypes.data_as(S.P_D),
                                               s.sysp, s.retp, S.SEP_ALL))
        if resident:
            s.gpu_sync()
        return np.array(s.view["f"]), R.ret_vec(s)
    finally:
        s.close()


def _check(f, r, gf, gr):
    assert force_error(f, gf) <= FP64_TOL, force_error(f, gf)
    assert abs(r[0] - gr[0]) <= FP64_TOL * abs(gr[0]), (r[0], gr[0])
    assert tensor_error(r[3:12], gr[3:12]) <= FP64_TOL


@pytest.mark.parametrize("resident", [False, True])
def test_table_forward_matches_golden_ka(product, resident):
    """500 atoms, 2 cells per axis (the reference's list holds duplicate
    pairs there; the table pass walks the same list)."""
    g = np.load(os.path.join(GOLD, "g5_ka.npz"), allow_pickle=False)
    f, r = _table_forward(product, R.workload_from(g), THETA, resident)
    _check(f, r, g["forces"].sum(axis=0), g["rets"][-1])


def test_table_forward_matches_oracle_three_cells(product, oracle):
    w = W.ka_mixture(4000, 1.2, 0.8, seed=3)
    gf, gr = R.run_ka(oracle, w)
    f, r = _table_forward(product, w, THETA)
    _check(f, r, gf.sum(axis=0), gr[-1])


def test_table_rejects_bad_arguments(product):
    w = W.ka_mixture(1000, 1.2, 0.8, seed=3)
    s = SepSystem.from_workload(product, w, cf=2.5, dt=0.001, skin=0.3, nneighb=FM.NNEIGHB)
    try:
        prm = TABLE.params(THETA)
        big = prm.copy()
        big[0] = 2.6  # AA cutoff above sys.cf
        assert product.lib.sep_gpu_force_lj_table(s.atoms, 2, b"AB", big.ctypes.data_as(S.P_D),
                                                  s.sysp, s.retp, S.SEP_ALL) != 0
        assert product.lib.sep_gpu_force_lj_table(s.atoms, 2, b"AA", prm.ctypes.data_as(S.P_D),
                                                  s.sysp, s.retp, S.SEP_ALL) != 0
    finally:
        s.close()


def _ka_device(product, n, seed):
    w = W.ka_mixture(n, 1.2, 0.8, seed=seed)
    tidx = np.array([0 if t == b"A" else 1 for t in w["type"]])
    s = SepSystem.from_workload(product, w, cf=2.5, dt=0.001, skin=0.3, nneighb=FM.NNEIGHB)
    return w, tidx, s, FM.DeviceForces(s, TABLE)


def test_adjoint_matches_torch_reference(product):
    w, tidx, s, dev = _ka_device(product, 2000, 8)
    try:
        theta = THETA * np.array([1.04, 0.97, 1.02, 0.99, 1.01, 1.03])
        F = dev.forces(w["x"], theta)
        rng = np.random.default_rng(4)
        g = rng.normal(size=F.shape)
        dx, dth = dev.adjoint(theta, g)
        x = torch.from_numpy(w["x"]).clone().requires_grad_(True)
        th = torch.from_numpy(theta).clone().requires_grad_(True)
        Ft = FM.reference_forces_torch(x, tidx, th, TABLE, w["L"])
        assert force_error(F, Ft.detach().numpy()) <= FP64_TOL
        gx, gth = torch.autograd.grad((Ft * torch.from_numpy(g)).sum(), (x, th))
        gx, gth = gx.numpy(), gth.numpy()
        assert np.abs(dx - gx).max() <= 1e-10 * np.abs(gx).max()
        assert np.abs(dth - gth).max() <= 1e-10 * np.abs(gth).max(), (dth, gth)
    finally:
        s.close()


def test_adjoint_matches_finite_differences(product):
    w, tidx, s, dev = _ka_device(product, 1000, 9)
    try:
        theta = THETA * np.array([0.95, 1.05, 1.03, 1.02, 0.98, 1.01])
        target = dev.forces(w["x"], THETA)

        def loss(x, th):
            return float(((dev.forces(x, th) - target) ** 2).sum())
        F = dev.forces(w["x"], theta)
        dx, dth = dev.adjoint(theta, 2.0 * (F - target))
        h = 1e-6
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            fd = (loss(w["x"], theta + e) - loss(w["x"], theta - e)) / (2 * h)
            assert abs(fd - dth[k]) <= 1e-5 * np.abs(dth).max(), (k, fd, dth[k])
        for (i, c) in [(5, 0), (321, 1), (777, 2)]:
            e = np.zeros_like(w["x"])
            e[i, c] = h
            fd = (loss(w["x"] + e, theta) - loss(w["x"] - e, theta)) / (2 * h)
            assert abs(fd - dx[i, c]) <= 1e-5 * np.abs(dx).max(), (i, c, fd, dx[i, c])
    finally:
        s.close()


def test_jacobian_columns_match_forward_mode(product):
    w, tidx, s, dev = _ka_device(product, 1000, 12)
    try:
        theta = THETA * np.array([1.02, 0.98, 1.01, 1.03, 0.99, 1.02])
        dev.forces(w["x"], theta)
        J = dev.jacobian(theta)
        Jt = torch.func.jacfwd(lambda t: FM.reference_forces_torch(
            torch.from_numpy(w["x"]), tidx, t, TABLE, w["L"]))(torch.from_numpy(theta)).numpy()
        for k in range(6):
            assert force_error(J[..., k], Jt[..., k]) <= FP64_TOL, k
    finally:
        s.close()


def test_autograd_function_and_fit_recover_parameters(product):
    w, tidx, frames = FM.make_frames(product, 1000, 1.2, 0.8, nframes=4, spacing=50,
                                     seed=6, equil=200)
    table = FM.ka_table()
    s = SepSystem.from_workload(product, w, cf=2.5, dt=0.004, skin=0.3, nneighb=FM.NNEIGHB)
    try:
        dev = FM.DeviceForces(s, table)
        targets = [dev.forces(x, FM.KA_THETA) for x in frames]
        assert all(np.isfinite(t).all() for t in targets)
        fn = FM.autograd_function()
        th = torch.tensor(FM.KA_THETA * 1.02, dtype=torch.float64, requires_grad=True)
        x = torch.from_numpy(frames[0])
        loss = ((fn.apply(x, th, dev) - torch.from_numpy(targets[0])) ** 2).sum()
        loss.backward()
        _, gref = dev.adjoint(FM.KA_THETA * 1.02,
                              2.0 * (dev.forces(frames[0], FM.KA_THETA * 1.02) - targets[0]))
        assert np.abs(th.grad.numpy() - gref).max() <= 1e-12 * np.abs(gref).max()
    finally:
        s.close()
    rng = np.random.default_rng(1)
    theta0 = FM.KA_THETA * (1.0 + rng.uniform(-0.1, 0.1, size=6))
    theta, hist = FM.fit(product, w, frames, targets, theta0)
    assert hist[-1] < 1e-16 * hist[0]
    assert (np.abs(theta - FM.KA_THETA) / FM.KA_THETA).max() <= 1e-3, theta
