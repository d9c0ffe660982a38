"""The FP64 torch restatement of the multi-type LJ forces used as the
reference gradient of the force-matching adjoint, pinned on the CPU:

  * its forces equal the oracle's three sep_force_lj calls (AA, AB, BB; the
    oracle is bitwise to the compiled reference on the KA golden vector) on
    a Kob-Andersen system with 3 cells per axis (the golden vector's own
    500 atoms give 2 cells per axis, where the reference's list holds
    duplicate pairs and the forces are not the physical ones);
  * its autograd gradient in theta and x equals central finite differences
    of its own loss.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from seplib_amd import forcematch as FM
from seplib_amd import workloads as W
from tests.golden import runs as R
from tests.parity import force_error

# Kob-Andersen parameters of the golden KA run: AA rc 2.5, AB rc 2.0, BB rc 2.2
TABLE = FM.LJTable("AB", [R.KA["AA"][0], R.KA["AB"][0], R.KA["BB"][0]])
THETA = np.array([R.KA[p][1] for p in ("AA", "AB", "BB")] +
                 [R.KA[p][2] for p in ("AA", "AB", "BB")])


def _ka(n):
    w = W.ka_mixture(n, 1.2, 0.8, seed=11)
    tidx = np.array([0 if t == b"A" else 1 for t in w["type"]])
    return w, tidx


def test_torch_reference_forces_match_oracle_ka():
    w, tidx = _ka(1000)
    forces, _ = R.run_ka(pytest.importorskip("oracle.checkers").OracleAPI(), w)
    F = FM.reference_forces_torch(torch.from_numpy(w["x"]), tidx, torch.from_numpy(THETA),
                                  TABLE, w["L"]).numpy()
    assert force_error(F, forces.sum(axis=0)) <= 1e-12
    # a wrong sigma for one pair shows
    bad = THETA.copy()
    bad[4] *= 1.001
    Fb = FM.reference_forces_torch(torch.from_numpy(w["x"]), tidx, torch.from_numpy(bad),
                                   TABLE, w["L"]).numpy()
    assert force_error(Fb, forces.sum(axis=0)) > 1e-6


def test_torch_reference_gradient_against_finite_differences():
    w, tidx = _ka(500)
    rng = np.random.default_rng(2)
    target = torch.from_numpy(rng.normal(size=(500, 3)) * 10.0)
    x = torch.from_numpy(w["x"]).clone().requires_grad_(True)
    th = torch.from_numpy(THETA * 1.03).clone().requires_grad_(True)

    def loss(xx, tt):
        F = FM.reference_forces_torch(xx, tidx, tt, TABLE, w["L"])
        return ((F - target) ** 2).sum()

    gx, gth = torch.autograd.grad(loss(x, th), (x, th))
    h = 1e-6
    with torch.no_grad():
        for k in range(6):
            e = torch.zeros(6, dtype=torch.float64)
            e[k] = h
            fd = float(loss(x, th + e) - loss(x, th - e)) / (2 * h)
            assert abs(fd - float(gth[k])) <= 1e-6 * float(gth.abs().max())
        for (i, c) in [(3, 0), (77, 1), (400, 2)]:
            e = torch.zeros_like(x)
            e[i, c] = h
            fd = float(loss(x + e, th) - loss(x - e, th)) / (2 * h)
            assert abs(fd - float(gx[i, c])) <= 1e-5 * float(gx.abs().max())


def test_column_tables_give_the_exact_jacobian():
    """LJTable.column_params(theta, k): a table whose forces are dF/dtheta_k
    (eps' = 14 eps/sigma, aw' = 4 aw/7 for a sigma column), against forward
    mode autodiff of the torch restatement."""
    w, tidx = _ka(300)
    th = THETA * np.array([1.02, 0.98, 1.01, 1.03, 0.99, 1.02])
    J = torch.func.jacfwd(lambda t: FM.reference_forces_torch(
        torch.from_numpy(w["x"]), tidx, t, TABLE, w["L"]))(torch.from_numpy(th)).numpy()
    for k in range(6):
        col = FM.reference_table_forces(w["x"], tidx, TABLE.column_params(th, k), 2, w["L"])
        assert np.abs(col - J[..., k]).max() <= 1e-11 * np.abs(J[..., k]).max(), k
    # and the full table gives the forces themselves
    F = FM.reference_table_forces(w["x"], tidx, TABLE.params(th), 2, w["L"])
    Ft = FM.reference_forces_torch(torch.from_numpy(w["x"]), tidx, torch.from_numpy(th),
                                   TABLE, w["L"]).numpy()
    assert np.abs(F - Ft).max() <= 1e-12 * np.abs(Ft).max()


def test_chunked_reference_gradient_equals_full():
    w, tidx = _ka(500)
    rng = np.random.default_rng(5)
    g = rng.normal(size=(500, 3))
    gx, gth = FM.reference_gradient_torch(w["x"], tidx, THETA, TABLE, w["L"], g, chunk=128)
    x = torch.from_numpy(w["x"]).clone().requires_grad_(True)
    th = torch.from_numpy(THETA).clone().requires_grad_(True)
    F = FM.reference_forces_torch(x, tidx, th, TABLE, w["L"])
    ax, ath = torch.autograd.grad((F * torch.from_numpy(g)).sum(), (x, th))
    assert np.abs(gx - ax.numpy()).max() <= 1e-12 * np.abs(ax.numpy()).max()
    assert np.abs(gth - ath.numpy()).max() <= 1e-12 * np.abs(ath.numpy()).max()
