"""Bound-constrained CG solvers (host logic, CPU) on the oracle's operators:
the primal (gap) and dual (pixel force) contact objectives of a periodic
sphere reach the constrained minimum that scipy's L-BFGS-B finds and that
the oracle's Polonsky-Keer solve gives (reference
``test/test_numpi_ccgs.py``, ``test/test_primal_dual_objective.py``); the
mean-value constraint holds; KKT conditions at the solution."""

import numpy as np
import pytest
import scipy.optimize as optim

from contactmechanics_amd.Optimization import CCGWithoutRestart, CCGWithRestart
from oracle.ccg import constrained_conjugate_gradients as pk
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

N, S, R, ES, OFFSET = 48, 1.0, 10.0, 50.0, 0.005


def sphere(n=N, s=S):
    x = (np.arange(n) + 0.5) * s / n - s / 2
    return -(x[:, None] ** 2 + x[None, :] ** 2) / (2 * R)


def primal(hs, h, offset):
    def fun(gap):
        disp = gap.reshape(h.shape) + h + offset
        f = hs.evaluate_force(disp)
        return -0.5 * np.sum(f * disp), -f.ravel()

    def hessp(d):
        return -hs.evaluate_force(d.reshape(h.shape)).reshape(d.shape)
    return fun, hessp


def dual(hs, h, offset):
    def fun(p):
        p = p.reshape(h.shape)
        disp = hs.evaluate_disp(-p)
        return 0.5 * np.sum(p * disp) - np.sum(p * (h + offset)), (disp - h - offset).ravel()

    def hessp(d):
        return hs.evaluate_disp(-d.reshape(h.shape)).reshape(d.shape)
    return fun, hessp


def lbfgsb(fun, x0, gtol):
    # L-BFGS-B as in the reference tests (ftol 1e-20: it stops on gtol or on
    # a failed line search, which at these tolerances is the minimiser to
    # round-off)
    r = optim.minimize(fun, x0.ravel(), method="L-BFGS-B", jac=True, bounds=optim.Bounds(lb=np.zeros(x0.size)),
                       options=dict(gtol=gtol, ftol=1e-20, maxiter=20000))
    return r.x.reshape(x0.shape)


@pytest.mark.parametrize("solve", [CCGWithoutRestart.constrained_conjugate_gradients,
                                   CCGWithRestart.constrained_conjugate_gradients_with_restart])
def test_primal_reaches_constrained_minimum(solve):
    hs = PeriodicHalfSpace((N, N), ES, (S, S))
    h = sphere()
    fun, hessp = primal(hs, h, OFFSET)
    gap0 = np.zeros((N, N)) - h - OFFSET
    r = solve(fun, hessp, x0=gap0, gtol=1e-9, maxiter=2000)
    assert r.success, r.message
    gap = r.x.reshape(N, N)
    # the oracle's Polonsky-Keer solve under the same offset
    ref = pk(hs, h, offset=OFFSET, pentol=1e-12, forcetol=1e-14, maxiter=5000)
    np.testing.assert_allclose(gap, np.maximum(ref.x - h - OFFSET, 0), atol=1e-8)
    np.testing.assert_allclose(gap, lbfgsb(fun, gap0, 1e-8), atol=1e-6)
    # KKT: gap >= 0; the gradient (the contact pressure force, -f of the
    # substrate) vanishes off contact and pushes into the bound in contact
    p = r.jac.reshape(N, N)
    assert gap.min() >= 0
    assert np.max(np.abs(p[gap > 0])) <= 1e-9
    assert np.all(p[gap == 0] >= -1e-9) and p.max() > 0


@pytest.mark.parametrize("solve", [CCGWithoutRestart.constrained_conjugate_gradients,
                                   CCGWithRestart.constrained_conjugate_gradients_with_restart])
def test_dual_reaches_constrained_minimum(solve):
    hs = PeriodicHalfSpace((N, N), ES, (S, S))
    h = sphere()
    fun, hessp = dual(hs, h, OFFSET)
    disp = np.maximum(h + OFFSET, 0)
    p0 = np.maximum(hs.evaluate_force(disp), 0)
    r = solve(fun, hessp, x0=p0, gtol=1e-10, maxiter=3000)
    assert r.success, r.message
    p = r.x.reshape(N, N)
    ref = pk(hs, h, offset=OFFSET, pentol=1e-12, forcetol=1e-14, maxiter=5000)
    np.testing.assert_allclose(p, ref.jac, atol=1e-7 * np.max(ref.jac))
    np.testing.assert_allclose(p, lbfgsb(fun, p0, 1e-10), atol=1e-4 * np.max(ref.jac))


def test_dual_mean_value_constraint():
    # load control through the mean of the pixel forces
    hs = PeriodicHalfSpace((N, N), ES, (S, S))
    h = sphere()
    ref = pk(hs, h, external_force=2e-4, pentol=1e-12, forcetol=1e-14, maxiter=5000)
    fun, hessp = dual(hs, h, ref.offset)
    p0 = np.full((N, N), 2e-4 / N ** 2)
    for solve in (CCGWithoutRestart.constrained_conjugate_gradients,
                  CCGWithRestart.constrained_conjugate_gradients_with_restart):
        r = solve(fun, hessp, x0=p0, gtol=1e-11, mean_val=2e-4 / N ** 2, maxiter=3000)
        assert r.success, r.message
        np.testing.assert_allclose(np.sum(r.x), 2e-4, rtol=1e-12)
        np.testing.assert_allclose(r.x.reshape(N, N), ref.jac, atol=1e-6 * np.max(ref.jac))


def test_free_primal_with_bounds_from_heights():
    # displacements on the padded grid, lower bounds h + offset on the
    # topography and none on the padding (reference
    # test/test_ccg_without_restart_nonperiodic.py)
    n = 24
    hs = FreeHalfSpace((n, n), 0.75, (4.0, 4.0))
    x = (np.arange(n) + 0.5) * 4.0 / n - 2.0
    h = -(x[:, None] ** 2 + x[None, :] ** 2) / 2
    pen = 0.5
    lb = np.full((2 * n, 2 * n), -np.inf)
    lb[:n, :n] = h + pen

    def fun(u):
        f = hs.evaluate_force(u.reshape(2 * n, 2 * n))
        return -0.5 * np.sum(f * u.reshape(2 * n, 2 * n)), -f.ravel()

    def hessp(d):
        return -hs.evaluate_force(d.reshape(2 * n, 2 * n)).reshape(d.shape)
    u0 = np.maximum(np.zeros((2 * n, 2 * n)), lb)
    r = CCGWithoutRestart.constrained_conjugate_gradients(fun, hessp, u0.ravel(), gtol=1e-10, bounds=lb.ravel(),
                                                          maxiter=3000)
    assert r.success, r.message
    ref = optim.minimize(fun, u0.ravel(), method="L-BFGS-B", jac=True,
                         bounds=optim.Bounds(lb=lb.ravel()), options=dict(gtol=1e-10, ftol=1e-20, maxiter=20000))
    assert ref.success
    assert np.max(np.abs(r.x - ref.x)) < 1e-5
