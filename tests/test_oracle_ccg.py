"""Pin the oracle's constrained conjugate gradients (CPU).

* Hertz (free half-space) with the reference test's tolerances
  (``test/test_hertz.py:39-209``), here at 128^2 with load and offset control;
* Westergaard sinusoid, periodic 256x16 (``test/test_westergaard.py:42-84``);
* KKT conditions at convergence on a random rough surface;
* an independent bound-constrained minimisation (L-BFGS-B) of the elastic
  energy on a small free grid (``test/test_primal_dual_objective.py:216-254``);
* the plastic cone log recorded in ``examples/plastic_simulation_cone.ipynb``
  (``tests/golden/plastic_cone_notebook.json``);
* the bearing-area limit at vanishing hardness (``test/test_plastic.py:45-92``);
* maxiter and NaN error paths (``ConstrainedConjugateGradients.py:491-509``).
"""

import json
import os

import numpy as np
import pytest
import scipy.optimize as optim

from oracle import reference_solutions as rs
from oracle.ccg import constrained_conjugate_gradients as ccg
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def sphere(nx, ny, sx, sy, R):
    x = ((np.arange(nx) - nx / 2) * sx / nx)[:, None]
    y = ((np.arange(ny) - ny / 2) * sy / ny)[None, :]
    return -(x ** 2 + y ** 2) / (2 * R), np.sqrt(x ** 2 + y ** 2)


@pytest.mark.parametrize("disp0,normal_force", [(0.1, None), (None, 15.0)])
def test_hertz_free(disp0, normal_force):
    R, Es, sx, n = 20.0, 102.0, 5.0, 128
    hs = FreeHalfSpace((n, n), Es, (sx, sx))
    h, r = sphere(n, n, sx, sx, R)
    res = ccg(hs, h, external_force=normal_force, offset=disp0, pentol=1e-10)
    assert res.success
    forces = res.jac
    F = forces.sum()
    if normal_force is not None:
        np.testing.assert_allclose(F, normal_force, rtol=1e-7)
        np.testing.assert_allclose(res.offset, rs.hertz_penetration(normal_force, R, Es), rtol=1e-2)
    else:
        np.testing.assert_allclose(res.offset, disp0, rtol=1e-7)
        np.testing.assert_allclose(F, rs.hertz_normal_load(disp0, R, Es), rtol=2e-2)
    a, p0 = rs.hertz_radius_and_pressure(F, R, Es)
    area = np.sum(forces > 0) * hs.area_per_pt
    np.testing.assert_allclose(area, np.pi * a ** 2, rtol=1e-1)
    p = forces / hs.area_per_pt
    m = r < 0.99 * a
    assert np.max(np.abs(rs.hertz_pressure(r, a, p0)[m] - p[m])) / Es < 2e-3
    e_ref = rs.hertz_elastic_energy(res.offset, R, Es)
    np.testing.assert_allclose(res.fun, e_ref, rtol=3e-2)


@pytest.mark.parametrize("disp0", [-0.9, -0.1])
def test_westergaard_periodic(disp0):
    sx, sy, Es, nx, ny = 30.0, 1.0, 3.56, 256, 16
    hs = PeriodicHalfSpace((nx, ny), Es, (sx, sy))
    profile = np.resize(np.cos(2 * np.pi * np.arange(nx) / nx), (ny, nx)).T
    res = ccg(hs, profile, offset=disp0, pentol=1e-9)
    assert res.success
    x = np.arange(nx) * sx / nx
    mean_p = np.mean(res.jac) / hs.area_per_pt
    pth = mean_p * rs.westergaard_pressure(x / sx, sx * mean_p / Es)
    assert np.allclose(res.jac[: nx // 2, 0] / hs.area_per_pt, pth[: nx // 2], atol=1e-2)


def rough(n, seed, H=0.8, slope=0.1):
    rng = np.random.default_rng(seed)
    w = np.fft.rfft2(rng.standard_normal((n, n)))
    qx = np.fft.fftfreq(n)[:, None]
    qy = np.fft.rfftfreq(n)[None, :]
    q = np.hypot(qx, qy)
    amp = np.zeros_like(q)
    m = (q >= 1.0 / n) & (q <= 0.25)
    amp[m] = q[m] ** (-(1 + H))
    h = np.fft.irfft2(w * amp, s=(n, n))
    gx = np.gradient(h, axis=0) * n
    gy = np.gradient(h, axis=1) * n
    return h * slope / np.sqrt(np.mean(gx ** 2 + gy ** 2))


def test_kkt_at_convergence():
    n = 64
    h = rough(n, 7)
    hs = PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0))
    res = ccg(hs, h, external_force=0.01, pentol=1e-12, forcetol=1e-12)
    assert res.success
    f = -res.jac  # compressive negative
    gap = res.x - h - res.offset
    scale = np.max(np.abs(f))
    assert np.all(f <= 1e-12 * scale)
    assert np.min(gap) > -1e-11
    assert np.max(np.abs(f * gap)) < 1e-10 * scale
    np.testing.assert_allclose(res.jac.sum(), 0.01, rtol=1e-9)


def test_agrees_with_lbfgsb_free():
    # minimise the elastic energy 1/2 u.K.u over u >= h + offset (comp only)
    nx = ny = 16
    sx = sy = 1.0
    Es, R, pen = 50.0, 10.0, 0.005
    hs = FreeHalfSpace((nx, ny), Es, (sx, sy))
    x = (np.arange(nx) * sx / nx - sx / 2)[:, None]
    y = (np.arange(ny) * sy / ny - sy / 2)[None, :]
    h = -(x ** 2 + y ** 2) / (2 * R)
    M = hs.nb_domain_grid_pts

    def energy(uflat):
        u = uflat.reshape(M)
        f = hs.evaluate_force(u)
        return -0.5 * np.sum(f * u), -f.ravel()

    lb = np.full(M, -np.inf)
    lb[:nx, :ny] = h + pen
    u0 = np.maximum(np.ones(M) * 0.0, lb)
    u0[np.isinf(u0)] = 0.0
    opt = optim.minimize(energy, u0.ravel(), jac=True, method="L-BFGS-B",
                         bounds=optim.Bounds(lb.ravel(), np.full(lb.size, np.inf)),
                         options=dict(gtol=1e-12, ftol=1e-20, maxiter=20000))
    res = ccg(hs, h, offset=pen, pentol=1e-10)
    assert res.success
    np.testing.assert_allclose(opt.x.reshape(M), res.x, atol=1e-6)


def _cone_setup():
    with open(os.path.join(GOLDEN, "plastic_cone_notebook.json")) as fh:
        data = json.load(fh)
    s = data["setup"]
    nx, ny, sx, sy = s["nx"], s["ny"], s["sx"], s["sy"]
    x = np.arange(nx).reshape(-1, 1) * sx / nx - sx / 2
    y = np.arange(ny).reshape(1, -1) * sy / ny - sy / 2
    h = -np.sqrt(x ** 2 + y ** 2) * s["cone_slope"]
    return data, s, h


def test_plastic_cone_matches_notebook_log():
    data, s, h = _cone_setup()
    nx, ny = s["nx"], s["ny"]
    hs = FreeHalfSpace((nx, ny), s["young"], (s["sx"], s["sy"]))
    hardness = s["hardness"] * hs.area_per_pt
    disp0 = np.zeros(hs.nb_domain_grid_pts)
    disp0[:nx, :ny] = h + s["initial_penetration"]
    disp0 = np.where(disp0 > 0, disp0, 0)
    plastic = np.zeros((nx, ny))
    for load, F in enumerate(s["external_forces"]):
        rows = [r for r in data["rows"] if r["load"] == load]
        trace = []
        res = ccg(hs, h + plastic, hardness=hardness, external_force=F, initial_displacements=disp0,
                  pentol=s["pentol"], trace=trace)
        assert res.success
        if load == 0:
            # same start state: the whole logged trajectory agrees (the
            # notebook's older release stopped at it=44 on a looser force test)
            for ours, rec in zip(trace, rows[:-1]):
                assert ours["status"] == rec["status"]
                assert ours["area"] == rec["area"]
                assert ours["plastic_area"] == rec["plastic_area"]
                np.testing.assert_allclose(ours["total_force"], rec["total_force"], rtol=1e-9)
                np.testing.assert_allclose(ours["offset"], rec["offset"], rtol=1e-9)
        conv = rows[-1]
        last = trace[-1]
        assert abs(last["area"] - conv["area"]) <= 0.01 * conv["area"] + 1
        np.testing.assert_allclose(last["offset"], conv["offset"], rtol=2e-5)
        np.testing.assert_allclose(last["total_force"], conv["total_force"], rtol=1e-4)
        disp0 = res.x
        gap = res.x[:nx, :ny] - (h + plastic + res.offset)
        plastic = plastic + np.where(res.plastic, gap, 0.0)


def test_bearing_area_limit_at_vanishing_hardness():
    n = 64
    h = rough(n, 11, slope=0.05)
    hs = PeriodicHalfSpace((n, n), 1.0, (float(n), float(n)))
    res = ccg(hs, h, hardness=1e-12 * hs.area_per_pt, offset=-0.002 + h.max() - 0.01 * np.ptp(h),
              maxiter=1000, pentol=1e-6)
    assert res.success
    ncontact = np.sum(res.active_set)
    bearing = optim.bisect(lambda x: np.sum(h > x) - ncontact, h.min(), h.max())
    cba = h > bearing
    assert np.sum(res.active_set != cba) < 25


def test_maxiter_and_nan_paths():
    n = 32
    h = rough(n, 2)
    hs = PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0))
    res = ccg(hs, h, external_force=0.01, pentol=1e-12, maxiter=1)
    assert not res.success and res.message == "Reached maxiter = 1" and res.nit == 1
    bad = h.copy()
    bad[3, 3] = np.nan
    with pytest.raises(RuntimeError, match="nan encountered"):
        ccg(hs, bad, external_force=0.01, pentol=1e-12)
