"""GPU parity of the constrained conjugate gradient solve against the oracle.

FP64 plans follow the oracle's iterate trajectory: same iteration count and
forces equal to 1e-10 relative (L-infinity) on the problems below.  Plastic
solves that take mixing steps (discrete bisection on the bearing area) agree
to tolerance: 1e-5 relative on forces and offsets.
"""

import json
import os

import numpy as np
import pytest

from contactmechanics_amd import (FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace, PlasticTopography,
                                  Topography, make_plastic_system, make_system)
from contactmechanics_amd import constrained_conjugate_gradients as ccg_gpu
from oracle.ccg import constrained_conjugate_gradients as ccg_cpu
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def paraboloid(n, R=1.0):
    x = (np.arange(n) + 0.5) / n
    return -((x[:, None] - 0.5) ** 2 + (x[None, :] - 0.5) ** 2) / (2 * R)


def rough(n, seed, H=0.8, slope=0.1):
    rng = np.random.default_rng(seed)
    w = np.fft.rfft2(rng.standard_normal((n, n)))
    q = np.hypot(np.fft.fftfreq(n)[:, None], np.fft.rfftfreq(n)[None, :])
    amp = np.zeros_like(q)
    m = (q >= 1.0 / n) & (q <= 0.25)
    amp[m] = q[m] ** (-(1 + H))
    h = np.fft.irfft2(w * amp, s=(n, n))
    gx, gy = np.gradient(h, axis=0) * n, np.gradient(h, axis=1) * n
    return h * slope / np.sqrt(np.mean(gx ** 2 + gy ** 2))


@pytest.mark.parametrize("n", [64, 256])
def test_periodic_hertz_load_control(gpu, n):
    h = paraboloid(n)
    kw = dict(external_force=1e-3, pentol=1e-12, forcetol=1e-14, maxiter=5000)
    rg = ccg_gpu(PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0)), Topography(h, (1.0, 1.0)), **kw)
    rc = ccg_cpu(PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0)), h, **kw)
    assert rg.success and rc.success
    assert rg.nit == rc.nit and rg.nfev == rc.nfev
    assert rel(rg.jac, rc.jac) < 1e-10
    np.testing.assert_allclose(rg.offset, rc.offset, rtol=1e-12)
    np.testing.assert_allclose(rg.fun, rc.fun, rtol=1e-10)
    assert np.array_equal(rg.active_set, rc.active_set)


@pytest.mark.parametrize("kw", [dict(external_force=15.0), dict(offset=0.1)])
def test_free_hertz(gpu, kw):
    n, R, Es, sx = 128, 20.0, 102.0, 5.0
    x = ((np.arange(n) - n / 2) * sx / n)[:, None]
    y = ((np.arange(n) - n / 2) * sx / n)[None, :]
    h = -(x ** 2 + y ** 2) / (2 * R)
    rg = ccg_gpu(FreeFFTElasticHalfSpace((n, n), Es, (sx, sx)), h, pentol=1e-10, **kw)
    rc = ccg_cpu(FreeHalfSpace((n, n), Es, (sx, sx)), h, pentol=1e-10, **kw)
    # the free operator matches the oracle to ~1e-11 (FFT round-off of the
    # Love kernel), which can move the iteration where maxdu first drops
    # below pentol by one; both iterates are then converged to pentol
    assert rg.success and rc.success and abs(rg.nit - rc.nit) <= 2
    assert rel(rg.x, rc.x) < 1e-7
    assert rel(rg.jac, rc.jac) < 1e-5


def test_displacement_control_rough_periodic(gpu):
    n = 128
    h = rough(n, 3)
    kw = dict(offset=np.max(h) - 1.5 * np.std(h), pentol=1e-10, forcetol=1e-12, maxiter=3000)
    rg = ccg_gpu(PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0)), h, **kw)
    rc = ccg_cpu(PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0)), h, **kw)
    assert rg.success and rc.success and rg.nit == rc.nit
    assert rel(rg.jac, rc.jac) < 1e-9


def test_masked_topography_flat_punch(gpu):
    # masked pixels leave the computational region (reference :189-194,
    # test/test_flat_punch.py)
    n, r_s, E = 64, 20.0, 3.56
    sx = 2.5 * r_s
    xr, yr = np.arange(n)[:, None], np.arange(n)[None, :]
    r_sq = (sx / n * (xr - n // 2)) ** 2 + (sx / n * (yr - n // 2)) ** 2
    h = np.ma.masked_where(r_sq > r_s ** 2, np.zeros([n, n]))
    rg = ccg_gpu(FreeFFTElasticHalfSpace((n, n), E, (sx, sx)), h, offset=0.1, pentol=1e-8)
    rc = ccg_cpu(FreeHalfSpace((n, n), E, (sx, sx)), h, offset=0.1, pentol=1e-8)
    assert rg.success and rc.success and abs(rg.nit - rc.nit) <= 2
    assert rel(rg.x, rc.x) < 1e-6
    assert rel(rg.jac, rc.jac) < 1e-4
    np.testing.assert_allclose(rg.jac.sum() / 0.1 / (2 * r_s * E), 1.0, atol=0.05)


def test_initial_forces_and_displacements(gpu):
    n = 64
    h = rough(n, 5)
    sub_g = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0))
    sub_c = PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0))
    f0 = np.zeros((n, n))
    f0[h > np.percentile(h, 90)] = -0.01 / (0.1 * n * n)
    kw = dict(external_force=0.01, pentol=1e-10, forcetol=1e-12)
    rg = ccg_gpu(sub_g, h, initial_forces=f0, **kw)
    rc = ccg_cpu(sub_c, h, initial_forces=f0, **kw)
    assert rg.success and rc.success and rg.nit == rc.nit and rel(rg.jac, rc.jac) < 1e-9
    u0 = np.maximum(h - np.max(h) + 0.02, 0)
    rg = ccg_gpu(sub_g, h, initial_displacements=u0, **kw)
    rc = ccg_cpu(sub_c, h, initial_displacements=u0, **kw)
    assert rg.success and rc.success and rg.nit == rc.nit and rel(rg.jac, rc.jac) < 1e-9


def test_maxiter_nan_and_callback(gpu):
    n = 32
    h = rough(n, 2)
    sub = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0))
    res = ccg_gpu(sub, h, external_force=0.01, pentol=1e-12, maxiter=3)
    ref = ccg_cpu(PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0)), h, external_force=0.01, pentol=1e-12, maxiter=3)
    assert not res.success and res.message == "Reached maxiter = 3" and res.nit == 3
    assert rel(res.jac, ref.jac) < 1e-10
    seen = []
    res = ccg_gpu(sub, h, external_force=0.01, pentol=1e-10,
                  callback=lambda it, f, d: seen.append((it, d["area"])))
    assert res.success
    assert [s[0] for s in seen] == list(range(1, res.nit))
    # the iteration that reaches maxiter and the NaN iteration also reach the
    # callback (reference :478-492), as in the oracle
    seen_g, seen_c = [], []
    ccg_gpu(sub, h, external_force=0.01, pentol=1e-12, maxiter=3, callback=lambda it, f, d: seen_g.append(it))
    ccg_cpu(PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0)), h, external_force=0.01, pentol=1e-12, maxiter=3,
            callback=lambda it, f, d: seen_c.append(it))
    assert seen_g == seen_c == [1, 2, 3]
    bad = h.copy()
    bad[3, 3] = np.nan
    nan_g, nan_c = [], []
    with pytest.raises(RuntimeError, match="nan encountered"):
        ccg_gpu(sub, bad, external_force=0.01, pentol=1e-12, callback=lambda it, f, d: nan_g.append(it))
    with pytest.raises(RuntimeError, match="nan encountered"):
        ccg_cpu(PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0)), bad, external_force=0.01, pentol=1e-12,
                callback=lambda it, f, d: nan_c.append(it))
    assert nan_g == nan_c and len(nan_g) >= 1


def test_plastic_cone_notebook_first_load(gpu):
    with open(os.path.join(GOLDEN, "plastic_cone_notebook.json")) as fh:
        s = json.load(fh)["setup"]
    nx, ny, sx, sy = s["nx"], s["ny"], s["sx"], s["sy"]
    x = np.arange(nx).reshape(-1, 1) * sx / nx - sx / 2
    y = np.arange(ny).reshape(1, -1) * sy / ny - sy / 2
    topo = Topography(-np.sqrt(x ** 2 + y ** 2) * s["cone_slope"], (sx, sy))
    system = make_plastic_system(FreeFFTElasticHalfSpace((nx, ny), s["young"], (sx, sy)),
                                 PlasticTopography(topo, s["hardness"]))
    hs = FreeHalfSpace((nx, ny), s["young"], (sx, sy))
    disp0 = np.zeros(hs.nb_domain_grid_pts)
    disp0[:nx, :ny] = topo.heights() + s["initial_penetration"]
    disp0 = np.where(disp0 > 0, disp0, 0)
    plastic = np.zeros((nx, ny))
    d_cpu = disp0
    for load, F in enumerate(s["external_forces"][:3]):
        rg = system.minimize_proxy(external_force=F, initial_displacements=disp0, pentol=s["pentol"])
        rc = ccg_cpu(hs, topo.heights() + plastic, hardness=s["hardness"] * hs.area_per_pt,
                     external_force=F, initial_displacements=d_cpu, pentol=s["pentol"])
        assert rg.success and rc.success
        if load == 0:  # purely PK iterations: same trajectory
            assert rg.nit == rc.nit and rel(rg.jac, rc.jac) < 1e-10
        else:  # includes mixing steps with bisection
            assert rel(rg.jac, rc.jac) < 1e-4
            np.testing.assert_allclose(rg.offset, rc.offset, rtol=1e-5)
        assert np.max(system.force / system.area_per_pt) < s["hardness"] * (1 + 1e-12)
        disp0 = system.disp
        d_cpu = rc.x
        gap = rc.x[:nx, :ny] - (topo.heights() + plastic + rc.offset)
        plastic = plastic + np.where(rc.plastic, gap, 0.0)


def test_bearing_area_limit(gpu):
    n = 64
    h = rough(n, 11, slope=0.05)
    kw = dict(hardness=1e-12 / (n * n) * n * n, offset=-0.002 + h.max() - 0.01 * np.ptp(h), maxiter=1000,
              pentol=1e-6)
    sub = PeriodicFFTElasticHalfSpace((n, n), 1.0, (float(n), float(n)))
    rg = ccg_gpu(sub, h, **kw)
    rc = ccg_cpu(PeriodicHalfSpace((n, n), 1.0, (float(n), float(n))), h, **kw)
    assert rg.success and rc.success
    assert np.sum(rg.active_set != rc.active_set) <= 2
    assert rel(rg.jac, rc.jac) < 1e-6


def test_system_api(gpu):
    n = 64
    h = paraboloid(n)
    sys_ = make_system(PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0)), Topography(h, (1.0, 1.0)))
    res = sys_.minimize_proxy(external_force=1e-3, pentol=1e-10)
    assert res.success
    np.testing.assert_allclose(sys_.compute_normal_force(), 1e-3, rtol=1e-9)
    assert sys_.compute_contact_area() == np.sum(res.jac > 0) * sys_.area_per_pt
    gap = sys_.compute_gap(sys_.disp, sys_.offset)
    assert np.min(gap) > -1e-9


@pytest.mark.parametrize("bc,hard", [("periodic", False), ("free", False), ("free", True)])
def test_fp32_solve_matches_oracle(gpu, bc, hard):
    """FP32 transforms and conjugate directions, FP64 forces, gaps and
    reductions (displacements updated by the FP32 transform of the force
    change, recomputed with FP64 transforms every 8 iterations): at tight
    tolerances the solve converges like the FP64 oracle and the pressures
    agree to the FP32 rounding of the returned forces."""
    n = 256 if bc == "periodic" else 128
    h = rough(n, 31)
    gcls, ocls = ((PeriodicFFTElasticHalfSpace, PeriodicHalfSpace) if bc == "periodic"
                  else (FreeFFTElasticHalfSpace, FreeHalfSpace))
    kw = dict(external_force=0.01, pentol=1e-11, forcetol=1e-14, thermotol=1e-10, maxiter=20000)
    if hard:
        kw["hardness"] = 0.05 / (n * n)
    rg = ccg_gpu(gcls((n, n), 1.0, (1.0, 1.0), precision="fp32"), h, **kw)
    rc = ccg_cpu(ocls((n, n), 1.0, (1.0, 1.0)), h, **kw)
    assert rg.success and rc.success
    l2 = np.linalg.norm(rg.jac - rc.jac) / np.linalg.norm(rc.jac)
    print(f"fp32 {bc} hard={hard}: nit {rg.nit} (oracle {rc.nit}), L2 {l2:.2e}, Linf {rel(rg.jac, rc.jac):.2e}")
    assert l2 < 1e-6 and rel(rg.jac, rc.jac) < 1e-6
    assert np.sum((rg.jac > 0) != (rc.jac > 0)) <= 2
    np.testing.assert_allclose(rg.offset, rc.offset, rtol=1e-7)
    # elastic solves take the oracle's iteration count up to a few; with a
    # hardness the mixing steps' bisection counts pixels, so round-off can
    # shift when the solver switches between mixing and PK steps
    assert abs(rg.nit - rc.nit) <= (max(3, rc.nit // 20) if not hard else max(8, rc.nit // 5))


@pytest.mark.parametrize("bc", ["periodic", "free"])
def test_mixed_precision_reaches_fp64_accuracy(gpu, bc):
    # FP32 bulk iterations + FP64 finish: within 1e-6 (L2, relative) of the
    # FP64 solve at the same tight tolerance; pure FP32 stalls near 1e-5
    n = 128
    h = rough(n, 21)
    cls = PeriodicFFTElasticHalfSpace if bc == "periodic" else FreeFFTElasticHalfSpace
    kw = dict(external_force=0.01, pentol=1e-11, forcetol=1e-14, thermotol=1e-10, maxiter=20000)
    r64 = ccg_gpu(cls((n, n), 1.0, (1.0, 1.0), precision="fp64"), h, **kw)
    rmx = ccg_gpu(cls((n, n), 1.0, (1.0, 1.0), precision="mixed"), h, **kw)
    assert r64.success and rmx.success
    assert rmx.nit == rmx.nit_fp32 + rmx.nit_fp64 and rmx.nit_fp32 > 0
    err = np.linalg.norm(rmx.jac - r64.jac) / np.linalg.norm(r64.jac)
    assert err < 1e-6, err


@pytest.mark.parametrize("bc", ["periodic", "free"])
def test_full_size_iterations_match_oracle(gpu, bc):
    """At the benchmark's grid size (2048^2 computational points: C2 periodic,
    C3-shaped free 1024^2 topography) a fixed number of iterations with
    unreachable tolerances follows the oracle's trajectory."""
    from contactmechanics_amd import workloads
    if bc == "periodic":
        n, h = 2048, workloads.config2()["heights"]
        gcls, ocls = PeriodicFFTElasticHalfSpace, PeriodicHalfSpace
        F = 0.01
    else:
        n, h = 1024, workloads.self_affine(1024, 0.7, 0.2, 1024)
        gcls, ocls = FreeFFTElasticHalfSpace, FreeHalfSpace
        F = 1e-3
    kw = dict(external_force=F, pentol=0.0, forcetol=0.0, thermotol=0.0, maxiter=6)
    rg = ccg_gpu(gcls((n, n), 1.0, (1.0, 1.0)), Topography(h, (1.0, 1.0)), **kw)
    rc = ccg_cpu(ocls((n, n), 1.0, (1.0, 1.0)), h, **kw)
    assert rg.nit == rc.nit == 6 and not rg.success and not rc.success
    # periodic: same FFT conventions, round-off only (2e-14 after 30
    # iterations at this size); free: the Love-kernel spectrum carries ~1e-11
    # relative FFT round-off, so a pixel at the contact edge may flip
    tol = 1e-11 if bc == "periodic" else 1e-8
    assert rel(rg.jac, rc.jac) < tol
    assert rel(rg.x, rc.x) < tol
    np.testing.assert_allclose(rg.offset, rc.offset, rtol=tol)
    flips = int(np.sum(rg.active_set != rc.active_set))
    assert flips == 0 if bc == "periodic" else flips <= 2


def test_compensated_sums_match_fsum(gpu):
    """The FP64 sums of the solver are compensated (Neumaier per thread,
    TwoSum across lanes and blocks): the total force and the elastic energy
    of a converged solve equal math.fsum of the same terms to within 2 eps of
    their magnitude sum."""
    import math

    from contactmechanics_amd.session import CcgSession, make_params
    n = 256
    h = rough(n, 41)
    sub = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0), precision="fp64")
    s = CcgSession(sub, [h], [make_params(external_force=0.01, pentol=1e-10, forcetol=1e-14, maxiter=5000)])
    st = s.run()[0]
    assert st["status"] == 0
    f = s.f[0].cpu().numpy()[:n, :n].ravel()
    u = s.u[0].cpu().numpy()[:n, :n].ravel()
    s.close()
    eps = np.finfo(np.float64).eps
    tf = -math.fsum(f)
    assert abs(st["total_force"] - tf) <= 2 * eps * math.fsum(np.abs(f)), (st["total_force"], tf)
    prods = f * u  # the same IEEE products the kernel forms
    fun = -math.fsum(prods) / 2
    assert abs(st["fun"] - fun) <= 2 * eps * math.fsum(np.abs(prods)) / 2, (st["fun"], fun)
