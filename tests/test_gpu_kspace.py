"""k-space helpers on the GPU against the oracle, reference
``FFTElasticHalfSpace.py:613-885``: the forward transform into the
reference's fourier layout, evaluate_k_disp / evaluate_k_force /
evaluate_k_force_k, the Hermitian-weighted scalar product and the elastic
energies, periodic and free, FP64 and FP32; and the Hertz energy check of the
reference (``test/test_hertz.py:117-125``): the k-space and the real-space
energy of a converged free Hertz contact match the closed form to 1 %.
"""

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace, Topography, make_system
from oracle import reference_solutions as hz
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("bc,res", [("periodic", (96, 64)), ("periodic", (75, 48)), ("free", (40, 30))])
def test_kspace_helpers_match_oracle(gpu, precision, bc, res):
    gcls, ocls = ((PeriodicFFTElasticHalfSpace, PeriodicHalfSpace) if bc == "periodic"
                  else (FreeFFTElasticHalfSpace, FreeHalfSpace))
    g = gcls(res, 2.2, (1.3, 2.1), precision=precision)
    c = ocls(res, 2.2, (1.3, 2.1))
    rng = np.random.default_rng(sum(res))
    f, u = rng.standard_normal(g.nb_domain_grid_pts), rng.standard_normal(g.nb_domain_grid_pts)
    tol = 1e-12 if precision == "fp64" else 2e-5
    assert g.fft(f).shape == c.nb_fourier_grid_pts
    assert rel(g.fft(f), c.fft(f)) < tol
    assert rel(g.evaluate_k_disp(f), c.evaluate_k_disp(f)) < tol * (1 if bc == "periodic" else 100)
    assert rel(g.evaluate_k_force(u), c.evaluate_k_force(u)) < tol * (1 if bc == "periodic" else 100)
    ku = c.fft(u)
    assert rel(g.evaluate_k_force_k(ku), c.evaluate_k_force_k(ku)) < tol * (1 if bc == "periodic" else 100)
    kf = c.fft(f)
    sp, sp_ref = g.evaluate_scalar_product_k_space(kf, ku), c.evaluate_scalar_product_k_space(kf, ku)
    assert abs(sp - sp_ref) < tol * np.sum(np.abs(f * u))
    e, e_ref = g.evaluate_elastic_energy(f, u), c.evaluate_elastic_energy(f, u)
    assert abs(e - e_ref) < tol * np.sum(np.abs(f * u))
    for forces in (False, True):
        ge = g.evaluate(u, pot=True, forces=forces)[0]
        ce = c.evaluate(u, pot=True, forces=forces)[0]
        np.testing.assert_allclose(ge, ce, rtol=tol * (10 if bc == "periodic" else 1000))


def test_kspace_device_tensors_stay_on_device(gpu):
    import torch
    g = PeriodicFFTElasticHalfSpace((64, 64), 1.0, (1.0, 1.0))
    f = torch.randn(64, 64, dtype=torch.float64, device="cuda")
    k = g.evaluate_k_disp(f)
    assert k.is_cuda and k.dtype == torch.complex128 and tuple(k.shape) == (33, 64)
    e, fk = g.evaluate_k(k, pot=True, forces=False)
    assert fk is not None and fk.is_cuda  # the reference returns the forces whenever pot is set


@pytest.mark.parametrize("kw", [dict(offset=0.1), dict(external_force=15.0)])
def test_hertz_energy_kspace_and_realspace(gpu, kw):
    n, r_s, E_s, sx = 256, 20.0, 102.0, 5.0
    x = ((np.arange(n) - n / 2) * sx / n)[:, None]
    y = ((np.arange(n) - n / 2) * sx / n)[None, :]
    h = -(x ** 2 + y ** 2) / (2 * r_s)
    sub = FreeFFTElasticHalfSpace((n, n), E_s, (sx, sx))
    system = make_system(sub, Topography(h, (sx, sx)))
    res = system.minimize_proxy(pentol=1e-10, **kw)
    assert res.success
    d = kw.get("offset", None)
    if d is None:
        d = hz.hertz_penetration(kw["external_force"], r_s, E_s)
    e_ref = hz.hertz_elastic_energy(d, r_s, E_s)
    e_k = sub.evaluate(res.x, pot=True, forces=False)[0]
    e_r = sub.evaluate(res.x, pot=True, forces=True)[0]
    np.testing.assert_allclose(e_k, e_ref, rtol=1e-2)
    np.testing.assert_allclose(e_r, e_ref, rtol=1e-2)
    np.testing.assert_allclose(e_k, e_r, rtol=1e-10)
