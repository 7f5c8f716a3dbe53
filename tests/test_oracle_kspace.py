"""Pin the oracle's k-space helpers (CPU), reference
``FFTElasticHalfSpace.py:613-885``.

* one sine-wave mode: the k-space displacement field has the two DFT
  coefficients N (1 -+ i)/2 at +-(k0, k1), the k-space pressure is
  ``-E*/2 |q| u_k`` and the elastic energy ``E* |q| sx sy / 4`` in both real
  and k-space (reference ``test/test_mpi_periodic_elastic_half_space.py:120-170``);
* the power theorem: the half-spectrum scalar product equals ``sum a b``
  for random real fields, also on grids with odd sides (brute force);
* real-space and k-space energies agree for random displacements
  (reference ``test/test_fft_elastic_half_space.py:209-220``, same grids);
* the k-space displacement is the spectrum of ``evaluate_disp`` (inverse DFT);
* free half-space: the k-space energy is real and equals the real-space one.
"""

import numpy as np
import pytest

from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace


@pytest.mark.parametrize("nx,ny,k0,k1", [(16, 12, 3, 2), (15, 20, 2, 5), (32, 9, 0, 3)])
def test_sine_wave_kspace_laws(nx, ny, k0, k1):
    sx, sy, Es = 2.5, 1.5, 3.3
    hs = PeriodicHalfSpace((nx, ny), Es, (sx, sy))
    X, Y = np.meshgrid(np.arange(nx) * sx / nx, np.arange(ny) * sy / ny, indexing="ij")
    qx, qy = 2 * np.pi * k0 / sx, 2 * np.pi * k1 / sy
    q = np.hypot(qx, qy)
    disp = np.cos(qx * X + qy * Y) + np.sin(qx * X + qy * Y)
    expected = np.zeros(hs.nb_fourier_grid_pts, dtype=complex)
    expected[k0, k1] += (0.5 - 0.5j) * nx * ny
    if k0 == 0:  # the mirror (-k0, -k1) lies in the half spectrum too
        expected[0, (-k1) % ny] += (0.5 + 0.5j) * nx * ny
    np.testing.assert_allclose(hs.fft(disp), expected, atol=1e-9 * nx * ny)
    kp = hs.evaluate_k_force(disp) / hs.area_per_pt
    np.testing.assert_allclose(kp, -Es / 2 * q * expected, atol=1e-9 * nx * ny * Es * q)
    e_ref = Es * q * sx * sy / 4
    np.testing.assert_allclose(hs.evaluate(disp, pot=True, forces=False)[0], e_ref, rtol=1e-12)
    np.testing.assert_allclose(hs.evaluate(disp, pot=True, forces=True)[0], e_ref, rtol=1e-12)
    e_k, f_k = hs.evaluate_k(hs.fft(disp), pot=True, forces=True)
    np.testing.assert_allclose(e_k, e_ref, rtol=1e-12)
    np.testing.assert_allclose(f_k, hs.evaluate_k_force(disp), atol=1e-12 * np.max(np.abs(f_k)))


@pytest.mark.parametrize("nx,ny", [(8, 8), (9, 8), (8, 7), (5, 3), (2, 6), (3, 4)])
def test_power_theorem_brute_force(nx, ny):
    rng = np.random.default_rng(nx * 10 + ny)
    a, b = rng.standard_normal((nx, ny)), rng.standard_normal((nx, ny))
    hs = PeriodicHalfSpace((nx, ny), 1.0, (1.0, 1.0))
    direct = sum(a[i, j] * b[i, j] for i in range(nx) for j in range(ny))
    np.testing.assert_allclose(hs.evaluate_scalar_product_k_space(hs.fft(a), hs.fft(b)), direct, rtol=1e-12)


@pytest.mark.parametrize("res", [(16, 16), (16, 15), (15, 16), (15, 9)])
def test_real_and_kspace_energy_agree(res):
    disp = np.random.default_rng(sum(res)).normal(size=res)
    hs = PeriodicHalfSpace(res, 1.7, (2.0, 3.0))
    np.testing.assert_allclose(hs.evaluate(disp, pot=True, forces=True)[0],
                               hs.evaluate(disp, pot=True, forces=False)[0], rtol=1e-12)


@pytest.mark.parametrize("cls,res", [(PeriodicHalfSpace, (12, 10)), (FreeHalfSpace, (6, 5))])
def test_kdisp_is_spectrum_of_evaluate_disp(cls, res):
    hs = cls(res, 2.0, (1.0, 1.5))
    f = np.random.default_rng(3).standard_normal(hs.nb_domain_grid_pts)
    u = hs.evaluate_disp(f)
    np.testing.assert_allclose(hs.evaluate_k_disp(f), hs.fft(u), atol=1e-12 * np.max(np.abs(hs.fft(u))))


def test_free_energy_real_and_parseval():
    hs = FreeHalfSpace((8, 6), 3.0, (2.0, 1.0))
    f = np.zeros(hs.nb_domain_grid_pts)
    f[:8, :6] = np.random.default_rng(5).random((8, 6))
    f[:8, :6] -= f[:8, :6].mean()
    u = hs.evaluate_disp(f)
    e_real = hs.evaluate_elastic_energy(f, u)
    e_k = hs.evaluate_elastic_energy_k_space(hs.fft(f), hs.evaluate_k_disp(f))
    np.testing.assert_allclose(e_k, e_real, rtol=1e-12)
    assert e_real > 0
