"""GPU parity of the elastic half-space operators against the oracle.

Tolerances: FP64 plans agree with the NumPy FP64 oracle to round-off
(periodic 1e-13 relative; free 1e-9 for the force, whose stiffness 1/G
amplifies the FFT round-off of the real-space Love kernel); FP32 plans to
2e-6 relative (single-precision transforms).
"""

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

pytestmark = pytest.mark.gpu

SIZES = [(16, 16), (64, 32), (48, 40), (96, 160), (128, 120), (256, 256)]


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_periodic_operators(gpu, nx, ny, prec):
    rng = np.random.default_rng(nx * 1000 + ny)
    g = PeriodicFFTElasticHalfSpace((nx, ny), 1.3, (2.0, 3.0), precision=prec)
    c = PeriodicHalfSpace((nx, ny), 1.3, (2.0, 3.0))
    f = rng.standard_normal((nx, ny))
    tol = 1e-13 if prec == "fp64" else 2e-6
    assert rel(g.evaluate_disp(f), c.evaluate_disp(f)) < tol
    assert rel(g.evaluate_force(f), c.evaluate_force(f)) < tol


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_free_operators(gpu, nx, ny, prec):
    rng = np.random.default_rng(nx * 1000 + ny + 7)
    g = FreeFFTElasticHalfSpace((nx, ny), 1.3, (2.0, 3.0), precision=prec)
    c = FreeHalfSpace((nx, ny), 1.3, (2.0, 3.0))
    f = rng.standard_normal(g.nb_domain_grid_pts)
    tol_d, tol_f = (1e-10, 1e-9) if prec == "fp64" else (2e-6, 2e-6)
    assert rel(g.evaluate_disp(f), c.evaluate_disp(f)) < tol_d
    assert rel(g.evaluate_force(f), c.evaluate_force(f)) < tol_f
    # unpadded input: zero-padded, output cropped unless bIncludePadding
    fu = rng.standard_normal((nx, ny))
    assert g.evaluate_disp(fu).shape == (nx, ny)
    assert rel(g.evaluate_disp(fu), c.evaluate_disp(fu)) < tol_d
    assert rel(g.evaluate_disp(fu, bIncludePadding=True), c.evaluate_disp(fu, bIncludePadding=True)) < tol_d


@pytest.mark.parametrize("kw", [dict(), dict(stiffness_q0=0.0), dict(stiffness_q0=1.43),
                                dict(thickness=3.45, poisson=0.3)])
def test_periodic_green_function_matches_oracle(gpu, kw):
    g = PeriodicFFTElasticHalfSpace((32, 24), 2.1, (3.0, 2.0), **kw)
    c = PeriodicHalfSpace((32, 24), 2.1, (3.0, 2.0), **kw)
    np.testing.assert_allclose(g.greens_function, c.greens_function, rtol=1e-13, atol=0)
    u = -np.ones((32, 24))
    np.testing.assert_allclose(g.evaluate_force(u), c.evaluate_force(u), rtol=1e-12, atol=1e-14)


def test_free_green_function_matches_oracle(gpu):
    g = FreeFFTElasticHalfSpace((32, 24), 1.5, (100.0, 200.0))
    c = FreeHalfSpace((32, 24), 1.5, (100.0, 200.0))
    np.testing.assert_allclose(g.greens_function, c.greens_function.real,
                               rtol=1e-11, atol=1e-13 * np.max(np.abs(c.greens_function)))


def test_torch_device_path_and_errors(gpu):
    import torch
    g = PeriodicFFTElasticHalfSpace((64, 64), 1.0, (1.0, 1.0))
    c = PeriodicHalfSpace((64, 64), 1.0, (1.0, 1.0))
    f = np.random.default_rng(1).standard_normal((64, 64))
    ud = g.evaluate_disp(torch.from_numpy(f).to(gpu))
    assert ud.is_cuda and ud.dtype == torch.float64
    assert rel(ud.cpu().numpy(), c.evaluate_disp(f)) < 1e-13
    with pytest.raises(g.Error):
        g.evaluate_disp(np.zeros((63, 64)))
    # a row of 16411 complex points (prime) needs more than 512 threads x 32
    # elements: the plan refuses loudly (no CPU fallback)
    with pytest.raises(PeriodicFFTElasticHalfSpace.Error):
        PeriodicFFTElasticHalfSpace((4, 2 * 16411), 1.0, (1.0, 1.0)).evaluate_disp(np.zeros((4, 2 * 16411)))


@pytest.mark.parametrize("nx,ny", [(4000, 32), (32, 4000), (2430, 48)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_long_mixed_radix_lines(gpu, nx, ny, prec):
    # free grids of 8000-point lines (8000 = 16*4*5^3) take the 16/32-element
    # kernels with blocks of more than 256 threads; 2430 = 2*3^5*5 periodic
    rng = np.random.default_rng(nx + 3 * ny)
    if nx == 2430:
        g = PeriodicFFTElasticHalfSpace((nx, ny), 1.3, (2.0, 3.0), precision=prec)
        c = PeriodicHalfSpace((nx, ny), 1.3, (2.0, 3.0))
        f = rng.standard_normal((nx, ny))
        tol_d = tol_f = 1e-12 if prec == "fp64" else 3e-6
    else:
        g = FreeFFTElasticHalfSpace((nx, ny), 1.3, (2.0, 3.0), precision=prec)
        c = FreeHalfSpace((nx, ny), 1.3, (2.0, 3.0))
        f = rng.standard_normal(g.nb_domain_grid_pts)
        # evaluate_force divides by the Love-kernel spectrum, whose smallest
        # (high-frequency) values carry the FFT round-off of an 8000-point
        # line relative to its largest: 2e-8 observed, against 1e-12 for disp
        tol_d, tol_f = (1e-9, 1e-7) if prec == "fp64" else (3e-6, 3e-6)
    assert rel(g.evaluate_disp(f), c.evaluate_disp(f)) < tol_d
    assert rel(g.evaluate_force(f), c.evaluate_force(f)) < tol_f
