"""Pin the oracle's half-space operators (CPU).

Each check is independent of the oracle's own formula: brute-force DFT sums,
closed-form sine-wave laws, Love's closed form for a loaded rectangle, a dense
real-space convolution, and the reference's recorded Hertz values.
"""

import numpy as np
import pytest

from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace, love_rectangle_kernel
from oracle import reference_solutions as rs


def test_fft_layout_is_brute_force_dft():
    # reference FFTElasticHalfSpace.py:119-128 / test_fft_elastic_half_space.py:356-371:
    # the half spectrum is r2c along axis 0, shape (nx//2+1, ny), unnormalised
    rng = np.random.default_rng(3)
    nx, ny = 6, 5
    a = rng.standard_normal((nx, ny))
    spec = np.fft.rfft2(a, axes=(1, 0))
    assert spec.shape == (nx // 2 + 1, ny)
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    for k in range(nx // 2 + 1):
        for m in range(ny):
            ref = np.sum(a * np.exp(-2j * np.pi * (k * i / nx + m * j / ny)))
            assert abs(spec[k, m] - ref) < 1e-12


@pytest.mark.parametrize("nx,ny", [(8, 15), (8, 4), (9, 4), (113, 765)])
def test_periodic_sine_wave_pressure(nx, ny):
    # reference test_mpi_periodic_elastic_half_space.py:76-161: p = -E*/2 q u
    sx, sy, E = 2.45, 1.0, 1.0
    hs = PeriodicHalfSpace((nx, ny), E, (sx, sy))
    for k in [(1, 0), (0, 1), (1, 2), (nx // 2, 0), (1, ny // 2), (0, 2), (nx // 2, ny // 2), (0, ny // 2)]:
        qx, qy = k[0] * 2 * np.pi / sx, k[1] * 2 * np.pi / sy
        q = np.hypot(qx, qy)
        Y, X = np.meshgrid(np.linspace(0, sy, ny + 1)[:-1], np.linspace(0, sx, nx + 1)[:-1])
        disp = np.cos(qx * X + qy * Y) + np.sin(qx * X + qy * Y)
        p = hs.evaluate_force(disp) / hs.area_per_pt
        np.testing.assert_allclose(p, -disp * E / 2 * q, atol=1e-10 * nx * ny, rtol=1e-7)


@pytest.mark.parametrize("nx,ny", [(64, 33), (65, 32), (64, 64)])
def test_periodic_sine_wave_displacement(nx, ny):
    # reference test_mpi_periodic_elastic_half_space.py:253-299: u = -2 p/(E* q)
    sx, sy, E = 2.0, 1.0, 1.0
    Y, X = np.meshgrid(np.linspace(0, sy, ny + 1)[:-1], np.linspace(0, sx, nx + 1)[:-1])
    qx, qy = 2 * np.pi / sx, 4 * 2 * np.pi / sy
    p = np.cos(qx * X + qy * Y)
    hs = PeriodicHalfSpace((nx, ny), E, (sx, sy))
    u = hs.evaluate_disp(p * hs.area_per_pt)
    np.testing.assert_allclose(u, -p / E * 2 / np.hypot(qx, qy), atol=1e-7, rtol=1e-10)


@pytest.mark.parametrize("res", [(32, 32), (33, 32), (32, 33)])
def test_periodic_reversibility(res):
    # reference test_fft_elastic_half_space.py:132-158 (zero-mean fields)
    rng = np.random.default_rng(0)
    hs = PeriodicHalfSpace(res, 3.3, (1.5, 2.5))
    d = rng.random(res)
    d -= d.mean()
    np.testing.assert_allclose(hs.evaluate_disp(hs.evaluate_force(d)), d, atol=1e-10)
    f = rng.random(res)
    f -= f.mean()
    np.testing.assert_allclose(hs.evaluate_force(hs.evaluate_disp(f)), f, atol=1e-10)


def test_periodic_stiffness_q0_laws():
    # reference test_fft_elastic_half_space.py:291-309
    res, E, size = (16, 12), 2.1, (3.0, 2.0)
    hs = PeriodicHalfSpace(res, E, size, stiffness_q0=1.43)
    f = hs.evaluate_force(-np.ones(res))
    assert abs(f.sum() / np.prod(size) - 1.43) < 1e-12
    nu, h0 = 0.3, 3.45
    hs = PeriodicHalfSpace(res, E, size, thickness=h0, poisson=nu)
    f = hs.evaluate_force(-np.ones(res))
    M = (1 - nu) / ((1 - 2 * nu) * (1 + nu)) * E
    assert abs(f.sum() / np.prod(size) - M / h0) < 1e-10
    # stiffness_q0 None: mean of the two lowest stiffnesses (:546-551)
    hs = PeriodicHalfSpace(res, E, size)
    s = 1.0 / hs.greens_function[0, 0]
    assert abs(s - (np.pi * E / size[0] + np.pi * E / size[1]) / 2) < 1e-12


@pytest.mark.parametrize("nx,ny", [(64, 32), (65, 33)])
def test_free_uniform_pressure_patch_love(nx, ny):
    # reference test_mpi_free_elastic_half_space.py:128-186: superposition of
    # pixels equals Love's closed form for the whole loaded rectangle
    sx, sy, E = 100.0, 200.0, 1.5
    forces = np.zeros((2 * nx, 2 * ny))
    x = ((np.arange(2 * nx) - (nx - 1) / 2) * sx / nx)[:, None]
    y = ((np.arange(2 * ny) - (ny - 1) / 2) * sy / ny)[None, :]
    forces[1:nx - 1, 1:ny - 1] = -(sx * sy) / (nx * ny)
    a = (nx - 2) / 2 * sx / nx
    b = (ny - 2) / 2 * sy / ny
    ref = love_rectangle_kernel(x, y, a, b, E)
    hs = FreeHalfSpace((nx, ny), E, (sx, sy))
    u = hs.evaluate_disp(forces)
    np.testing.assert_allclose(u[1:nx - 1, 1:ny - 1], ref[1:nx - 1, 1:ny - 1], rtol=1e-7)


def test_free_matches_dense_real_space_convolution():
    # 2N zero padding makes the FFT convolution exactly the aperiodic sum
    rng = np.random.default_rng(5)
    nx, ny, sx, sy, E = 7, 6, 1.4, 0.9, 2.0
    hs = FreeHalfSpace((nx, ny), E, (sx, sy))
    dx, dy = sx / nx, sy / ny
    f = rng.standard_normal((nx, ny))
    u = np.zeros((nx, ny))
    for i in range(nx):
        for j in range(ny):
            for k in range(nx):
                for m in range(ny):
                    u[i, j] -= love_rectangle_kernel((i - k) * dx, (j - m) * dy, dx / 2, dy / 2, E) * f[k, m]
    u /= hs.area_per_pt
    np.testing.assert_allclose(hs.evaluate_disp(f), u, rtol=1e-12, atol=1e-14)


def test_free_green_is_real_positive_even():
    hs = FreeHalfSpace((32, 24), 1.0, (1.0, 1.0))
    g = hs.greens_function
    assert np.max(np.abs(g.imag)) < 1e-12 * np.max(np.abs(g.real))
    assert np.all(g.real > 0)
    # even in the full (non-r2c) axis
    np.testing.assert_allclose(g.real[:, 1:], g.real[:, 1:][:, ::-1], rtol=1e-12)


def test_hertz_recorded_values():
    # reference test/ReferenceSolutions/test_reference_solutions.py:83-96
    np.testing.assert_allclose(rs.hertz_penetration(7, 4, 5), 0.6507879989531481, rtol=1e-7)
    np.testing.assert_allclose(rs.hertz_radius_and_pressure(7, 4, 5),
                               (1.6134286460245437, 1.2839257217043503), rtol=1e-7)
    Es, R, F = 3.0, 7.0, 11.0
    d = rs.hertz_penetration(F, R, Es)
    np.testing.assert_allclose(rs.hertz_normal_load(d, R, Es), F, rtol=1e-12)
    ds = np.linspace(0, 3, 2001)
    np.testing.assert_allclose(rs.hertz_elastic_energy(ds[-1], 5, 7),
                               np.trapezoid(rs.hertz_normal_load(ds, 5, 7), x=ds), rtol=1e-5)
