"""Grid maps for odd contiguous dimensions (CPU): the plan grid's fields and
fourier layouts against numpy's real FFT of the caller's grid, and the
tiled-period convolution identity with a brute-force circular convolution."""

import numpy as np
import pytest
import torch

from contactmechanics_amd.gridmap import GridMap


def rfft_half(a):
    return np.fft.rfft2(a, axes=(1, 0))


CASES = [((8, 15), (1, 1)), ((64, 33), (1, 1)), ((9, 15), (1, 1)), ((113, 45), (1, 1)), ((11, 7), (0, 1)),
         ((7, 11), (1, 0)), ((12, 10), (1, 1)), ((13, 9), (0, 0))]


@pytest.mark.parametrize("shape,per", CASES)
def test_fields_and_fourier_layouts(shape, per):
    m = GridMap(*shape, *per)
    assert m.plan_shape[1] % 2 == 0
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    ap = m.fwd(torch.from_numpy(a))
    assert tuple(ap.shape) == m.plan_shape
    np.testing.assert_array_equal(m.inv(ap).numpy(), a)
    F, Fp = rfft_half(a), rfft_half(ap.numpy())
    np.testing.assert_allclose(m.fourier_fwd(torch.from_numpy(F)).numpy(), Fp, atol=1e-12 * np.abs(F).max())
    np.testing.assert_allclose(m.fourier_inv(torch.from_numpy(Fp)).numpy(), F, atol=1e-12 * np.abs(F).max())
    b = rng.standard_normal(shape)
    assert m.sum_inv(float((ap * m.fwd(torch.from_numpy(b))).sum())) == pytest.approx(float((a * b).sum()))


@pytest.mark.parametrize("shape,per", [((6, 9), (1, 1)), ((9, 7), (1, 1)), ((5, 7), (0, 1)), ((7, 5), (1, 0)),
                                       ((7, 9), (0, 0))])
def test_tiled_convolution_is_the_one_period_convolution(shape, per):
    """A kernel k on the caller's grid; the plan convolves with the mapped
    kernel spectrum (one-period G at even ky, zero at odd): the result is
    the caller's circular convolution, mapped (brute force)."""
    m = GridMap(*shape, *per)
    rng = np.random.default_rng(3)
    k = rng.standard_normal(shape)
    k = 0.25 * (k + k[::-1] + k[:, ::-1] + k[::-1, ::-1])  # even in both axes: real spectrum
    k = np.roll(np.roll(k, 1, 0), 1, 1)  # ... about index 0
    k = 0.25 * (k + np.roll(k[::-1], 1, 0) + np.roll(k[:, ::-1], 1, 1) + np.roll(np.roll(k[::-1, ::-1], 1, 0), 1, 1))
    f = rng.standard_normal(shape)
    mx, my = shape
    direct = np.zeros(shape)
    for i in range(mx):
        for j in range(my):
            direct += f[i, j] * np.roll(np.roll(k, i, 0), j, 1)
    G = rfft_half(k).real
    Gfull = np.fft.fft2(k).real
    # the plan's kernel: G of the caller at the plan's frequencies
    pmx, pmy = m.plan_shape
    if m.transpose:
        Gfull = Gfull.T
    Gp = np.zeros((pmx, pmy))
    Gp[:, :: m.tile] = Gfull
    fp = m.fwd(torch.from_numpy(f)).numpy()
    up = np.fft.irfft2(rfft_half(fp) * Gp[: pmx // 2 + 1], s=(pmy, pmx), axes=(1, 0))
    np.testing.assert_allclose(m.inv(torch.from_numpy(up)).numpy(), direct, atol=1e-12 * np.abs(direct).max())
    np.testing.assert_allclose(m.green_inv(torch.from_numpy(Gp[: pmx // 2 + 1])).numpy(), G, atol=1e-12)


def test_solvable():
    assert GridMap(8, 15, 0, 0).solvable  # transposed, no repetition
    assert GridMap(9, 15, 1, 1).solvable and GridMap(9, 15, 1, 0).solvable  # repeated along a periodic direction
    assert not GridMap(9, 15, 0, 0).solvable  # repeated along a free direction
