"""Pin the oracle's semi-periodic half-space (CPU), reference
``FFTElasticHalfSpace.py:1286-1615``, by what its definition fixes:

* no periodic direction and no images: the free (2N padded) half-space on
  the topography (the padding only has to exceed 2n - 1; reference
  ``test/test_fft_elastic_half_space.py:652-674``);
* exchanging the periodic direction together with the grid (reference :582-604);
* periodic in both directions: converges to the periodic half-space (with
  ``stiffness_q0 = 0``, rigid motion removed) as images are added (reference
  :676-790);
* a brute-force real-space sum of the image-summed Love kernel over all
  pairs of points on a tiny grid;
* units: lengths x100 and modulus x1e-4 scale the displacement by 100
  (reference :630-650);
* evaluate_force inverts evaluate_disp on the padded grid (reference :606-628).
"""

import numpy as np
import pytest

from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace, SemiPeriodicHalfSpace, love_rectangle_kernel


def patch_force(res, value=1e8):
    f = np.zeros(res)
    f[: res[0] // 2, : res[1] // 2] = value
    return f


def test_no_periodicity_no_images_is_free():
    res, size, E = (16, 16), (0.1, 0.1), 210e9
    semi = SemiPeriodicHalfSpace(res, E, size, periodicity=(False, False), n_images=0)
    free = FreeHalfSpace(res, E, size)
    f = patch_force(res)
    a, b = semi.evaluate_disp(f), free.evaluate_disp(f)
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("grid", [(12, 12), (12, 13), (13, 12), (13, 13)])
def test_direction_exchange(grid):
    size, E = (0.1, 0.1), 210e9
    h1 = SemiPeriodicHalfSpace(grid, E, size, periodicity=(True, False), n_images=4)
    h2 = SemiPeriodicHalfSpace(grid[::-1], E, size, periodicity=(False, True), n_images=4)
    f = patch_force(grid)
    a, b = h1.evaluate_disp(f), h2.evaluate_disp(f.T).T
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("res", [(100, 101), (101, 100)])
def test_periodic_limit(res):
    size, E = (0.1, 0.1), 210e9
    f = patch_force(res)
    per = PeriodicHalfSpace(res, E, size, stiffness_q0=0.0).evaluate_disp(f)
    errs = []
    for n in (0, 5, 20):
        u = SemiPeriodicHalfSpace(res, E, size, periodicity=(True, True), n_images=n).evaluate_disp(f)
        u = u - u.mean()
        errs.append(np.mean(np.abs(u - per)))
    assert errs[0] > errs[1] > errs[2]
    denom = np.maximum(np.abs(per), np.mean(np.abs(per)))
    assert np.max(np.abs(u - per) / denom) < 1e-2


@pytest.mark.parametrize("per", [(False, True), (True, False), (True, True), (False, False)])
def test_brute_force_image_sum(per):
    res, size, E, nimg = (4, 5), (1.0, 1.3), 2.0, 2
    h = SemiPeriodicHalfSpace(res, E, size, periodicity=per, n_images=nimg)
    mx, my = h.nb_domain_grid_pts
    dx, dy = size[0] / res[0], size[1] / res[1]
    f = np.random.default_rng(1).standard_normal((mx, my))
    u = np.zeros((mx, my))

    def wrap(d, m):  # numpy fftfreq convention: 0..(m-1)//2 positive, the rest negative
        d %= m
        return d if d <= (m - 1) // 2 else d - m

    imx = range(-nimg, nimg + 1) if per[0] else [0]
    imy = range(-nimg, nimg + 1) if per[1] else [0]
    for i in range(mx):
        for j in range(my):
            for k in range(mx):
                for m in range(my):
                    x, y = wrap(i - k, mx) * dx, wrap(j - m, my) * dy
                    kern = sum(love_rectangle_kernel(x + a * size[0], y + b * size[1], dx / 2, dy / 2, E)
                               for a in imx for b in imy)
                    u[i, j] += kern * (-f[k, m])
    u /= h.area_per_pt
    np.testing.assert_allclose(h.evaluate_disp(f), u, rtol=1e-10, atol=1e-12 * np.max(np.abs(u)))


def test_length_units():
    res = (12, 10)
    f = patch_force(res, 1e7)
    for per in [(False, True), (True, False), (True, True), (False, False)]:
        a = SemiPeriodicHalfSpace(res, 210e9, (0.1, 0.1), periodicity=per, n_images=3).evaluate_disp(f)
        b = SemiPeriodicHalfSpace(res, 210e5, (10.0, 10.0), periodicity=per, n_images=3).evaluate_disp(f)
        assert np.max(np.abs(a - b / 100)) < 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("per", [(False, True), (True, False), (False, False)])
def test_force_inverts_disp(per):
    res = (10, 11)
    h = SemiPeriodicHalfSpace(res, 210e9, (0.1, 0.1), periodicity=per, n_images=3)
    f = patch_force(res, 1e7)
    u = h.evaluate_disp(f, bIncludePadding=True)
    assert u.shape == h.nb_domain_grid_pts
    rec = h.evaluate_force(u)
    fp = np.zeros(h.nb_domain_grid_pts)
    fp[: res[0], : res[1]] = f
    np.testing.assert_allclose(rec, fp, atol=1e-6 * 1e7)
