"""Pin the small host-side functions of the oracle and of the product (CPU).

* ``rms_height_from_area`` (SurfaceTopography, absent from the reference
  tree; used for the default ``pentol``, reference
  ``ConstrainedConjugateGradients.py:150-162``) by closed forms: a sinusoid of
  amplitude A over whole periods has RMS A/sqrt(2), a two-level field +-a has
  RMS a, a constant has RMS 0, and masked pixels do not count;
* the default ``pentol`` rule with its two fall-backs (flat surface: offset
  plus mean height over 1000; then 1e-3);
* ``check_boundaries`` of the free half-space (reference
  ``FFTElasticHalfSpace.py:1217-1269``): a force on any of the four edges of
  the topography block raises ``FreeBoundaryError``, interior forces and
  forces up to ``tol`` do not, and in a masked force array masked pixels are
  ignored and an edge that is masked entirely passes (:1236-1243).

The product's versions are checked on the same cases (their host logic needs
no GPU).
"""

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, Topography
from contactmechanics_amd.Optimization.ConstrainedConjugateGradients import default_pentol
from oracle.ccg import rms_height_from_area as oracle_rms
from oracle.halfspace import FreeHalfSpace


def product_rms(h):
    return Topography(h, (1.0, 1.0)).rms_height_from_area()


RMS_IMPLS = [oracle_rms, product_rms]


@pytest.mark.parametrize("rms", RMS_IMPLS)
@pytest.mark.parametrize("n,periods,A", [(64, 1, 1.0), (48, 3, 0.25), (100, 7, 3e-4)])
def test_rms_of_sinusoid(rms, n, periods, A):
    x = np.arange(n) / n
    h = A * np.sin(2 * np.pi * periods * x)[:, None] * np.ones((1, 5)) + 0.7  # offset does not count
    np.testing.assert_allclose(rms(h), A / np.sqrt(2), rtol=1e-13)


@pytest.mark.parametrize("rms", RMS_IMPLS)
def test_rms_of_two_levels_and_constant(rms):
    h = np.full((8, 6), -0.5)
    h[:4] = 1.5  # +-1 about the mean 0.5
    np.testing.assert_allclose(rms(h), 1.0, rtol=1e-15)
    assert rms(np.full((7, 9), 3.25)) == 0.0


@pytest.mark.parametrize("rms", RMS_IMPLS)
def test_rms_ignores_masked_pixels(rms):
    n = 32
    x = np.arange(n) / n
    h = np.tile(2.0 * np.cos(2 * np.pi * 2 * x)[:, None], (1, 10))
    garbage = h.copy()
    garbage[:, 3:6] = 1e6  # masked columns hold anything
    hm = np.ma.masked_array(garbage, mask=np.zeros_like(garbage, dtype=bool))
    hm.mask[:, 3:6] = True
    np.testing.assert_allclose(rms(hm), 2.0 / np.sqrt(2), rtol=1e-13)


def test_default_pentol_rule_and_fallbacks():
    n = 40
    x = np.arange(n) / n
    h = 0.3 * np.sin(2 * np.pi * x)[:, None] * np.ones((1, n))
    topo = Topography(h, (1.0, 1.0))
    np.testing.assert_allclose(default_pentol(topo, h, (n, n), None), (0.3 / np.sqrt(2)) / (10 * n),
                               rtol=1e-13)
    # a plain array (no rms_height_from_area): the same value
    np.testing.assert_allclose(default_pentol(h, h, (n, n), None), (0.3 / np.sqrt(2)) / (10 * n),
                               rtol=1e-13)
    flat = np.full((n, n), 0.25)  # exactly representable mean: RMS exactly 0
    np.testing.assert_allclose(default_pentol(flat, flat, (n, n), 0.1), (0.1 + 0.25) / 1000, rtol=1e-13)
    zero = np.zeros((n, n))
    assert default_pentol(zero, zero, (n, n), 0.0) == 1e-3
    # the topography's own rms_height_from_area is used when it has one

    class Rough:
        def rms_height_from_area(self):
            return 5.0

    np.testing.assert_allclose(default_pentol(Rough(), h, (n, n), None), 5.0 / (10 * n), rtol=1e-15)


def _substrates(nx, ny):
    return [FreeHalfSpace((nx, ny), 1.0, (1.0, 1.0)), FreeFFTElasticHalfSpace((nx, ny), 1.0, (1.0, 1.0))]


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("edge", ["left", "right", "top", "bottom"])
def test_check_boundaries_edge_forces_raise(which, edge):
    nx, ny = 12, 9
    hs = _substrates(nx, ny)[which]
    f = np.zeros((nx, ny))
    f[3:9, 2:7] = 1.0  # interior only: fine
    hs.check_boundaries(f)
    g = f.copy()
    idx = {"left": (5, 0), "right": (5, ny - 1), "top": (0, 4), "bottom": (nx - 1, 4)}[edge]
    g[idx] = 1e-3
    with pytest.raises(type(hs).FreeBoundaryError):
        hs.check_boundaries(g)
    hs.check_boundaries(g, tol=1e-3)  # |f| <= tol passes
    with pytest.raises(type(hs).FreeBoundaryError):
        hs.check_boundaries(g, tol=0.5e-3)


@pytest.mark.parametrize("which", [0, 1])
def test_check_boundaries_masked(which):
    nx, ny = 10, 10
    hs = _substrates(nx, ny)[which]
    f = np.zeros((nx, ny))
    f[0, 4] = 2.0  # on the top edge ...
    mask = np.zeros((nx, ny), dtype=bool)
    mask[0, 4] = True  # ... but masked: ignored
    hs.check_boundaries(np.ma.masked_array(f, mask=mask))
    mask[0, 4] = False
    with pytest.raises(type(hs).FreeBoundaryError):
        hs.check_boundaries(np.ma.masked_array(f, mask=mask))
    # an edge masked entirely passes whatever it holds
    f2 = np.zeros((nx, ny))
    f2[:, ny - 1] = 1.0
    m2 = np.zeros((nx, ny), dtype=bool)
    m2[:, ny - 1] = True
    hs.check_boundaries(np.ma.masked_array(f2, mask=m2))


@pytest.mark.parametrize("which", [0, 1])
def test_check_only_when_enabled(which):
    nx, ny = 6, 6
    f = np.ones((nx, ny))
    off = [FreeHalfSpace((nx, ny), 1.0, (1.0, 1.0)),
           FreeFFTElasticHalfSpace((nx, ny), 1.0, (1.0, 1.0))][which]
    off.check(f)  # check_boundaries=False: no check
    on = [FreeHalfSpace((nx, ny), 1.0, (1.0, 1.0), check_boundaries=True),
          FreeFFTElasticHalfSpace((nx, ny), 1.0, (1.0, 1.0), check_boundaries=True)][which]
    with pytest.raises(type(on).FreeBoundaryError):
        on.check(f)
