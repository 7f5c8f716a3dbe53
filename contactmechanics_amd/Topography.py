"""Minimal topography containers (stand-ins for the SurfaceTopography types
the reference solver consumes).

The reference reads from SurfaceTopography (absent here): ``heights()``,
``nb_grid_pts``, ``physical_sizes``, ``rms_height_from_area()``
(``ConstrainedConjugateGradients.py:141-160``) and, for plasticity,
``PlasticTopography.hardness`` / ``plastic_displ`` / ``plastic_area``
(``PlasticSystemSpecialisations.py:65-68``).  Only those members are provided.
``rms_height_from_area`` is the area RMS of the valid heights about their mean
(SurfaceTopography's definition for uniform grids); a SurfaceTopography object
passed to the solver supplies its own.
"""

import numpy as np


class Topography:
    """Uniform 2D height map on ``nb_grid_pts`` with ``physical_sizes``
    (default: the grid shape, one length unit per pixel) and a length
    ``unit`` (informational)."""

    def __init__(self, heights, physical_sizes=None, unit=None, periodic=False, **kwargs):
        self._heights = heights if np.ma.isMaskedArray(heights) else np.asarray(heights, dtype=float)
        if self._heights.ndim != 2:
            raise ValueError("only 2D topographies are supported")
        if physical_sizes is None:
            physical_sizes = self._heights.shape
        if not hasattr(physical_sizes, "__iter__"):
            physical_sizes = (physical_sizes, physical_sizes)
        self.physical_sizes = tuple(float(s) for s in physical_sizes)
        self.unit = unit
        self._periodic = bool(periodic)

    def __getitem__(self, item):
        return self.heights()[item]

    def scale(self, factor):
        """The topography with heights multiplied by ``factor``."""
        return Topography(self.heights() * factor, self.physical_sizes, unit=self.unit, periodic=self._periodic)

    def detrend(self, detrend_mode="height"):
        """The topography minus its least-squares polynomial fit in the
        pixel-centre coordinates: ``height`` a constant, ``slope`` a plane,
        ``curvature`` a quadratic surface (SurfaceTopography's detrend modes;
        the fit is restated here)."""
        h = np.ma.asarray(self.heights())
        nx, ny = h.shape
        x, y = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny, indexing="ij")
        terms = {"height": [np.ones_like(x)], "slope": [np.ones_like(x), x, y],
                 "curvature": [np.ones_like(x), x, y, x * x, x * y, y * y]}
        if detrend_mode not in terms:
            raise ValueError(f"detrend_mode must be one of {sorted(terms)}")
        basis = np.stack([t.ravel() for t in terms[detrend_mode]], axis=1)
        valid = ~np.ma.getmaskarray(h).ravel()
        coef, *_ = np.linalg.lstsq(basis[valid], np.ma.filled(h, 0.0).ravel()[valid], rcond=None)
        fit = (basis @ coef).reshape(h.shape)
        out = h - fit if np.ma.isMaskedArray(self.heights()) else np.asarray(h) - fit
        return Topography(out, self.physical_sizes, unit=self.unit, periodic=self._periodic)

    def rms_gradient(self):
        """RMS of the height gradient from central differences (one-sided at
        the edges of a non-periodic map)."""
        h = np.ma.filled(np.ma.asarray(self.heights()), np.nan)
        dx, dy = (s / n for s, n in zip(self.physical_sizes, h.shape))
        if self._periodic:
            gx = (np.roll(h, -1, 0) - np.roll(h, 1, 0)) / (2 * dx)
            gy = (np.roll(h, -1, 1) - np.roll(h, 1, 1)) / (2 * dy)
        else:
            gx, gy = np.gradient(h, dx, dy)
        return float(np.sqrt(np.nanmean(gx ** 2 + gy ** 2)))

    @property
    def nb_grid_pts(self):
        return tuple(self._heights.shape)

    @property
    def nb_subdomain_grid_pts(self):
        return self.nb_grid_pts

    @property
    def subdomain_locations(self):
        return (0, 0)

    @property
    def subdomain_slices(self):
        return tuple(slice(0, n) for n in self.nb_grid_pts)

    @property
    def dim(self):
        return 2

    @property
    def area_per_pt(self):
        return np.prod(self.physical_sizes) / np.prod(self.nb_grid_pts)

    @property
    def is_periodic(self):
        return self._periodic

    def heights(self):
        return self._heights

    def rms_height_from_area(self):
        h = np.ma.compressed(np.ma.asarray(self.heights()))
        return float(np.sqrt(np.mean((h - h.mean()) ** 2)))


class PlasticTopography(Topography):
    """Topography with a hardness and accumulated plastic displacement."""

    def __init__(self, topography, hardness, plastic_displ=None):
        self.parent = topography
        super().__init__(topography.heights(), topography.physical_sizes, unit=getattr(topography, "unit", None),
                         periodic=getattr(topography, "is_periodic", False))
        self.hardness = hardness
        if plastic_displ is None:
            plastic_displ = np.zeros(topography.nb_grid_pts)
        self.plastic_displ = np.asarray(plastic_displ, dtype=float)

    def heights(self):
        return self.parent.heights() + self.plastic_displ

    @property
    def plastic_area(self):
        return np.count_nonzero(self.plastic_displ) * self.area_per_pt
