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
    """Uniform 2D height map on ``nb_grid_pts`` with ``physical_sizes``."""

    def __init__(self, heights, physical_sizes, **kwargs):
        self._heights = heights if np.ma.isMaskedArray(heights) else np.asarray(heights, dtype=float)
        if self._heights.ndim != 2:
            raise ValueError("only 2D topographies are supported")
        if not hasattr(physical_sizes, "__iter__"):
            physical_sizes = (physical_sizes, physical_sizes)
        self.physical_sizes = tuple(float(s) for s in physical_sizes)

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
        return False

    def heights(self):
        return self._heights

    def rms_height_from_area(self):
        h = np.ma.compressed(np.ma.asarray(self.heights()))
        return float(np.sqrt(np.mean((h - h.mean()) ** 2)))


class PlasticTopography(Topography):
    """Topography with a hardness and accumulated plastic displacement."""

    def __init__(self, topography, hardness, plastic_displ=None):
        self.parent = topography
        super().__init__(topography.heights(), topography.physical_sizes)
        self.hardness = hardness
        if plastic_displ is None:
            plastic_displ = np.zeros(topography.nb_grid_pts)
        self.plastic_displ = np.asarray(plastic_displ, dtype=float)

    def heights(self):
        return self.parent.heights() + self.plastic_displ

    @property
    def plastic_area(self):
        return np.count_nonzero(self.plastic_displ) * self.area_per_pt
