"""System constructors (``ContactMechanics/Factory.py:115-156``)."""

from .FFTElasticHalfSpace import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from .PlasticSystemSpecialisations import PlasticNonSmoothContactSystem
from .Systems import NonSmoothContactSystem
from .Topography import PlasticTopography


def _substrate(substrate, surface, young, precision, **kwargs):
    if isinstance(substrate, str):
        cls = {"periodic": PeriodicFFTElasticHalfSpace, "free": FreeFFTElasticHalfSpace}[substrate]
        return cls(surface.nb_grid_pts, young, surface.physical_sizes, precision=precision, **kwargs)
    return substrate


def make_system(substrate, surface, young=None, precision="fp64", **kwargs):
    sub = _substrate(substrate, surface, young, precision, **kwargs)
    if isinstance(surface, PlasticTopography):
        return PlasticNonSmoothContactSystem(sub, surface)
    return NonSmoothContactSystem(sub, surface)


def make_plastic_system(substrate, surface, young=None, precision="fp64", **kwargs):
    sub = _substrate(substrate, surface, young, precision, **kwargs)
    return PlasticNonSmoothContactSystem(sub, surface)
