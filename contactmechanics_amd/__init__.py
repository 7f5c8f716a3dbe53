"""MI355X (gfx950) contact solver with the ContactMechanics hot-path API.

Polonsky-Keer constrained conjugate gradients on periodic, free (2N
zero-padded) and semi-periodic FFT elastic half-spaces, with the plastic
hardness cap, running
in hand-written HIP kernels behind the C ABI ``include/cmhip.h``.
"""

from .FFTElasticHalfSpace import (FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace,  # noqa: F401
                                  SemiPeriodicFFTElasticHalfSpace)
from .Factory import make_plastic_system, make_system  # noqa: F401
from .Optimization import constrained_conjugate_gradients  # noqa: F401
from .PlasticSystemSpecialisations import PlasticNonSmoothContactSystem  # noqa: F401
from .Systems import NonSmoothContactSystem  # noqa: F401
from .Topography import PlasticTopography, Topography  # noqa: F401

__version__ = "0.1.0"
