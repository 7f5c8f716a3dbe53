from . import BoundConstrainedCG  # noqa: F401
from .ConstrainedConjugateGradients import constrained_conjugate_gradients  # noqa: F401

# NuMPI-style module names for the bound-constrained solvers the reference's
# Systems.py drives (NuMPI.Optimization.CCGWithoutRestart / CCGWithRestart)
CCGWithoutRestart = BoundConstrainedCG
CCGWithRestart = BoundConstrainedCG
