from .ConstrainedConjugateGradients import constrained_conjugate_gradients  # noqa: F401
