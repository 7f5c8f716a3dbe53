"""CPU FP64 oracle for the constrained-conjugate-gradient contact solve.

TEST INFRASTRUCTURE ONLY.  This package is the checker: a plain NumPy/SciPy
restatement of the reference algorithm (ContactMechanics v1.8.3), used by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``.  The product path (``contactmechanics_amd``) never imports it.

Modules
-------
halfspace
    Periodic and free (2N zero-padded) FFT elastic half-space operators:
    Green's functions, ``evaluate_disp`` and ``evaluate_force``
    (reference ``FFTElasticHalfSpace.py``).
ccg
    Polonsky-Keer constrained conjugate gradients with load/offset control and
    the plastic hardness cap (reference ``ConstrainedConjugateGradients.py``).
reference_solutions
    Closed forms used to pin the oracle (Hertz, Westergaard, Love's
    rectangle-patch displacement).

Parity status of each function (what pins it) is listed in the module
docstrings.  Functions whose behaviour lives in third-party packages absent
from the reference tree (SurfaceTopography's ``rms_height_from_area``) are
restated from their published definition and pinned by closed forms.
"""
