"""GPU: a logger alone no longer makes the solver stop after every iteration.

The displacement pass of every iteration stores the scalar record in a ring of
64 records per instance; the host reads the ring after each chunk of
iterations and logs every iteration from it.  The log must be the one of a
solve that stops after every iteration (``iters_per_sync=1``), line for line,
including the verbose columns, mixing steps and a solve that reaches maxiter
(reference ``ConstrainedConjugateGradients.py:436-453``: every iteration that
has not converged is logged, except the one that reaches maxiter).
"""

import io

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from contactmechanics_amd import constrained_conjugate_gradients as ccg
from contactmechanics_amd import workloads
from contactmechanics_amd.Tools.Logger import Logger

pytestmark = pytest.mark.gpu


def logged(sub, h, **kw):
    buf = io.StringIO()
    r = ccg(sub, h, logger=Logger(buf), verbose=True, **kw)
    return r, buf.getvalue()


@pytest.mark.parametrize("case", ["periodic", "plastic", "maxiter"])
def test_ring_log_equals_per_iteration_log(gpu, case):
    n = 128
    h = workloads.self_affine(n, 0.8, 0.1, 11)
    if case == "plastic":
        sub = FreeFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0))
        kw = dict(external_force=2e-3, hardness=0.05 / (n * n), pentol=1e-9, maxiter=3000)
    elif case == "maxiter":  # tolerances it cannot reach: 70 iterations, more than the ring holds
        sub = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0))
        kw = dict(external_force=1e-2, pentol=0.0, forcetol=0.0, thermotol=0.0, maxiter=70)
    else:
        sub = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0))
        kw = dict(external_force=1e-2, pentol=1e-10, forcetol=1e-14, maxiter=5000)
    r1, log1 = logged(sub, h, iters_per_sync=1, **kw)
    r2, log2 = logged(sub, h, **kw)
    assert r1.nit == r2.nit > 8  # several chunks (4, 8, 16, ... iterations)
    if case == "maxiter":
        assert r2.nit == 70  # more iterations than the ring holds: it wrapped
    rows = [ln for ln in log2.splitlines() if ln and not ln.startswith("#")]
    print(f"{case}: nit {r2.nit}, {len(rows)} logged rows, mixing steps {sum(' mix ' in ln for ln in rows)}",
          flush=True)
    assert log1 == log2
    np.testing.assert_array_equal(r1.jac, r2.jac)
