"""GPU: the plastic mixing step's bisection (``ConstrainedConjugateGradients.py
:316-318``, scipy ``bisect``) on grids whose gaps span several blocks of the
batched bisection rounds, against the oracle over 40 iterations of which 29-30
are mixing steps.

Cases: a free 256² self-affine surface (65536 gaps, 8 blocks per round); the
same problem scaled by 1e5 (heights, load, hardness), whose bracket is about
2^50 xtol wide, so each bisection runs past the batched rounds into the last
block's own passes; heights rounded to 1e-4 (276 distinct values), so many
gaps tie.  Free grids carry ~1e-11 relative FFT round-off in the Love-kernel
spectrum, so the trajectories agree to that round-off as long as no gap sits
within it of a bisection point (the C3 first-load test uses the same bound).
"""

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, Topography, workloads
from contactmechanics_amd import constrained_conjugate_gradients as ccg_gpu
from oracle.ccg import constrained_conjugate_gradients as ccg_cpu
from oracle.halfspace import FreeHalfSpace

pytestmark = pytest.mark.gpu

N = 256


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("case", ["plain", "wide", "ties"])
def test_mixing_steps_follow_oracle(gpu, case):
    h = workloads.self_affine(N, 0.8, 0.1, 5)
    scale = 1e5 if case == "wide" else 1.0
    h = np.round(h, 4) if case == "ties" else h * scale
    kw = dict(external_force=3e-3 * scale, hardness=8e-8 * scale, pentol=0.0, forcetol=0.0, thermotol=0.0,
              maxiter=40)
    trace = []
    rc = ccg_cpu(FreeHalfSpace((N, N), 1.0, (1.0, 1.0)), h, trace=trace, **kw)
    rg = ccg_gpu(FreeFFTElasticHalfSpace((N, N), 1.0, (1.0, 1.0)), Topography(h, (1.0, 1.0)), **kw)
    nmix = sum(t["status"] == "mix" for t in trace)
    print(f"{case}: {nmix} mixing steps, offset gpu {rg.offset:.15e} oracle {rc.offset:.15e}", flush=True)
    assert nmix >= 29  # (29 in the tied case)
    assert rg.nit == rc.nit == 40
    assert rel(rg.jac, rc.jac) < 1e-8
    assert rel(rg.x, rc.x) < 1e-8
    np.testing.assert_allclose(rg.offset, rc.offset, rtol=1e-8)
