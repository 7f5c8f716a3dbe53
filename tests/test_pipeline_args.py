"""Argument checks of the load-sweep driver (CPU: raised before any GPU use),
reference ``PipelineFunction.py:301-330``."""

import numpy as np
import pytest

from contactmechanics_amd import Topography
from contactmechanics_amd.PipelineFunction import _next_offset, contact_mechanics


def test_exactly_one_sweep_mode():
    t = Topography(np.zeros((8, 8)), (1.0, 1.0))
    with pytest.raises(ValueError):
        contact_mechanics(t, substrate="periodic")
    with pytest.raises(ValueError):
        contact_mechanics(t, substrate="periodic", nsteps=3, pressures=[1e-3])
    with pytest.raises(ValueError):
        contact_mechanics(t, substrate="periodic", offsets=[0.1], pressures=[1e-3])


def test_automatic_offsets_follow_the_reference_rule():
    rng = np.random.default_rng(0)
    t = Topography(rng.standard_normal((16, 16)), (1.0, 1.0))
    h = t.heights()
    assert _next_offset(t, None) == -np.mean(h)
    hist = (np.array([-np.mean(h)]), None, None, np.array([0.3]), None)
    assert _next_offset(t, hist) == -np.max(h) + 0.01 * (np.max(h) - np.mean(h))
    # step >= 2: bisect the widest log-area gap, or extrapolate below the deepest offset
    hist = (np.array([0.0, -1.0]), None, None, np.array([0.5, 0.01]), None)
    off = _next_offset(t, hist)
    assert np.isfinite(off)
