"""Closed-form contact solutions used to pin the oracle (TEST INFRASTRUCTURE ONLY).

Hertz (sphere on half-space): ``ReferenceSolutions/Hertz.py:35-130``; pinned
by the reference's recorded values
(``test/ReferenceSolutions/test_reference_solutions.py:83-96``).

Westergaard (periodic 1D sinusoid): ``ReferenceSolutions/Westergaard.py:52-97``.
"""

import math

import numpy as np


def hertz_penetration(N, R, Es):
    """Rigid-body penetration for load N (Johnson eq. 4.23)."""
    return (9 * N ** 2 / (16 * R * Es ** 2)) ** (1 / 3)


def hertz_normal_load(d, R, Es):
    return 4 / 3 * Es * np.sqrt(d ** 3 * R)


def hertz_elastic_energy(d, R, Es):
    return 8 / 15 * np.sqrt(R) * Es * d ** (5 / 2)


def hertz_radius_and_pressure(N, R, Es):
    """Contact radius a and peak pressure p0 for load N."""
    a = R * (3.0 / 4 * (N / (Es * R ** 2))) ** (1.0 / 3)
    p0 = 3 * N / (2 * math.pi * a * a)
    return a, p0


def hertz_pressure(r, a, p0):
    """Hertz pressure p0 sqrt(1-(r/a)^2) inside the contact, 0 outside."""
    r = np.asarray(r, dtype=float)
    p = np.zeros_like(r)
    m = r < a
    p[m] = p0 * np.sqrt(1 - (r[m] / a) ** 2)
    return p


def westergaard_pressure(x, mean_pressure):
    """Westergaard pressure profile for a unit-amplitude sinusoid.

    ``x`` is position in units of the period; ``mean_pressure`` is
    non-dimensionalised by ``pi E* h0 / lambda`` (``Westergaard.py:56-97``).
    """
    x = np.abs(np.mod(np.asarray(x, dtype=float) + 0.5, 1.0) - 0.5)
    psi = np.pi * x
    psia = np.arcsin(np.sqrt(mean_pressure / np.pi))
    p = np.zeros_like(x)
    m = np.cos(psi) > np.cos(psia)
    if m.sum() > 0:
        p[m] = 2 * np.cos(psi[m]) / np.sin(psia) ** 2 * np.sqrt(
            np.sin(psia) ** 2 - np.sin(psi[m]) ** 2)
    return p
