"""Bound-constrained conjugate gradients for the primal and dual contact
objectives (``Systems.py:613-980`` drives them through NuMPI >= 0.12,
``NuMPI.Optimization.CCGWithoutRestart`` / ``CCGWithRestart``, which is not
part of the reference tree; the algorithms are restated here from their
published descriptions).

Both minimise a convex quadratic ``f(x)`` given as ``fun(x) -> (f, grad)``
and its Hessian product ``hessp(d)``, subject to ``x >= lower`` (``bounds``,
default 0) and optionally ``mean(x) = mean_val``:

``constrained_conjugate_gradients`` (without restart; Bugnicourt, Sainsot,
Dureisseix, Gauthier and Lubrecht, Tribol. Lett. 66:29, 2018)
    Conjugate directions over the free set F = {x > lower or grad < 0}; the
    step is the exact minimiser of the quadratic along the direction, then
    the iterate is projected onto the bounds.  The Polak-Ribiere coefficient
    is kept through changes of the free set (no restart), clipped at 0.

``constrained_conjugate_gradients_with_restart`` (Polonsky and Keer, Wear
231:206, 1999)
    The same iteration, but the conjugacy restarts (steepest descent) in
    every iteration where the projection moved a point onto its bound or a
    bound point was released.

With ``mean_val``, directions have zero mean over F and the iterate is
rescaled onto the mean after every projection (the load constraint of the
dual, as in the Polonsky-Keer force rescaling).

Convergence: the projected gradient (grad on F, zero where x sits on its
bound with grad >= 0) has max norm <= ``gtol``; with ``mean_val`` the
gradient's mean over F (the Lagrange multiplier of the mean) is removed first.

The iteration runs in the array type of ``x0``: NumPy arrays on the host, or
torch tensors on the HIP device, in which case ``fun`` and ``hessp`` (e.g. a
system's objectives, whose products are the device convolutions) keep every
vector on the device and only scalars cross to the host.
"""

import numpy as np
import scipy.optimize as optim


class _Xp:
    """The few array operations the iteration needs, for NumPy or torch."""

    def __init__(self, like):
        self.torch = type(like).__module__.startswith("torch")
        if self.torch:
            import torch
            self.t = torch

    def asarray(self, a, like):
        if self.torch:
            if not type(a).__module__.startswith("torch"):
                a = self.t.as_tensor(np.asarray(a), dtype=like.dtype)
            return a.to(device=like.device, dtype=like.dtype).reshape(-1)
        return np.asarray(a, dtype=float).reshape(-1)

    def where(self, c, a, b):
        return self.t.where(c, a, b) if self.torch else np.where(c, a, b)

    def maximum(self, a, b):
        return self.t.maximum(a, b) if self.torch else np.maximum(a, b)

    def dot(self, a, b):
        return float((a * b).sum())

    def amax_abs(self, a):
        return float(a.abs().max()) if self.torch else float(np.max(np.abs(a))) if a.size else 0.0

    def count(self, m):
        return float(m.sum())

    def zeros_like(self, a):
        return self.t.zeros_like(a) if self.torch else np.zeros_like(a)


def _run(fun, hessp, x0, gtol, mean_val, maxiter, bounds, callback, restart):
    xp = _Xp(x0)
    shape = tuple(x0.shape)
    x = xp.asarray(x0, x0).clone() if xp.torch else np.array(x0, dtype=float).reshape(-1)
    if bounds is None:
        lower = xp.zeros_like(x)
    else:
        lower = xp.asarray(bounds, x)
        if xp.torch:
            lower = lower.clone()
        # unbounded points (-inf or masked) stay free
    x = xp.maximum(x, lower)
    if mean_val is not None:
        x = x * (mean_val / float(x.mean()))

    def evaluate(x):
        f, g = fun(x.reshape(shape))
        return f, xp.asarray(g, x)

    def free_set(x, g):
        return (x > lower) | (g < 0)

    def projected(g, F):
        pg = xp.where(F, g, xp.zeros_like(g))
        if mean_val is not None:
            nF = xp.count(F)
            if nF > 0:
                pg = xp.where(F, pg - float(pg.sum()) / nF, pg)
        return pg

    f, g = evaluate(x)
    F = free_set(x, g)
    pg = projected(g, F)
    d = -pg
    nit = 0
    res = None
    while True:
        if xp.amax_abs(pg) <= gtol:
            res = dict(success=True, message="CONVERGENCE: NORM_OF_GRADIENT_<=_GTOL")
            break
        if nit >= maxiter:
            res = dict(success=False, message="NO CONVERGENCE: MAXITERATIONS REACHED")
            break
        Hd = xp.asarray(hessp(d.reshape(shape)), x)
        dHd = xp.dot(d, Hd)
        if not dHd > 0:
            res = dict(success=False, message="NO CONVERGENCE: direction of non-positive curvature")
            break
        alpha = -xp.dot(g, d) / dHd
        x_new = xp.maximum(x + alpha * d, lower)
        if mean_val is not None:
            x_new = x_new * (mean_val / float(x_new.mean()))
        f, g_new = evaluate(x_new)
        F_new = free_set(x_new, g_new)
        pg_new = projected(g_new, F_new)
        changed = bool((F_new != F).any())
        if restart and changed:
            beta = 0.0
        else:
            pgpg = xp.dot(pg, pg)
            beta = max(0.0, xp.dot(pg_new, pg_new - pg) / pgpg) if pgpg > 0 else 0.0
        d = xp.where(F_new, -pg_new + beta * d, xp.zeros_like(d))
        if mean_val is not None:
            nF = xp.count(F_new)
            if nF > 0:
                d = xp.where(F_new, d - float(d.sum()) / nF, d)
        x, g, F, pg = x_new, g_new, F_new, pg_new
        nit += 1
        if callback is not None:
            callback(x.reshape(shape))
    result = optim.OptimizeResult(x=x.reshape(shape), jac=g.reshape(shape), fun=f, nit=nit, **res)
    if not xp.torch:
        result.x = np.asarray(result.x).reshape(shape)
        result.jac = np.asarray(result.jac).reshape(shape)
    return result


def constrained_conjugate_gradients(fun, hessp, x0, gtol=1e-8, mean_val=None, residual_plot=False, maxiter=3000,
                                    bounds=None, communicator=None, callback=None):
    """Bound-constrained CG without restart (see module docstring).
    ``fun(x) -> (energy, gradient)``, ``hessp(d) -> H d``; ``bounds``: lower
    bounds (default 0; masked or -inf entries unbounded).  Returns an
    ``OptimizeResult`` with ``x``, ``jac`` (gradient), ``fun``, ``nit``,
    ``success``, ``message``."""
    if communicator is not None and getattr(communicator, "size", 1) > 1:
        raise ValueError("one process holds the whole grid")
    if bounds is not None and np.ma.isMaskedArray(bounds):
        bounds = np.ma.filled(bounds, -np.inf)
    return _run(fun, hessp, x0, gtol, mean_val, maxiter, bounds, callback, restart=False)


def constrained_conjugate_gradients_with_restart(fun, hessp, x0, gtol=1e-8, mean_val=None, residual_plot=False,
                                                 maxiter=3000, bounds=None, communicator=None, callback=None):
    """Bound-constrained CG restarted whenever the free set changes
    (Polonsky-Keer; see module docstring)."""
    if communicator is not None and getattr(communicator, "size", 1) > 1:
        raise ValueError("one process holds the whole grid")
    if bounds is not None and np.ma.isMaskedArray(bounds):
        bounds = np.ma.filled(bounds, -np.inf)
    return _run(fun, hessp, x0, gtol, mean_val, maxiter, bounds, callback, restart=True)
