"""Oracle Polonsky-Keer constrained conjugate gradients (TEST INFRASTRUCTURE ONLY).

Restatement in NumPy FP64 of ``constrained_conjugate_gradients``
(reference ``ContactMechanics/Optimization/ConstrainedConjugateGradients.py:47-509``)
with the serial ``NuMPI.Tools.Reduction`` semantics (plain ``np.sum``/``np.max``),
every branch in the reference's order:

* set-up: default ``pentol`` (:150-162), masks (:179-200), displacement clamp
  (:201), initial force/displacement (:211-219);
* Polonsky-Keer step (:260-299), plastic mixing step with ``scipy.optimize.bisect``
  on the bearing-area load (:300-330);
* projection and active-set update (:332-389);
* displacement update and convergence tests (:391-434);
* result assembly (:454-475, :494-509).

Pinned by (``tests/test_oracle_ccg.py``): Hertz and Westergaard closed forms
with the reference tests' tolerances, KKT conditions at convergence, an
independent bound-constrained QP (L-BFGS-B) on small grids, the plastic cone
trajectory recorded in the reference notebook
(``examples/plastic_simulation_cone.ipynb`` cell 15) and the bearing-area limit
at vanishing hardness.

``rms_height_from_area`` (SurfaceTopography >= 1.20, absent from the
reference tree) is restated as the area RMS of the valid heights about their
mean, pinned by closed forms in ``tests/test_oracle_pins.py`` (sinusoid
A/sqrt(2), two-level field, constant, masked pixels excluded).
"""

from math import isnan, sqrt

import numpy as np
import scipy.optimize as optim


class Result(dict):
    """Minimal stand-in for ``scipy.optimize.OptimizeResult`` (attribute dict)."""

    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError as err:
            raise AttributeError(name) from err

    def __setattr__(self, name, value):
        self[name] = value


def rms_height_from_area(heights):
    """Area RMS height about the mean of the valid pixels (SurfaceTopography's
    ``rms_height_from_area`` on a uniform grid; closed-form pins in the module
    docstring)."""
    h = np.asarray(np.ma.filled(heights, np.nan), dtype=float)
    h = h[np.isfinite(h)]
    return float(np.sqrt(np.mean((h - h.mean()) ** 2)))


def constrained_conjugate_gradients(substrate, heights, hardness=None,
                                    external_force=None, offset=None,
                                    initial_displacements=None,
                                    initial_forces=None, pentol=None,
                                    forcetol=1e-5, thermotol=1e-6, mixfac=0.1,
                                    mixdecfac=0.95, minmixsteps=10,
                                    maxiter=100000, callback=None,
                                    trace=None):
    """FP64 restatement of the reference CCG; see module docstring.

    ``heights`` is an ``(nx, ny)`` array, optionally a masked array (masked
    pixels are excluded from the computational region, :189-194).
    ``trace``, if a list, receives one dict per iteration with the logged
    scalars (status, area, total force, offset, plastic area, ...).
    """
    heights_arr = heights
    nb_grid = substrate.nb_grid_pts
    nb_surface_pts = int(np.prod(nb_grid))  # :149

    if pentol is None:  # :150-162
        pentol = rms_height_from_area(heights_arr) / (10 * np.mean(nb_grid))
        if pentol == 0:
            pentol = ((offset or 0) + np.sum(heights_arr[...]) / nb_surface_pts) / 1000
        if pentol == 0:
            pentol = 1e-3

    if offset is None:
        offset = 0

    shape = substrate.nb_subdomain_grid_pts
    if initial_displacements is None:
        u = np.zeros(shape)
    else:
        u = np.array(initial_displacements, dtype=float, copy=True)

    # :179-200 computational / surface / padding masks
    comp_slice = tuple(slice(0, min(nb_grid[i], shape[i])) for i in range(2))
    comp = np.zeros(shape, dtype=bool)
    comp[comp_slice] = True
    smask = np.ma.getmask(heights_arr)
    if smask is np.ma.nomask:
        surf_mask = np.ones(nb_grid, dtype=bool)
    else:
        comp[comp_slice][smask] = False
        surf_mask = np.logical_not(smask)
    h = np.asarray(heights_arr[surf_mask])  # 1-D gather over comp
    max_h = np.max(h)
    pad = np.logical_not(comp)
    n_pad = np.sum(pad * 1)

    # :201 clamp u >= h + offset on comp
    uc = u[comp]
    u[comp] = np.where(uc < h + offset, h + offset, uc)

    result = Result(nfev=0, nit=0, success=False, message="Not Converged (yet)")

    if initial_forces is None:  # :211-215
        f = substrate.evaluate_force(u)
    else:
        f = np.array(initial_forces, dtype=float, copy=True)
        u = substrate.evaluate_disp(f)
    result.nfev += 1
    f[pad] = 0.0  # :219

    delta = 0
    delta_str = "reset"
    G_old = 1.0
    t = np.zeros_like(u)
    tau = 0.0
    mix = mixfac
    mixsteps = 0
    last_max_height = None
    last_total_force = None
    A_fl = 0

    for it in range(1, maxiter + 1):  # :235
        result.nit = it
        c = f < 0.0                                   # :239
        A_contact = np.sum(c * 1)                     # :242
        if hardness is not None:                      # :247-249
            c = np.logical_and(c, f > -hardness)
        A_cg = np.sum(c * 1)                          # :253

        if hardness is None or (A_contact > 0 and A_cg / A_contact > 0.01 and mixsteps <= 0):
            # ---- Polonsky-Keer step (:260-299)
            if delta_str == "mix":
                delta_str = "sd"
            g = u[comp] - h
            if external_force is not None:
                offset = 0
                if A_cg > 0:
                    offset = np.sum(g[c[comp]]) / A_cg
            g -= offset
            G = np.sum(c[comp] * g * g)
            if delta > 0 and G_old > 0:
                t[comp] = c[comp] * (g + delta * (G / G_old) * t[comp])
            else:
                t[comp] = c[comp] * g
            r = substrate.evaluate_disp(t)
            result.nfev += 1
            tau = 0.0
            if A_cg > 0:
                x = -np.sum(c * r * t)
                if x > 0.0:
                    tau = np.sum(c[comp] * g * t[comp]) / x
                else:
                    G = 0.0
            f += tau * c * t
            mix = mixfac
        else:
            # ---- plastic mixing step (:300-330)
            if delta_str != "mix":
                mixsteps = minmixsteps
                delta_str = "mix"
            g = u[comp] - h
            if external_force is not None:
                offset = optim.bisect(
                    lambda x: np.sum((g - x) < 0.0) * hardness - external_force,
                    np.max(g), np.min(g))
            g -= offset
            c[comp] = g < 0.0
            f = (1 - mix) * f - mix * hardness * c
            mix *= mixdecfac
            if mixsteps > 0:
                mixsteps -= 1

        # ---- projection (:332-389)
        tensile = f >= 0.0
        nc = np.logical_and(tensile[comp], g < 0.0)
        if hardness is not None:
            flowing = f <= -hardness
            nc = np.logical_or(nc, np.logical_and(flowing[comp], g > 0.0))
        pad_pres = 0
        if n_pad > 0:
            pad_pres = np.max(abs(f[pad]))
        max_pres = 0
        if np.sum(tensile * 1) > 0:
            max_pres = np.max(f[tensile] * 1)
        if hardness:
            A_fl = np.sum(flowing)
            if A_fl > 0:
                max_pres = max(max_pres, -np.min(f[flowing] + hardness))
        f[tensile] = 0.0
        if external_force is not None:
            total_force = -np.sum(f[comp])
            if total_force != 0:
                f *= external_force / total_force
            else:
                f = -external_force / nb_surface_pts * np.ones_like(f)
                f[pad] = 0.0
        if hardness is not None:
            f[flowing] = -hardness
        if delta_str == "mix":
            delta = 0
            G = 0
        else:
            if np.sum(nc * 1) > 0:
                f[comp] += tau * nc * g
                delta = 0
                delta_str = "sd"
            else:
                delta = 1
                delta_str = "cg"

        # ---- displacement update and convergence (:391-434)
        u_new = substrate.evaluate_disp(f)
        maxdu = np.max(abs(u_new - u))
        u = u_new
        result.nfev += 1
        G_old = G
        rms_pen = sqrt(G / A_cg) if A_cg > 0 else sqrt(G)
        max_pen = max(0.0, np.max(c[comp] * (h + offset - u[comp])))
        result.maxcv = {"max_pen": max_pen, "max_pres": max_pres}

        converged = True
        total_force = -np.sum(f[comp])
        with np.errstate(divide="ignore", invalid="ignore"):
            if external_force is not None:
                converged = converged and abs((total_force - external_force) / total_force) < thermotol
            elif last_total_force is not None:
                converged = converged and abs((total_force - last_total_force) / total_force) < thermotol
            last_total_force = total_force
            max_height = max_h + offset
            if external_force is not None and last_max_height is not None:
                converged = converged and abs((max_height - last_max_height) / max_height) < thermotol
            last_max_height = max_height
        if delta_str == "mix":
            converged = converged and maxdu < pentol and max_pres < forcetol and pad_pres < forcetol
        else:
            converged = (converged and rms_pen < pentol and max_pen < pentol and maxdu < pentol
                         and max_pres < forcetol and pad_pres < forcetol)

        if trace is not None:
            trace.append(dict(status="CONVERGED" if converged else delta_str, it=it,
                              area=int(A_contact), cg_area=int(A_cg),
                              total_force=float(total_force), offset=float(offset),
                              plastic_area=int(A_fl) if hardness else 0,
                              rms_pen=float(rms_pen), max_pen=float(max_pen),
                              max_pres=float(max_pres), pad_pres=float(pad_pres),
                              maxdu=float(maxdu), tau=float(tau)))

        if converged:  # :454-475
            result.x = u
            result.jac = -f[comp_slice]
            result.active_set = c
            if hardness is not None:
                plastic = np.zeros_like(c)
                plastic[comp] = g < 0.0
                result.plastic = plastic[comp_slice]
            if delta_str == "mix":
                result.active_set[comp] = g < 0.0
            result.fun = -np.sum(f[comp_slice] * u[comp_slice]) / 2
            result.offset = offset
            result.success = True
            result.message = "Polonsky converged"
            return result

        if callback is not None:
            callback(it, f, dict(area=int(A_contact), rms_penetration=float(rms_pen),
                                 max_penetration=float(max_pen), max_pressure=float(max_pres),
                                 pad_pressure=float(pad_pres), penetration_tol=float(pentol),
                                 pressure_tol=float(forcetol)))
        if isnan(G) or isnan(rms_pen):
            raise RuntimeError("nan encountered.")

    result.x = u  # :494-509
    result.jac = -f[comp_slice]
    result.active_set = c
    result.fun = -np.sum(f[comp_slice] * u[comp_slice]) / 2
    result.offset = offset
    result.message = "Reached maxiter = {}".format(maxiter)
    return result
