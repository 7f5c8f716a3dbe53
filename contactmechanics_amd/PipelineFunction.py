"""Load-sweep driver on the GPU solver.

Mirrors ``ContactMechanics/PipelineFunction.py``:

* ``_contact_calculation`` (:37-135): one solve at an offset or external force,
  appending mean displacement, mean gap, mean pressure, fractional contact
  area and the convergence flag to the history;
* ``_next_contact_step`` (:138-227): automatic offsets spreading the contact
  areas evenly on a log scale (plastic displacement reset before each step);
* ``contact_mechanics`` (:229-449): offsets, pressures (units of E*) or
  ``nsteps`` automatic steps, ``pentol = max(rms_height/(10 mean(n)), 1e-12)``,
  contact modulus 1, hardness through a plastic system.

Extra keyword ``warm_start`` (default False, as in the reference): start each
step from the previous step's displacements, the pattern of reference
``test/test_plastic.py:159-171`` used by the plastic load sweep benchmark.
"""

import logging

import numpy as np

from .FFTElasticHalfSpace import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from .Factory import make_plastic_system, make_system
from .Topography import PlasticTopography

_log = logging.getLogger(__name__)


def _contact_calculation(system, offset=None, external_force=None, history=None, pentol=None,
                         maxiter=None, optimizer_kwargs={}, initial_displacements=None):
    topography = system.surface
    substrate = system.substrate
    middle = np.mean(topography.heights())
    if history is None:
        mean_displacements, mean_gaps, mean_pressures, total_contact_areas = [], [], [], []
        converged = np.array([], dtype=bool)
    else:
        mean_displacements, mean_gaps, mean_pressures, total_contact_areas, converged = history
    kw = dict(optimizer_kwargs)
    if initial_displacements is not None:
        kw["initial_displacements"] = initial_displacements
    opt = system.minimize_proxy(offset=offset, external_force=external_force, pentol=pentol, maxiter=maxiter,
                                **kw)
    force_xy = opt.jac
    displacement_xy = opt.x[: force_xy.shape[0], : force_xy.shape[1]]
    mean_displacements = np.append(mean_displacements, [opt.offset])
    mean_gaps = np.append(mean_gaps, [np.mean(displacement_xy) - middle - opt.offset])
    mean_load = force_xy.sum() / np.prod(topography.physical_sizes)
    mean_pressures = np.append(mean_pressures, [mean_load])
    total_contact_area = (force_xy > 0).sum() / np.prod(topography.nb_grid_pts)
    total_contact_areas = np.append(total_contact_areas, [total_contact_area])
    converged = np.append(converged, np.array([opt.success], dtype=bool))
    pressure_xy = force_xy / substrate.area_per_pt
    gap_xy = displacement_xy - topography.heights() - opt.offset
    gap_xy[gap_xy < 0.0] = 0.0
    return (displacement_xy, gap_xy, pressure_xy, force_xy > 0, opt.offset, mean_load, total_contact_area,
            (mean_displacements, mean_gaps, mean_pressures, total_contact_areas, converged), opt)


def _next_offset(topography, history):
    heights = topography.heights()
    top, middle, bot = np.max(heights), np.mean(heights), np.min(heights)
    step = 0 if history is None else len(history[0])
    if step == 0:
        return -middle
    if step == 1:
        return -top + 0.01 * (top - middle)
    mean_displacements, _, _, total_contact_areas, _ = history
    order = sorted(zip(mean_displacements, total_contact_areas), key=lambda x: x[1])
    sorted_disp = np.array([o[0] for o in order])
    sorted_area = np.array([o[1] for o in order])
    ref_area = np.log10(sorted_area + 1 / np.prod(topography.nb_grid_pts))
    darea = np.append(ref_area[1:] - ref_area[:-1], -ref_area[-1])
    i = int(np.argmax(darea))
    if i == step - 1:
        return bot + 2 * (sorted_disp[-1] - bot)
    return (sorted_disp[i] + sorted_disp[i + 1]) / 2


def contact_mechanics(topography, substrate=None, nsteps=None, offsets=None, pressures=None, hardness=None,
                      maxiter=100, results_callback=None, optimizer_kwargs={}, warm_start=False,
                      precision="fp64", device=None):
    """Sweep offsets, pressures or ``nsteps`` automatic steps; see module docstring.

    Returns (mean_pressure, total_contact_area, mean_displacement, mean_gap, converged).
    """
    periodic = bool(getattr(topography, "is_periodic", False))
    if substrate is None:
        substrate = "periodic" if periodic else "nonperiodic"
    if periodic != (substrate == "periodic"):
        _log.warning("Topography periodicity does not match the requested substrate (%s)", substrate)
    given = [x is not None for x in (nsteps, offsets, pressures)]
    if not any(given):
        raise ValueError("Either `nsteps`, `offsets` or `pressures` must be given for a contact "
                         "mechanics calculation.")
    if sum(given) > 1:
        raise ValueError("Only one of `nsteps`, `offsets` and `pressures` may be given.")
    force_conv = np.prod(topography.physical_sizes)
    plastic = hardness is not None and hardness > 0
    surface = PlasticTopography(topography, hardness) if plastic else topography
    cls = dict(periodic=PeriodicFFTElasticHalfSpace, nonperiodic=FreeFFTElasticHalfSpace)[substrate]
    sub = cls(surface.nb_grid_pts, 1.0, surface.physical_sizes, precision=precision, device=device)
    system = (make_plastic_system if plastic else make_system)(substrate=sub, surface=surface)
    h = np.ma.compressed(np.ma.asarray(surface.heights()))
    pentol = max(float(np.sqrt(np.mean((h - h.mean()) ** 2))) / (10 * np.mean(surface.nb_grid_pts)), 1e-12)
    if pressures is not None:
        nsteps = len(pressures)
    if offsets is not None:
        nsteps = len(offsets)
    history = None
    disp0 = None
    for i in range(nsteps):
        kw = dict(history=history, pentol=pentol, maxiter=maxiter, optimizer_kwargs=optimizer_kwargs,
                  initial_displacements=disp0 if warm_start else None)
        if offsets is not None:
            out = _contact_calculation(system, offset=offsets[i], **kw)
        elif pressures is not None:
            out = _contact_calculation(system, external_force=pressures[i] * force_conv, **kw)
        else:
            if plastic:  # every automatic step computes a fresh plastic displacement
                surface.plastic_displ = np.zeros_like(surface.plastic_displ)
            out = _contact_calculation(system, offset=_next_offset(surface, history), **kw)
        history = out[7]
        opt = out[8]
        disp0 = opt.x if opt.success else None
        if results_callback is not None:
            results_callback(*out[:7])
    mean_displacement, mean_gap, mean_pressure, total_contact_area, converged = history
    return (np.array(mean_pressure), np.array(total_contact_area), np.array(mean_displacement),
            np.array(mean_gap), np.array(converged))
