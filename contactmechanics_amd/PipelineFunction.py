"""Load sweeps on the GPU solver (drop-in for ``ContactMechanics/PipelineFunction.py``).

``contact_mechanics`` (reference :229-449) runs a sequence of contact solves
on one topography -- at given offsets, at given mean pressures (units of E*),
or at ``nsteps`` automatically chosen offsets -- and returns, per step, the
mean pressure, the fractional contact area, the mean displacement (the
offset), the mean gap and the convergence flag.  ``results_callback``
receives the fields of every step (displacements, gaps, pressures and
contacting pixels of the topography, then offset, mean pressure and area).

The sweep is resident on the device (``DeviceSweep``):

* one substrate plan (Green's function, twiddles, launch shapes) serves every
  step;
* the effective heights (topography plus accumulated plastic displacement)
  live in HBM as FP64 and the plastic update of a converged plastic step
  (``PlasticSystemSpecialisations.py:68``: a plastic pixel takes its gap) is
  applied there;
* the statistics of a step (mean displacement over the topography, force
  sum, number of contacting pixels) are reduced on the device: a step returns
  three scalars to the host, and fields only when a callback asks for them;
* with ``warm_start=True`` a step starts from the device displacement field of
  the previous converged step (the pattern of ``test/test_plastic.py:159-171``);
  by default every step starts from zero, as in the reference.

Automatic offsets (``nsteps``, reference :138-227): the first step sits at
minus the mean height, the second just below the highest point; after that
the steps so far are ordered by contact area and the next offset halves the
widest gap between neighbouring log10(area + 1/N) values, or, when the widest
gap is the one between the largest area and full contact, goes twice as far
below the lowest point as the deepest step so far.  Each automatic step
starts from a fresh plastic displacement, because the steps are not ordered
by load.
"""

import logging

import numpy as np

from .FFTElasticHalfSpace import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from .Optimization import constrained_conjugate_gradients

_log = logging.getLogger(__name__)

MIN_PENTOL = 1e-12  # lower bound of the sweep's penetration tolerance (reference :370-376)


def automatic_offset(step, top, middle, bottom, offsets, areas, nb_pts):
    """Offset of automatic step ``step`` from the heights' extremes and mean
    and the (offset, fractional area) pairs of the steps done so far."""
    if step == 0:
        return -middle
    if step == 1:
        return -top + 0.01 * (top - middle)
    order = np.argsort(np.asarray(areas, dtype=float), kind="stable")
    d = np.asarray(offsets, dtype=float)[order]
    la = np.log10(np.asarray(areas, dtype=float)[order] + 1.0 / nb_pts)
    # widths of the log-area intervals, the last one up to full contact (log 1 = 0)
    widths = np.concatenate([np.diff(la), [0.0 - la[-1]]])
    k = int(np.argmax(widths))
    if k == len(d) - 1:
        return bottom + 2.0 * (d[-1] - bottom)
    return 0.5 * (d[k] + d[k + 1])


class DeviceSweep:
    """State of one load sweep kept on the device between steps."""

    def __init__(self, topography, substrate, hardness=None, pentol=None, maxiter=100,
                 optimizer_kwargs=None, warm_start=False):
        import torch
        h = topography.heights()
        if np.ma.is_masked(h):
            raise NotImplementedError("the device load sweep takes unmasked topographies")
        self.substrate = substrate
        self.nx, self.ny = substrate.nb_grid_pts
        self.dev = substrate._device_str()
        self.h_parent = torch.as_tensor(np.ascontiguousarray(h, dtype=np.float64)).to(self.dev)
        self.h_plastic = torch.zeros_like(self.h_parent)
        self.hardness_force = None if hardness is None else hardness * substrate.area_per_pt
        self.physical_area = float(np.prod(topography.physical_sizes))
        self.nb_pts = self.nx * self.ny
        self.pentol = pentol
        self.maxiter = maxiter
        self.optimizer_kwargs = dict(optimizer_kwargs or {})
        self.warm_start = warm_start
        self.u_last = None
        # per-step records on the host
        self.offsets, self.gaps, self.pressures, self.areas, self.converged = [], [], [], [], []

    def heights(self):
        """Effective heights on the device (topography + plastic displacement)."""
        return self.h_parent + self.h_plastic

    def height_extremes(self):
        h = self.heights()
        return float(h.max()), float(h.mean()), float(h.min())

    def reset_plastic(self):
        self.h_plastic.zero_()

    def solve(self, offset=None, external_force=None, want_fields=False):
        """One contact solve; returns the step's scalars and, with
        ``want_fields``, its fields as NumPy arrays."""
        import torch
        nx, ny = self.nx, self.ny
        h_eff = self.heights()
        mean_h = h_eff.mean()
        kw = dict(self.optimizer_kwargs)
        if self.warm_start and self.u_last is not None:
            kw["initial_displacements"] = self.u_last
        res = constrained_conjugate_gradients(
            self.substrate, h_eff, hardness=self.hardness_force, offset=offset,
            external_force=external_force, pentol=self.pentol, maxiter=self.maxiter,
            return_device=True, **kw)
        u_top = res.x[:nx, :ny].to(torch.float64)
        jac = res.jac.to(torch.float64)
        # statistics on the device; three scalars come back
        stats = torch.stack([u_top.mean(), jac.sum(), (jac > 0).sum().to(torch.float64)]).cpu().numpy()
        mean_u, force_sum, n_contact = (float(v) for v in stats)
        off = float(res.offset)
        if res.success and self.hardness_force is not None:
            # plastic pixels take their gap into the plastic displacement
            gap = u_top - (h_eff + off)
            self.h_plastic += torch.where(res.plastic, gap, torch.zeros_like(gap))
        self.u_last = res.x if res.success else None
        rec = dict(offset=off, mean_gap=mean_u - float(mean_h) - off,
                   mean_pressure=force_sum / self.physical_area, area=n_contact / self.nb_pts,
                   converged=bool(res.success))
        self.offsets.append(rec["offset"])
        self.gaps.append(rec["mean_gap"])
        self.pressures.append(rec["mean_pressure"])
        self.areas.append(rec["area"])
        self.converged.append(rec["converged"])
        fields = None
        if want_fields:
            gap = torch.clamp(u_top - self.heights() - off, min=0.0)
            fields = (u_top.cpu().numpy(), gap.cpu().numpy(),
                      (jac / self.substrate.area_per_pt).cpu().numpy(), (jac > 0).cpu().numpy())
        return rec, fields


def contact_mechanics(topography, substrate=None, nsteps=None, offsets=None, pressures=None, hardness=None,
                      maxiter=100, results_callback=None, optimizer_kwargs={}, warm_start=False,
                      precision="fp64", device=None):
    """Sweep offsets, pressures or ``nsteps`` automatic steps (see the module
    docstring).  Returns (mean_pressure, total_contact_area,
    mean_displacement, mean_gap, converged) as arrays.

    Extra keywords: ``warm_start`` (start each step from the previous
    converged displacements), ``precision`` and ``device`` of the substrate.
    """
    periodic = bool(getattr(topography, "is_periodic", False))
    if substrate is None:
        substrate = "periodic" if periodic else "nonperiodic"
    if periodic != (substrate == "periodic"):
        _log.warning("Topography is %s, but the analysis is configured for %s boundaries.",
                     "periodic" if periodic else "not periodic",
                     "free" if substrate != "periodic" else "periodic")
    modes = {"nsteps": nsteps, "offsets": offsets, "pressures": pressures}
    given = [k for k, v in modes.items() if v is not None]
    if not given:
        raise ValueError("Either `nsteps`, `offsets` or `pressures` must be given for a contact "
                         "mechanics calculation.")
    if len(given) > 1:
        raise ValueError("Only one of `nsteps`, `offsets` and `pressures` may be given (got {}).".format(
            ", ".join(f"`{k}`" for k in given)))

    cls = {"periodic": PeriodicFFTElasticHalfSpace, "nonperiodic": FreeFFTElasticHalfSpace}[substrate]
    sub = cls(topography.nb_grid_pts, 1.0, topography.physical_sizes, precision=precision, device=device)
    plastic = hardness is not None and hardness > 0
    if hasattr(topography, "rms_height_from_area"):
        rms = float(topography.rms_height_from_area())
    else:
        h = np.asarray(topography.heights(), dtype=float)
        rms = float(np.sqrt(np.mean((h - h.mean()) ** 2)))
    pentol = max(rms / (10 * np.mean(topography.nb_grid_pts)), MIN_PENTOL)
    sweep = DeviceSweep(topography, sub, hardness=hardness if plastic else None, pentol=pentol,
                        maxiter=maxiter, optimizer_kwargs=optimizer_kwargs, warm_start=warm_start)

    n_steps = len(offsets) if offsets is not None else (len(pressures) if pressures is not None else nsteps)
    force_unit = float(np.prod(topography.physical_sizes))  # pressure in E* -> total force
    for step in range(n_steps):
        if offsets is not None:
            kw = dict(offset=offsets[step])
        elif pressures is not None:
            kw = dict(external_force=pressures[step] * force_unit)
        else:
            if plastic:
                sweep.reset_plastic()
            top, mid, bot = sweep.height_extremes()
            kw = dict(offset=automatic_offset(step, top, mid, bot, sweep.offsets, sweep.areas, sweep.nb_pts))
        rec, fields = sweep.solve(want_fields=results_callback is not None, **kw)
        if results_callback is not None:
            results_callback(*fields, rec["offset"], rec["mean_pressure"], rec["area"])
    return (np.array(sweep.pressures), np.array(sweep.areas), np.array(sweep.offsets), np.array(sweep.gaps),
            np.array(sweep.converged, dtype=bool))
