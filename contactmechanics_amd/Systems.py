"""Contact systems: the drop-in boundary of the solver.

Mirrors ``ContactMechanics/Systems.py``: ``NonSmoothContactSystem.__init__``
(:381-397), ``minimize_proxy`` (:572-611, the injectable ``solver`` is the
plugin point), ``compute_gap`` (:123-134), ``compute_normal_force`` (:421-423),
``compute_contact_area`` / ``compute_nb_contact_pts`` (:141-143, :425-430),
and the objectives other solvers minimise on the same operators (:459-570,
:613-980): ``evaluate``, ``objective`` / ``hessian_product`` (displacements),
``primal_objective`` / ``primal_hessian_product`` (gap),
``evaluate_dual`` / ``dual_objective`` / ``dual_hessian_product`` (pixel
forces), ``primal_minimize_proxy`` / ``dual_minimize_proxy`` and the scipy
helpers ``shape_minimisation_input`` / ``_reshape_bounds`` /
``_lbounds_from_heights`` (:162-220).

The objectives take NumPy arrays (returned as NumPy, like the reference) or
torch tensors on the HIP device (kept there: the convolutions are the device
operators and only the energy crosses to the host).
"""

import numpy as np
import scipy.optimize as optim

from .Optimization import BoundConstrainedCG, constrained_conjugate_gradients


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class IncompatibleFormulationError(Exception):
    pass


class IncompatibleResolutionError(Exception):
    pass


class SystemBase:
    def __init__(self, substrate, surface):
        self.substrate = substrate
        self.area_per_pt = substrate.area_per_pt
        self.surface = surface
        self.dim = None
        self.gap = None
        self.disp = None
        self.comp_slice = substrate.local_topography_subdomain_slices

    def compute_gap(self, disp, offset, *profile_args, **profile_kwargs):
        """Gap ``u - (h + offset)`` on the topography (non-penetration: gap >= 0)."""
        return disp[self.comp_slice] - (self.surface.heights(*profile_args, **profile_kwargs) + offset)

    def compute_contact_area(self):
        return self.compute_nb_contact_pts() * self.area_per_pt

    def compute_relative_contact_area(self):
        return self.compute_contact_area() / np.prod(self.substrate.physical_sizes)

    # ---- scipy helpers (reference :162-220)
    def shape_minimisation_input(self, in_array):
        if np.prod(self.substrate.nb_subdomain_grid_pts) == np.size(in_array):
            return in_array.reshape(-1) if _is_torch(in_array) else np.asarray(in_array).ravel()
        raise IncompatibleResolutionError()

    def shape_minimisation_output(self, in_array):
        if np.prod(self.substrate.nb_subdomain_grid_pts) == np.size(in_array):
            return in_array.reshape(self.substrate.nb_subdomain_grid_pts)
        raise IncompatibleResolutionError()

    @staticmethod
    def _reshape_bounds(lbounds=None, ubounds=None):
        if lbounds is not None and ubounds is not None:
            return optim.Bounds(lb=np.ma.filled(lbounds, -np.inf).ravel(), ub=np.ma.filled(ubounds, np.inf).ravel())
        if lbounds is not None:
            return optim.Bounds(lb=np.ma.filled(lbounds, -np.inf).ravel())
        if ubounds is not None:
            return optim.Bounds(ub=np.ma.filled(ubounds, np.inf).ravel())
        return None

    def _lbounds_from_heights(self, offset):
        """Lower bounds of the displacements for gap >= 0: h + offset on the
        topography, unbounded (masked) on the padding."""
        lbounds = np.ma.masked_all(self.substrate.nb_subdomain_grid_pts)
        lbounds.mask[self.comp_slice] = False
        lbounds[self.comp_slice] = np.asarray(self.surface.heights()) + offset
        lbounds.set_fill_value(-np.inf)
        return lbounds


class NonSmoothContactSystem(SystemBase):
    """Hard-wall (non-adhesive, non-penetrating) contact."""

    def __init__(self, substrate, surface):
        super().__init__(substrate, surface)
        if tuple(surface.nb_grid_pts) != tuple(substrate.nb_grid_pts):
            raise IncompatibleResolutionError(
                "the substrate ({}) and the surface ({}) have incompatible nb_grid_ptss.".format(
                    substrate.nb_grid_pts, surface.nb_grid_pts))
        self.dim = len(self.substrate.nb_grid_pts)
        self.energy = None
        self.force = None
        self.contact_zone = None
        self.offset = None

    @property
    def nb_grid_pts(self):
        return self.surface.nb_grid_pts

    # ---- objectives on the device operators (reference :459-570, :613-980)
    def _heights_like(self, like):
        """The heights as an array of the type and device of ``like``."""
        h = self.surface.heights()
        if _is_torch(like):
            key = (str(like.device), like.dtype)
            if getattr(self, "_h_dev_key", None) != key:
                import torch
                self._h_dev = torch.as_tensor(np.ma.filled(h, 0.0) if np.ma.isMaskedArray(h) else np.asarray(h),
                                              dtype=like.dtype).to(like.device)
                self._h_dev_key = key
            return self._h_dev
        return h

    def evaluate(self, disp, offset, pot=True, forces=False, logger=None):
        """Elastic energy and/or substrate forces of a displacement field
        (reference :459-505); the gap is stored in ``self.gap``."""
        self.gap = self.compute_gap(disp, offset)
        self.substrate.compute(disp, pot, forces)
        self.energy = self.substrate.energy if pot else None
        self.force = self.substrate.force if forces else None
        if logger is not None:
            logger.st(["energy", "max. force"], [self.energy, float(abs(self.force).max()) if forces else 0.0])
        return self.energy, self.force

    def compute_gap(self, disp, offset, *profile_args, **profile_kwargs):
        return disp[self.comp_slice] - (self._heights_like(disp) + offset)

    def objective(self, offset, disp0=None, gradient=False, logger=None):
        """scipy-style objective of the displacements (reference :507-548):
        ``fun(disp) -> (energy, -force)`` with ``gradient``, else the energy."""
        res = self.substrate.nb_subdomain_grid_pts
        if gradient:
            def fun(disp):
                self.evaluate(disp.reshape(res), offset, forces=True, logger=logger)
                return self.energy, -self.force.reshape(-1)
        else:
            def fun(disp):
                return self.evaluate(disp.reshape(res), offset, forces=False, logger=logger)[0]
        return fun

    def hessian_product(self, disp):
        """Hessian product of ``objective`` (reference :550-570)."""
        return self.primal_hessian_product(disp)

    def primal_objective(self, offset, gradient=True):
        """Objective of the gap (reference :613-660): the displacement is
        ``gap + h + offset``; returns the elastic energy and its gradient
        ``-force`` (flattened)."""
        res = self.substrate.nb_subdomain_grid_pts
        if gradient:
            def fun(gap):
                disp = gap.reshape(res) + self._heights_like(gap) + offset
                self.evaluate(disp, offset, forces=True)
                return self.energy, -self.force.reshape(-1)
        else:
            def fun(gap):
                disp = gap.reshape(res) + self._heights_like(gap) + offset
                return self.evaluate(disp, offset, forces=False)[0]
        return fun

    def primal_hessian_product(self, gap):
        """``-evaluate_force(gap)`` in the shape of ``gap`` (reference :662-684)."""
        res = self.substrate.nb_subdomain_grid_pts
        return -self.substrate.evaluate_force(gap.reshape(res)).reshape(gap.shape)

    def evaluate_dual(self, force, offset, logger=None):
        """Energy and gradient of the dual objective for pixel forces on the
        topography (reference :760-840): ``disp = evaluate_disp(-force)``,
        gradient ``disp - h - offset`` (the gap), energy
        ``sum(force disp)/2 - sum(force (h + offset))``."""
        disp = self.substrate.evaluate_disp(-force)
        h = self._heights_like(force)
        self.gradient = disp - h - offset
        if _is_torch(force):
            self.energy = float(0.5 * (force * disp).sum() - (force * (h + offset)).sum())
        else:
            self.energy = 0.5 * np.sum(force * disp) - np.sum(force * (h + offset))
        self.substrate.force = -force
        self.substrate.disp = disp
        if logger is not None:
            n = np.prod(self.nb_grid_pts)
            f = -self.substrate.force
            proj = self.gradient * (force > 0)
            nb_rep, nb_att = float((f > 0).sum()), float((f < 0).sum())
            logger.st(["energy", "force", "frac_rep_area", "frac_att_area", "max_residual", "rms_residual"],
                      [self.energy, float(-self.substrate.force.sum()), nb_rep / n, nb_att / n,
                       float(abs(proj).max()), float(((proj ** 2).sum() / n) ** 0.5)])
        return self.energy, self.gradient

    def dual_objective(self, offset, gradient=True, logger=None):
        """Objective of the pixel forces on the topography (reference
        :842-890): ``fun(force) -> (energy, gap)`` (flattened)."""
        res = self.substrate.topography_nb_subdomain_grid_pts
        if not gradient:
            raise NotImplementedError

        def fun(force):
            self.evaluate_dual(force.reshape(res), offset, logger=logger)
            return self.energy, self.gradient.reshape(-1)
        return fun

    def dual_hessian_product(self, force):
        """``evaluate_disp(-force)`` in the shape of ``force`` (reference :892-912)."""
        res = self.substrate.topography_nb_subdomain_grid_pts
        return self.substrate.evaluate_disp(-force.reshape(res)).reshape(force.shape)

    _SOLVERS = ("ccg-without-restart", "ccg-with-restart", "l-bfgs-b")

    def _bounded_solve(self, solver, fun, hessp, x0, gtol, maxiter):
        if solver not in self._SOLVERS:
            raise ValueError("Input correct solver name from {}".format(set(self._SOLVERS)))
        if solver == "ccg-without-restart":
            return BoundConstrainedCG.constrained_conjugate_gradients(fun, hessp, x0=x0, gtol=gtol, maxiter=maxiter)
        if solver == "ccg-with-restart":
            return BoundConstrainedCG.constrained_conjugate_gradients_with_restart(fun, hessp, x0=x0, gtol=gtol,
                                                                                  maxiter=maxiter)
        x0n = x0.detach().cpu().numpy() if _is_torch(x0) else np.asarray(x0)

        def fun_np(x):
            e, g = fun(x)
            return e, (g.detach().cpu().numpy() if _is_torch(g) else np.asarray(g)).ravel()
        return optim.minimize(fun_np, x0n.ravel(), method="L-BFGS-B", jac=True,
                              bounds=self._reshape_bounds(np.zeros(x0n.size)), options=dict(gtol=gtol, ftol=1e-20))

    def primal_minimize_proxy(self, offset, init_gap=None, solver="ccg-without-restart", gtol=1e-8, maxiter=1000):
        """Minimise the primal objective over gaps >= 0 (reference :686-758).
        On success sets ``offset``, ``gap``, ``force`` (= ``result.jac``),
        ``contact_zone`` (gap == 0) and ``disp``."""
        self.disp = self.force = self.contact_zone = None
        self.init_gap = init_gap
        result = self._bounded_solve(solver, self.primal_objective(offset, gradient=True),
                                     self.primal_hessian_product, init_gap, gtol, maxiter)
        if result.success:
            self.offset = offset
            self.gap = result.x
            self.force = self.substrate.force = result.jac
            self.contact_zone = result.x == 0
            self.disp = self.gap + offset + np.asarray(self.surface.heights()).reshape(np.shape(self.gap)) \
                if not _is_torch(self.gap) else self.gap + offset + self._heights_like(self.gap).reshape(self.gap.shape)
        return result

    def dual_minimize_proxy(self, offset, init_force=None, solver="ccg-without-restart", gtol=1e-8, maxiter=1000,
                            logger=None):
        """Minimise the dual objective over pixel forces >= 0 on the
        topography (reference :914-980).  On success sets ``offset``, ``gap``
        (= ``result.jac``), ``force`` (= ``result.x``), ``contact_zone``
        (force > 0) and ``disp``."""
        self.disp = self.force = self.contact_zone = None
        self.init_force = init_force
        if init_force is not None and tuple(np.shape(init_force)) != tuple(
                self.substrate.topography_nb_subdomain_grid_pts):
            raise IncompatibleResolutionError
        result = self._bounded_solve(solver, self.dual_objective(offset, gradient=True, logger=logger),
                                     self.dual_hessian_product, init_force, gtol, maxiter)
        if result.success:
            self.offset = offset
            self.gap = result.jac
            self.force = self.substrate.force = result.x
            self.contact_zone = result.x > 0
            h = self._heights_like(self.gap) if _is_torch(self.gap) else np.asarray(self.surface.heights())
            self.disp = self.gap + offset + h.reshape(self.gap.shape)
        return result

    def compute_normal_force(self):
        return np.sum(self.substrate.force)

    def compute_nb_contact_pts(self):
        return np.sum(self.contact_zone)

    def compute_contact_coordinates(self):
        return np.argwhere(self.contact_zone)

    def minimize_proxy(self, solver=constrained_conjugate_gradients, **kwargs):
        """Solve the contact problem with ``solver(substrate, surface, **kwargs)``."""
        self.disp = None
        self.force = None
        self.contact_zone = None
        result = solver(self.substrate, self.surface, **kwargs)
        if result.success:
            self.offset = result.offset
            self.disp = result.x
            self.force = self.substrate.force = result.jac
            self.contact_zone = result.jac > 0
            self.substrate.check()
        return result
