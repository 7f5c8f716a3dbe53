"""Contact systems: the drop-in boundary of the solver.

Mirrors ``ContactMechanics/Systems.py``: ``NonSmoothContactSystem.__init__``
(:381-397), ``minimize_proxy`` (:572-611, the injectable ``solver`` is the
plugin point), ``compute_gap`` (:123-134), ``compute_normal_force`` (:421-423),
``compute_contact_area`` / ``compute_nb_contact_pts`` (:141-143, :425-430).
"""

import numpy as np

from .Optimization import constrained_conjugate_gradients


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
