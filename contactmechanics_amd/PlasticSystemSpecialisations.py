"""Penetration-hardness contact system.

Mirrors ``ContactMechanics/PlasticSystemSpecialisations.py:37-69``: hardness
is converted from pressure to per-pixel force (:65), passed to the solver
(:66), and the gap of plastic pixels is accumulated into the topography's
plastic displacement on success (:68).
"""

import numpy as np

from .Systems import NonSmoothContactSystem


class PlasticNonSmoothContactSystem(NonSmoothContactSystem):
    def minimize_proxy(self, **kwargs):
        hardness = self.surface.hardness * self.surface.area_per_pt
        opt = super().minimize_proxy(hardness=hardness, **kwargs)
        if opt.success:
            self.surface.plastic_displ += np.where(opt.plastic, self.compute_gap(self.disp, self.offset), 0.)
        return opt
