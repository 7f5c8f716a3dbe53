"""Launch-shape autotuner with an on-disk cache.

``tune(substrate)`` times the FFT passes of the substrate's plan over the
candidate shapes (``cm_plan_autotune``) the first time a (device, grid,
boundary, precision) combination is seen and stores the winner in a JSON file;
later calls apply the cached shape (``cm_plan_set_shape``) without timing.
"""

import json
import os

from ._native import CmError

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "contactmechanics_amd", "autotune.json")


def _key(substrate):
    import torch
    dev = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(dev, "gcnArchName", dev.name)
    mx, my = substrate.nb_domain_grid_pts
    prec = "fp32" if substrate.precision == "fp32" else "fp64"
    return f"{arch}|{mx}x{my}|{'periodic' if substrate.is_periodic() else 'free'}|{prec}"


def tune(substrate, cache_path=DEFAULT_CACHE, reps=10, force=False):
    """Set the fastest launch shape on ``substrate.plan``; returns its description.
    FP32 plans keep their automatic shape (the library offers explicit shapes
    for FP64 plans only)."""
    plan = substrate.plan
    if substrate.precision != "fp64":
        return dict(plan.shape(), cached=False, tuned=False)
    key = _key(substrate)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if not force and key in cache:
        c = cache[key]
        try:
            plan.set_shape(c["Cb"], c["R"], c["erow"], c["ecol"])
            return dict(c, cached=True)
        except CmError:
            pass  # a shape this build no longer offers (32-element transforms): tune again
    desc = plan.autotune(reps)
    cache[key] = desc
    if cache_path:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(cache, fh, indent=1)
        os.replace(tmp, cache_path)
    return dict(desc, cached=False)
