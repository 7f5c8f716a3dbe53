"""Batched solves of many independent contact problems of uneven size.

Instances sharing a computational grid and boundary condition form a group;
a group is one resident ``CcgSession`` whose kernels cover all its instances
in a single launch (grid dimension y = instance).  Groups run concurrently on
separate HIP streams, so small grids fill the device together.  Every
instance keeps its own scalar state and reduction tree, sized by its own grid
only, so its result is bit-for-bit the same alone or in any batch.
"""

import time
from collections import OrderedDict

import numpy as np

from .FFTElasticHalfSpace import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from .session import CcgSession, make_params


class PlanCache:
    """One substrate (plan) per (n, bc, precision, physical size, modulus)."""

    def __init__(self, precision="fp64", device=None):
        self.precision = precision
        self.device = device
        self._subs = OrderedDict()

    def get(self, n, bc, size=1.0, young=1.0):
        key = (int(n), bc, float(size), float(young))
        sub = self._subs.get(key)
        if sub is None:
            cls = PeriodicFFTElasticHalfSpace if bc == "periodic" else FreeFFTElasticHalfSpace
            sub = cls((n, n), young, (size, size), precision=self.precision, device=self.device)
            sub.plan  # build now
            self._subs[key] = sub
        return sub


def solve_batch(problems, precision="fp64", chunk=8, max_chunk=64, streams=4, cache=None):
    """Solve ``problems`` (dicts with n, bc, heights, and the solver keywords
    external_force / offset / hardness / pentol / forcetol / thermotol /
    maxiter).  Returns (results, info): per problem a dict with the stats and
    the device force field ``jac``; info has wall time and grid-pt-iters."""
    import torch
    cache = cache or PlanCache(precision)
    groups = OrderedDict()
    for i, p in enumerate(problems):
        groups.setdefault((p["n"], p["bc"]), []).append(i)
    pool = [torch.cuda.Stream() for _ in range(max(1, streams))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sessions = []
    for gi, ((n, bc), idx) in enumerate(groups.items()):
        sub = cache.get(n, bc)
        params = []
        for i in idx:
            p = problems[i]
            params.append(make_params(hardness=p.get("hardness"), external_force=p.get("external_force"),
                                      offset=p.get("offset"), pentol=p["pentol"], forcetol=p["forcetol"],
                                      thermotol=p.get("thermotol", 1e-6), maxiter=p.get("maxiter", 100000)))
        s = CcgSession(sub, [problems[i]["heights"] for i in idx], params, stream=pool[gi % len(pool)])
        sessions.append([s, idx, chunk, False])
    active = len(sessions)
    while active:
        for rec in sessions:  # enqueue a chunk on every unfinished group (no host sync)
            if not rec[3]:
                rec[0].step(rec[2])
        for rec in sessions:
            if rec[3]:
                continue
            done, _ = rec[0].poll()
            if done:
                rec[3] = True
                active -= 1
            else:
                rec[2] = min(rec[2] * 2, max_chunk)
    results = [None] * len(problems)
    gpi = 0
    for s, idx, _, _ in sessions:
        stats = s.finish()
        M = int(np.prod(s.substrate.nb_domain_grid_pts))
        for b, i in enumerate(idx):
            results[i] = dict(stats=stats[b], jac=s.pressures(b).clone(), grid=M)
            gpi += M * stats[b]["nit"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for s, _, _, _ in sessions:
        s.close()
    return results, dict(seconds=wall, grid_pt_iters=gpi, groups=len(sessions))
