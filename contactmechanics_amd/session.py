"""Resident CCG solves over the C ABI session entry points.

A ``CcgSession`` holds B instances of one substrate plan: displacements,
forces and contact masks as torch tensors, the solver's own workspace inside
the library.  ``step(n)`` enqueues n iterations without synchronising, so
sessions on different streams overlap on the device; ``poll()`` synchronises
and returns per-instance statistics.  Used by the batched driver, the request
server and the benchmark.
"""

import ctypes

import numpy as np

from . import _native

# solve statistics that are sums over the grid (scaled by the number of
# periods a plan holds, see gridmap.py)
_EXTENSIVE = ("fun", "area", "cg_area", "plastic_area", "total_force", "nc_count")


def make_params(hardness=None, external_force=None, offset=None, pentol=1e-8, forcetol=1e-5,
                thermotol=1e-6, mixfac=0.1, mixdecfac=0.95, minmixsteps=10, maxiter=100000,
                init_mode=0, iters_per_sync=0, use_graph=True, refresh_every=0):
    """``cm_ccg_params`` from the reference keyword arguments."""
    return _native.CcgParams(
        load_control=1 if external_force is not None else 0,
        external_force=float(external_force) if external_force is not None else 0.0,
        offset=float(offset) if offset is not None else float("nan"),
        has_hardness=1 if hardness is not None else 0,
        hardness=float(hardness) if hardness is not None else 0.0,
        pentol=float(pentol), forcetol=float(forcetol), thermotol=float(thermotol),
        mixfac=float(mixfac), mixdecfac=float(mixdecfac), minmixsteps=int(minmixsteps),
        maxiter=int(maxiter), init_mode=int(init_mode), iters_per_sync=int(iters_per_sync),
        use_graph=int(bool(use_graph)), refresh_every=int(refresh_every))


class CcgSession:
    """B instances of one substrate, solved together (one launch per kernel)."""

    def __init__(self, substrate, heights, params, initial_displacements=None, stream=None):
        import torch
        self.substrate = substrate
        self.plan = substrate.plan
        self.lib = self.plan.lib
        self.B = len(heights)
        dev = substrate._device_str()
        dt = substrate.torch_dtype
        # fields live on the plan's grid (the caller's unless My is odd, see
        # gridmap.py); the load and the grid sums scale with its periods
        self.gm = gm = substrate.gridmap
        if not gm.solvable:
            raise substrate.Error(f"no solve on the computational grid {substrate.nb_domain_grid_pts}: both sides "
                                  "odd and no periodic direction")
        shape = gm.plan_shape
        self.h = [gm.fwd(torch.as_tensor(np.ascontiguousarray(h, dtype=np.float64)).to(dev) if not
                         isinstance(h, torch.Tensor) else h.to(device=dev, dtype=torch.float64).contiguous())
                  for h in heights]
        if initial_displacements is None:
            self.u = [torch.zeros(shape, dtype=dt, device=dev) for _ in range(self.B)]
        else:
            self.u = [gm.fwd(torch.as_tensor(u).to(device=dev, dtype=dt).contiguous().clone())
                      for u in initial_displacements]
        self.f = [torch.zeros(shape, dtype=dt, device=dev) for _ in range(self.B)]
        self.cm = [torch.zeros(shape, dtype=torch.uint8, device=dev) for _ in range(self.B)]
        params = list(params)
        if gm.tile > 1:
            scaled = []
            for p in params:
                q = _native.CcgParams()
                ctypes.pointer(q)[0] = p
                q.external_force = p.external_force * gm.tile
                scaled.append(q)
            params = scaled
        self._params = (_native.CcgParams * self.B)(*params)
        vp = lambda xs: (ctypes.c_void_p * self.B)(*[ctypes.c_void_p(x.data_ptr()) for x in xs])  # noqa: E731
        self._hp, self._up, self._fp, self._cp = vp(self.h), vp(self.u), vp(self.f), vp(self.cm)
        self.ctx = self.plan.ctx
        if stream is None:
            self.stream = self.ctx.torch_stream
        else:
            self.stream = stream
        self.stream.wait_stream(torch.cuda.current_stream())
        h = ctypes.c_void_p()
        _native.check(self.lib.cm_ccg_begin(self.plan.handle, self.B, self._params, self._hp, None,
                                            self._up, self._fp, self._cp,
                                            ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)),
                      self.ctx.handle)
        self.handle = h
        self.stats = (_native.CcgStats * self.B)()

    @staticmethod
    def workspace_bytes(substrate, B=1):
        return int(substrate.plan.lib.cm_ccg_session_bytes(substrate.plan.handle, int(B)))

    def step(self, n):
        _native.check(self.lib.cm_ccg_step(self.handle, int(n)), self.ctx.handle)

    def _stats(self):
        out = [s.as_dict() for s in self.stats]
        for d in out:
            for k in _EXTENSIVE:
                d[k] = d[k] / self.gm.tile
        return out

    def poll(self):
        done = ctypes.c_int()
        _native.check(self.lib.cm_ccg_poll(self.handle, self.stats, ctypes.byref(done)), self.ctx.handle)
        return bool(done.value), self._stats()

    def finish(self):
        _native.check(self.lib.cm_ccg_finish(self.handle, self.stats), self.ctx.handle)
        return self._stats()

    def displacements(self, b=0):
        """Displacements on the computational grid (``result.x``)."""
        return self.gm.inv(self.u[b])

    def profile(self, n):
        out = (ctypes.c_float * 9)()
        _native.check(self.lib.cm_ccg_profile(self.handle, int(n), out), self.ctx.handle)
        return dict(zip(_native.KERNEL_NAMES, [float(v) for v in out]))

    def run(self, chunk=4, max_chunk=64):
        """Iterate to completion (done flags polled after each chunk)."""
        while True:
            self.step(chunk)
            done, _ = self.poll()
            if done:
                return self.finish()
            chunk = min(chunk * 2, max_chunk)

    def pressures(self, b=0):
        """Pixel forces on the topography, compressive positive (``result.jac``)."""
        nx, ny = self.substrate.nb_grid_pts
        return -self.gm.inv(self.f[b])[:nx, :ny]

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cm_ccg_end(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
