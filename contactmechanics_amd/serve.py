"""Continuous-batching request server for contact solves on one GPU.

Requests (grid side, boundary condition, heights, load, optional hardness,
tolerances) arrive over time.  The server admits a request when its device
footprint fits the memory budget (plan shared per grid + the request's solver
workspace and fields), gives it a resident session on one of a pool of HIP
streams, and steps every active session a chunk of iterations per round
without host synchronisation; a session is polled only once its stream has
drained (event query), finished sessions release their memory and waiting
requests are admitted.  Work of concurrent requests overlaps on the device.
Latency = completion time - arrival time.

Each request runs in its own session (batch of one), so its kernels, launch
shapes and reduction order are the same whether it runs alone or in a mix:
results are bit-identical either way.
"""

import time
from collections import deque

import numpy as np

from .batch import PlanCache
from .session import CcgSession, make_params


def request_bytes(sub, B=1):
    """Device bytes a request needs beyond its (shared) plan."""
    M = int(np.prod(sub.nb_domain_grid_pts))
    nx, ny = sub.nb_grid_pts
    w = 8 if sub.precision == "fp64" else 4
    return CcgSession.workspace_bytes(sub, B) + B * (M * (2 * w + 1) + 8 * nx * ny)


class ContactServer:
    def __init__(self, mem_budget=288 * 2 ** 30, precision="fp32", streams=4, chunk=8, max_chunk=32,
                 max_active=64):
        import torch
        self.cache = PlanCache(precision)
        self.budget = int(mem_budget)
        self.pool = [torch.cuda.Stream() for _ in range(max(1, streams))]
        self.chunk, self.max_chunk, self.max_active = chunk, max_chunk, max_active
        self.used = 0
        self.rejected = []

    def _admit(self, req, now):
        sub = self.cache.get(req["n"], req["bc"])
        need = request_bytes(sub)
        if need > self.budget:
            self.rejected.append((req["index"], need))
            return "reject"
        if self.used + need > self.budget:
            return None
        params = make_params(hardness=req.get("hardness_force"), external_force=req["external_force"],
                             pentol=req["pentol"], forcetol=req["forcetol"], thermotol=req.get("thermotol", 1e-6),
                             maxiter=req.get("maxiter", 100000), use_graph=False)
        stream = self.pool[req["index"] % len(self.pool)]
        s = CcgSession(sub, [req["heights"]], [params], stream=stream)
        self.used += need
        import torch
        ev = torch.cuda.Event()
        return dict(req=req, s=s, need=need, chunk=self.chunk, event=ev, start=now)

    def run(self, requests, realtime=True, on_done=None):
        """Serve ``requests`` (each with an ``arrival`` time in seconds).
        Returns a list of per-request records and a summary."""
        pending = deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = deque()
        active = []
        done = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        sim = 0.0
        while pending or waiting or active:
            now = clock() if realtime else sim
            while pending and pending[0]["arrival"] <= now:
                waiting.append(pending.popleft())
            if not realtime and not active and not waiting and pending:
                sim = pending[0]["arrival"]
                continue
            while waiting and len(active) < self.max_active:
                rec = self._admit(waiting[0], now)
                if rec == "reject":
                    waiting.popleft()
                    continue
                if rec is None:
                    break
                waiting.popleft()
                active.append(rec)
            if not active:
                if realtime and pending:
                    time.sleep(max(0.0, min(0.001, pending[0]["arrival"] - now)))
                continue
            for rec in active:  # one chunk for every active request, no host sync
                rec["s"].step(rec["chunk"])
                rec["event"].record(rec["s"].stream)
            still = []
            for rec in active:
                rec["event"].synchronize()
                fin, _ = rec["s"].poll()
                if fin:
                    stats = rec["s"].finish()[0]
                    end = clock() if realtime else sim
                    out = dict(index=rec["req"]["index"], n=rec["req"]["n"], bc=rec["req"]["bc"],
                               arrival=rec["req"]["arrival"], start=rec["start"], end=end,
                               latency=end - rec["req"]["arrival"], nit=stats["nit"], status=stats["status"],
                               grid=int(np.prod(rec["s"].substrate.nb_domain_grid_pts)))
                    if on_done is not None:
                        on_done(out, rec["s"])
                    rec["s"].close()
                    self.used -= rec["need"]
                    done.append(out)
                else:
                    rec["chunk"] = min(rec["chunk"] * 2, self.max_chunk)
                    still.append(rec)
            active = still
            if not realtime:
                sim += 0.0
        wall = (time.perf_counter() - t0)
        lat = np.array([d["latency"] for d in done]) if done else np.array([np.nan])
        gpi = sum(d["grid"] * d["nit"] for d in done)
        summary = dict(requests=len(done), rejected=len(self.rejected), wall_s=wall,
                       p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                       throughput_req_s=len(done) / wall if wall > 0 else float("nan"),
                       grid_pt_iters_s=gpi / wall if wall > 0 else float("nan"))
        return done, summary
