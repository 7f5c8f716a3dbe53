"""Continuous-batching request server for contact solves on one GPU.

Requests (grid side, boundary condition, heights, load, optional hardness,
tolerances) arrive over time.  The server admits a request when its device
footprint fits the memory budget: the request's solver workspace and fields,
plus the plan of its grid the first time that grid is seen (plans are shared
by every later request of the same grid and stay resident).  An admitted
request gets a resident session on one of a pool of HIP streams.

The loop is event driven.  Every active request has at most one chunk of
iterations in flight; the loop queries the chunk's completion event without
blocking, and only a request whose chunk has completed is polled (its stream
is then idle, so the poll costs one small copy) and given its next chunk.  A
request that finished releases its memory at once and waiting requests are
admitted.  Chunks start at a few iterations and double, capped so that one
chunk of a large grid stays short (about ``chunk_ms`` of GPU time by a
per-point cost estimate): a long request then cannot hold a stream for long,
and short requests on the same stream are not gated by it.  ``cancel(index)``
drops a waiting request or ends an active one after its chunk in flight.

Each request runs in its own session (batch of one), so its kernels, launch
shapes and reduction order are the same whether it runs alone or in a mix:
results are bit-identical either way.  Latency = completion time - arrival.
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
    # GPU time per grid point and iteration used to size chunks (FP32 path on
    # MI355X at 2048^2: ~200 us / 4.2 M points), and the smallest per-iteration
    # time a small grid takes (launch-bound)
    NS_PER_POINT_ITER = 0.05
    MIN_ITER_MS = 0.035

    def __init__(self, mem_budget=288 * 2 ** 30, precision="fp32", streams=4, chunk=8, max_chunk=64,
                 max_active=64, chunk_ms=20.0):
        import torch
        self.cache = PlanCache(precision)
        self.budget = int(mem_budget)
        self.pool = [torch.cuda.Stream() for _ in range(max(1, streams))]
        self.chunk, self.max_chunk, self.max_active = chunk, max_chunk, max_active
        self.chunk_ms = chunk_ms
        self.used = 0
        self.plan_bytes = {}
        self.rejected = []
        self._cancel = set()
        self._next_stream = 0

    def cancel(self, index):
        """Cancel request ``index``: a waiting request is dropped, an active one
        ends after its chunk in flight; its record gets status "cancelled"."""
        self._cancel.add(index)

    def _chunk_cap(self, sub):
        M = int(np.prod(sub.nb_domain_grid_pts))
        it_ms = max(self.MIN_ITER_MS, M * self.NS_PER_POINT_ITER * 1e-6)
        return int(max(1, min(self.max_chunk, self.chunk_ms / it_ms)))

    def _admit(self, req, now):
        import torch
        sub = self.cache.get(req["n"], req["bc"])
        key = (req["n"], req["bc"])
        plan_new = 0 if key in self.plan_bytes else sub.plan_bytes()
        need = request_bytes(sub)
        if need + plan_new > self.budget:
            self.rejected.append((req["index"], need + plan_new))
            return "reject"
        if self.used + need + plan_new > self.budget:
            return None
        if plan_new:
            self.plan_bytes[key] = plan_new
            self.used += plan_new
        params = make_params(hardness=req.get("hardness_force"), external_force=req["external_force"],
                             pentol=req["pentol"], forcetol=req["forcetol"], thermotol=req.get("thermotol", 1e-6),
                             maxiter=req.get("maxiter", 100000), use_graph=False)
        stream = self.pool[self._next_stream % len(self.pool)]
        self._next_stream += 1
        s = CcgSession(sub, [req["heights"]], [params], stream=stream)
        self.used += need
        cap = self._chunk_cap(sub)
        return dict(req=req, s=s, need=need, chunk=min(self.chunk, cap), cap=cap, event=torch.cuda.Event(),
                    inflight=False, start=now)

    def _record(self, rec, end, status, nit):
        req = rec["req"]
        return dict(index=req["index"], n=req["n"], bc=req["bc"], arrival=req["arrival"], start=rec["start"],
                    end=end, latency=end - req["arrival"], nit=nit, status=status,
                    grid=int(np.prod(rec["s"].substrate.nb_domain_grid_pts)))

    def run(self, requests, realtime=True, on_done=None):
        """Serve ``requests`` (each with an ``arrival`` time in seconds and an
        ``index``).  Returns a list of per-request records and a summary."""
        pending = deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = deque()
        active = []
        done = []
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        sim = 0.0
        while pending or waiting or active:
            now = clock() if realtime else sim
            while pending and pending[0]["arrival"] <= now:
                waiting.append(pending.popleft())
            if self._cancel:
                kept = deque()
                for r in waiting:
                    if r["index"] in self._cancel:
                        done.append(dict(index=r["index"], n=r["n"], bc=r["bc"], arrival=r["arrival"], start=None,
                                         end=now, latency=now - r["arrival"], nit=0, status="cancelled", grid=0))
                    else:
                        kept.append(r)
                waiting = kept
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
            progressed = False
            still = []
            for rec in active:
                if rec["inflight"]:
                    if not rec["event"].query():
                        still.append(rec)
                        continue
                    rec["inflight"] = False
                    progressed = True
                    cancelled = rec["req"]["index"] in self._cancel
                    fin, stats = rec["s"].poll()
                    if fin or cancelled:
                        end = clock() if realtime else sim
                        if fin:
                            st = rec["s"].finish()[0]
                            out = self._record(rec, end, st["status"], st["nit"])
                            if on_done is not None:
                                on_done(out, rec["s"])
                        else:
                            out = self._record(rec, end, "cancelled", stats[0]["nit"])
                        rec["s"].close()
                        self.used -= rec["need"]
                        done.append(out)
                        continue
                    rec["chunk"] = min(rec["chunk"] * 2, rec["cap"])
                rec["s"].step(rec["chunk"])
                rec["event"].record(rec["s"].stream)
                rec["inflight"] = True
                progressed = True
                still.append(rec)
            active = still
            if not progressed:
                time.sleep(0.0002)  # every active request has a chunk running
        wall = clock()
        ok = [d for d in done if d["status"] != "cancelled"]
        lat = np.array([d["latency"] for d in ok]) if ok else np.array([np.nan])
        gpi = sum(d["grid"] * d["nit"] for d in ok)
        summary = dict(requests=len(ok), cancelled=len(done) - len(ok), rejected=len(self.rejected), wall_s=wall,
                       p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                       throughput_req_s=len(ok) / wall if wall > 0 else float("nan"),
                       grid_pt_iters_s=gpi / wall if wall > 0 else float("nan"))
        return done, summary
