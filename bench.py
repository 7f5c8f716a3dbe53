#!/usr/bin/env python3
"""Constrained-CG benchmark on one MI355X.

Workload (default, --workload c2): periodic 2048^2 self-affine surface
(H = 0.8, RMS slope 0.1, seed 2048), mean pressure 0.01 E*, E* = L = 1.

A "step" is one Polonsky-Keer iteration over the whole grid (two FFT
convolutions plus the fused projection and reductions).  The timed region
runs exactly --steps iterations of a resident solve (inputs already in HBM)
with unreachable tolerances, so no iteration is skipped; the reported value
is grid points x iterations / seconds.  A separate full solve with the mapped
tolerances (gtol = 1e-6: pentol = gtol L, forcetol = gtol area_per_pt)
gives time-to-solution and the pressure error of the FP32 path against the
FP64 path.  The CPU baseline times the FP64 NumPy oracle for a bounded number
of iterations of the same workload on the host.

Prints ONE JSON line.
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def kernel_bytes(name, M, Mt, precision="fp64"):
    """Bytes each fused kernel must read + write once per launch.

    FP64 path: every field 8 bytes.  FP32 path: conjugate direction, spectra
    and Green's function 4 bytes, forces and gaps 8 bytes (FP64 masters), and
    between refreshes the displacement pass adds the transform of the force
    change to the stored gap, so it reads no heights.  Contact masks one
    byte.  The spectrum holds M/2 complex values (M words); the Green's
    function (M/2 + Mx) real values, counted as M/2 words.  The start of an
    iteration (contact set, gap sums) is fused into the displacement pass,
    so "phase_a" moves no bytes per iteration."""
    if precision == "fp64":
        w, fw, gw = 8, 8, 8
        heights = 8 * Mt
    else:
        w, fw, gw = 4, 8, 8
        heights = 0
    return {
        "phase_a": 0,
        "bisect": 0,
        "dir_r2c": M * (gw + 2 * w + 1 + w),          # gap, t, mask -> t, S
        # forward column FFT of the direction and sum_k G|T|^2 (reads S and G,
        # writes nothing): the step length needs no inverse pass
        "col_conv1": int(1.5 * w * M),
        "tau_c2r": 0,
        "phase_b": M * (fw + w + gw + 1),              # f, t, gap, mask
        "force_r2c": M * (2 * fw + w + gw + 2 + w),    # f, t, gap, mask -> f, mask, S
        "col_conv2": int(2.5 * w * M),
        "disp_c2r": M * (w + 2 * gw + 1) + heights,  # S, gap, mask -> gap (+ heights)
    }[name]


def canonical_iteration_bytes(M, Mt, w, free):
    """Algorithmic bytes per CCG iteration: w (19 M) periodic, w (19.25 M) free
    (6 FFT passes x 2M words, state 6M + Mt words, + real G for free)."""
    return w * (12 * M + 6 * M + Mt + (M // 4 if free else 0))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_ccg_c2_2048_periodic.json")


def measured_traffic(role, precision="fp64"):
    """HBM bytes per launch of a kernel role ("dir_r2c", "col_conv1", "phase_b",
    "force_r2c", "col_conv2", "disp_c2r") of the FP64 or FP32 path from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (already scaled by the
    copy-kernel factors); None when no such profile is present."""
    try:
        with open(TRAFFIC_FILE) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    for k in d.get("kernels", {}).values():
        if k.get("role") == role and k.get("precision") == precision:
            return k["fetch_bytes"] + k["write_bytes"]
    return None


def throughput(sub, heights, F, steps, warmup, precision):
    import torch
    from contactmechanics_amd.session import CcgSession, make_params
    params = make_params(external_force=F, pentol=0.0, forcetol=0.0, thermotol=0.0,
                         maxiter=warmup + steps + 20, use_graph=False)
    s = CcgSession(sub, [heights], [params])
    s.step(max(warmup, 1))
    s.poll()
    nit0 = s.stats[0].nit
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.poll()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nit1 = s.stats[0].nit
    prof = s.profile(5)
    s.close()
    return dict(seconds=t1 - t0, iterations=nit1 - nit0, kernel_us=prof)


def full_solve(sub, heights, cfg, maxiter=20000):
    import torch
    from contactmechanics_amd import constrained_conjugate_gradients as ccg
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = ccg(sub, heights, external_force=cfg["external_force"], pentol=cfg["pentol"],
            forcetol=cfg["forcetol"], thermotol=cfg["thermotol"], maxiter=maxiter, return_device=True)
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def rel_errors(p, ref):
    import torch
    p = p.to(torch.float64)
    ref = ref.to(torch.float64)
    l2 = float(torch.linalg.norm(p - ref) / torch.linalg.norm(ref))
    linf = float(torch.max(torch.abs(p - ref)) / torch.max(torch.abs(ref)))
    a, b = int((p > 0).sum()), int((ref > 0).sum())
    return dict(l2=l2, linf=linf, area_rel=abs(a - b) / max(b, 1))


def cpu_baseline(cfg, iters, substrate_gpu, workers=None):
    """FP64 NumPy oracle on the same workload: time `iters` iterations (FFTs on
    `workers` threads via scipy.fft when given, else numpy on one thread)."""
    import numpy as np
    from oracle import halfspace as ohs
    from oracle.ccg import constrained_conjugate_gradients as ccg_cpu
    n = cfg["n"]
    hs = ohs.PeriodicHalfSpace((n, n), 1.0, (1.0, 1.0))
    kw = dict(external_force=cfg["external_force"], pentol=0.0, forcetol=0.0, thermotol=0.0)
    ohs.set_fft_workers(workers)
    try:
        # set-up + 1 iteration, then set-up + 1 + iters iterations: the
        # difference times exactly `iters` iterations
        t0 = time.perf_counter()
        ccg_cpu(hs, cfg["heights"], maxiter=1, **kw)
        t_one = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref = ccg_cpu(hs, cfg["heights"], maxiter=iters + 1, **kw)
        t_run = time.perf_counter() - t0 - t_one
    finally:
        ohs.set_fft_workers(None)
    # same iterations on the GPU (FP64) for a same-seed error figure
    from contactmechanics_amd import constrained_conjugate_gradients as ccg_gpu
    g = ccg_gpu(substrate_gpu, cfg["heights"], maxiter=iters + 1, **kw)
    err = float(np.max(np.abs(g.jac - ref.jac)) / np.max(np.abs(ref.jac)))
    return dict(seconds=t_run, iterations=iters, gpu_vs_oracle_linf=err)


def pressure_errors(p, ref, area_per_pt=None):
    """L2 / Linf relative pressure errors and relative contact-area error."""
    import numpy as np
    p = np.asarray(p, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    a, b = int((p > 0).sum()), int((ref > 0).sum())
    return dict(l2=float(np.linalg.norm(p - ref) / np.linalg.norm(ref)),
                linf=float(np.max(np.abs(p - ref)) / np.max(np.abs(ref))),
                area_rel=abs(a - b) / max(b, 1))


def oracle_jac(prob):
    from oracle.ccg import constrained_conjugate_gradients as ccg_cpu
    from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace
    n = prob["n"]
    cls = PeriodicHalfSpace if prob["bc"] == "periodic" else FreeHalfSpace
    r = ccg_cpu(cls((n, n), 1.0, (1.0, 1.0)), prob["heights"], external_force=prob["external_force"],
                hardness=prob.get("hardness"), pentol=prob["pentol"], forcetol=prob["forcetol"],
                thermotol=prob["thermotol"], maxiter=prob["maxiter"])
    return r.jac


def bench_sweep(args):
    """C3: warm-started plastic load sweep on the free 1024^2 topography (2048^2
    computational grid), FP64, each load from the previous solution with the
    plastic displacement carried over."""
    import numpy as np
    import torch
    from contactmechanics_amd import (FreeFFTElasticHalfSpace, PlasticTopography, Topography, make_plastic_system,
                                      workloads)
    cfg = workloads.config3()
    n = cfg["n"]
    h = cfg["heights"]
    system = make_plastic_system(FreeFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0)),
                                 PlasticTopography(Topography(h, (1.0, 1.0)), cfg["hardness_pressure"]))
    u0 = np.zeros((2 * n, 2 * n))
    u0[:n, :n] = np.maximum(h - h.max() + 1e-4, 0.0)  # contact at the highest point only
    # warm-up: the first load's first iterations (one-time module and allocation costs)
    system.minimize_proxy(external_force=cfg["loads"][0], initial_displacements=u0, pentol=cfg["pentol"],
                          forcetol=cfg["forcetol"], thermotol=1e-10, maxiter=3)
    system.surface.plastic_displ[...] = 0.0
    torch.cuda.synchronize()
    disp, loads, seconds = u0, [], 0.0
    for F in cfg["loads"]:
        # timed: the call a sweep makes per load (host arrays in and out, the
        # plastic update of the topography); the property checks below are not
        t0 = time.perf_counter()
        res = system.minimize_proxy(external_force=F, initial_displacements=disp, pentol=cfg["pentol"],
                                    forcetol=cfg["forcetol"], thermotol=1e-10, maxiter=20000)
        seconds += time.perf_counter() - t0
        disp = system.disp
        p = system.force / system.area_per_pt
        loads.append(dict(load=float(F), nit=int(res.nit), success=bool(res.success),
                          area=float(np.mean(p > 0)), plastic_px=int(np.count_nonzero(system.surface.plastic_displ)),
                          max_p_over_H=float(np.max(p) / cfg["hardness_pressure"]),
                          force_rel_err=float(abs(np.sum(system.force) - F) / F)))
        print(f"# c3 load {F:.3e}: {loads[-1]}", file=sys.stderr, flush=True)
    nit = sum(x["nit"] for x in loads)
    M = (2 * n) ** 2
    line = {
        "metric": "C3 plastic load sweep, grid-pt-iters/s (FP64, warm-started)",
        "value": M * nit / seconds, "unit": "grid-pt-iters/s", "n_gpus": 1, "steps": nit, "warmup": 3,
        "ms_per_step": 1e3 * seconds / nit, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": cfg["name"]},
        "seconds": seconds, "iterations": nit, "loads": loads,
    }
    print(json.dumps(line))


def bench_batch(args):
    """C4: 256 independent instances of uneven size in one batched run."""
    import numpy as np
    import torch
    from contactmechanics_amd import workloads
    from contactmechanics_amd.batch import PlanCache, solve_batch
    recs = workloads.config4(args.count)
    t0 = time.perf_counter()
    probs = [workloads.materialize(r, args.gtol) for r in recs]
    print(f"# c4: {len(probs)} instances prepared in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    out = {}
    jac = {}
    for prec in ("fp32", "fp64"):
        cache = PlanCache(prec)
        for n, bc in sorted({(p["n"], p["bc"]) for p in probs}):
            cache.get(n, bc)
        solve_batch(probs[:4], precision=prec, cache=cache)  # warm-up
        res, info = solve_batch(probs, precision=prec, cache=cache)
        ok = sum(r["stats"]["status"] == 0 for r in res)
        out[prec] = dict(seconds=info["seconds"], grid_pt_iters=info["grid_pt_iters"], groups=info["groups"],
                         value=info["grid_pt_iters"] / info["seconds"], instances_per_s=len(probs) / info["seconds"],
                         converged=ok, iterations=int(sum(r["stats"]["nit"] for r in res)))
        jac[prec] = [r["jac"].cpu().numpy() for r in res]
        print(f"# c4 {prec}: {out[prec]}", file=sys.stderr, flush=True)
        del res
        torch.cuda.synchronize()
    e = [pressure_errors(a, b) for a, b in zip(jac["fp32"], jac["fp64"])]
    errs = dict(fp32_vs_fp64_gpu={k: max(x[k] for x in e) for k in e[0]})
    small = [i for i, p in enumerate(probs) if p["n"] == 128][:2]
    if small and not args.no_cpu:
        eo = [pressure_errors(jac["fp64"][i], oracle_jac(probs[i])) for i in small]
        errs["fp64_gpu_vs_oracle_128"] = {k: max(x[k] for x in eo) for k in eo[0]}
    line = {
        "metric": "C4 batched uneven solves, grid-pt-iters/s (FP32 fields, FP64 reductions)",
        "value": out["fp32"]["value"], "unit": "grid-pt-iters/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"C4 {len(probs)} instances, sides {{128..1024}}, 25% free, gtol {args.gtol}"},
        "fp32": out["fp32"], "fp64": out["fp64"], "errors": errs,
    }
    print(json.dumps(line))


def bench_serve(args):
    """C5: Poisson stream of mixed-size requests through the request server."""
    import numpy as np
    import torch
    from contactmechanics_amd import workloads
    from contactmechanics_amd.batch import solve_batch
    from contactmechanics_amd.serve import ContactServer
    recs = workloads.config5(args.count, args.rate, side_max=args.side_max)
    t0 = time.perf_counter()
    reqs = [workloads.materialize(r, args.gtol) for r in recs]
    print(f"# c5: {len(reqs)} requests prepared in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    srv = ContactServer(precision="fp32", streams=4)
    for n, bc in sorted({(r["n"], r["bc"]) for r in reqs}):
        srv.cache.get(n, bc)
    got = {}

    def keep(rec, s):
        if rec["n"] <= args.check_side:
            got[rec["index"]] = s.pressures(0).cpu().numpy().copy()

    done, summary = srv.run(reqs, on_done=keep)
    print(f"# c5 summary: {summary}", file=sys.stderr, flush=True)
    chk = [r for r in reqs if r["index"] in got]
    errs = None
    if chk:
        # served FP32 results against FP64 re-solves at the served tolerance
        # (the difference includes where each run stops), and the precision of
        # the FP32 path itself: FP32 and FP64 re-solves at gtol 1e-10
        res, _ = solve_batch(chk, precision="fp64")
        e = [pressure_errors(got[r["index"]], x["jac"].cpu().numpy()) for r, x in zip(chk, res)]
        tight = []
        for r in chk:
            q = dict(r, **workloads.tolerances(1e-10, r["n"], r["n"]))
            q["thermotol"] = 1e-10
            tight.append(q)
        t32, _ = solve_batch(tight, precision="fp32")
        t64, _ = solve_batch(tight, precision="fp64")
        et = [pressure_errors(a["jac"].cpu().numpy(), b["jac"].cpu().numpy()) for a, b in zip(t32, t64)]
        errs = {"requests_checked": len(chk),
                "served_fp32_vs_fp64_same_gtol": {k: max(x[k] for x in e) for k in e[0]},
                "fp32_path_vs_fp64_at_gtol_1e-10": {k: max(x[k] for x in et) for k in et[0]}}
    lat = sorted(d["latency"] for d in done)
    line = {
        "metric": "C5 served request stream: p50/p99 solve latency at stated throughput",
        "value": summary["p99_latency_s"], "unit": "s (p99 latency)", "n_gpus": 1,
        "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"C5 {len(reqs)} Poisson requests at {args.rate}/s, sides 256..{args.side_max}, "
                               f"half free, 20% hardness, gtol {args.gtol}"},
        "summary": summary, "statuses": [d["status"] for d in done], "errors": errs,
        "max_latency_s": lat[-1] if lat else None,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2 (default): single 2048^2 solve; c3: plastic load sweep; c4: batched uneven solves; "
                         "c5: served stream")
    ap.add_argument("--count", type=int, default=256, help="c4 instances / c5 requests")
    ap.add_argument("--rate", type=float, default=2.0, help="c5 arrival rate (requests/s)")
    ap.add_argument("--side-max", type=int, default=4096, help="c5 largest grid side")
    ap.add_argument("--check-side", type=int, default=512, help="c5: re-solve requests up to this side in FP64")
    ap.add_argument("--gtol", type=float, default=1e-6)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048, help="grid side of the C2 workload")
    ap.add_argument("--cpu-iters", type=int, default=30, help="oracle iterations for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent solves shard across GPUs, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from contactmechanics_amd import PeriodicFFTElasticHalfSpace, workloads

    torch.cuda.set_device(0)
    if args.workload == "c3":
        return bench_sweep(args)
    if args.workload == "c4":
        return bench_batch(args)
    if args.workload == "c5":
        return bench_serve(args)
    if args.calib_copy:
        src = torch.ones(32 * 2 ** 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("# calibration copy: 268435456 bytes read + written per launch", file=sys.stderr, flush=True)
    cfg = workloads.config2(args.n)
    n = cfg["n"]
    M = Mt = n * n
    print(f"# workload: {cfg['name']}", file=sys.stderr, flush=True)
    sub64 = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0), precision="fp64")
    sub32 = PeriodicFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0), precision="fp32")

    tp = {}
    for prec, sub in (("fp64", sub64), ("fp32", sub32)):
        tp[prec] = throughput(sub, cfg["heights"], cfg["external_force"], args.steps, args.warmup, prec)
        print(f"# {prec}: {tp[prec]['iterations']} its in {tp[prec]['seconds']:.4f}s; kernel us "
              f"{ {k: round(v, 2) for k, v in tp[prec]['kernel_us'].items()} }", file=sys.stderr, flush=True)
        assert tp[prec]["iterations"] == args.steps, "iterations were skipped"
    rate = {p: M * args.steps / tp[p]["seconds"] for p in tp}

    # precision of the two paths: full solves at the benchmark tolerance
    # (gtol 1e-6) and at a tight one (gtol 1e-10), FP32 against FP64
    extra = {}
    verified = None
    if not args.no_solve:
        solves = {}
        # untimed 3-iteration solves first: the first solve of a process pays
        # one-time set-up (code-object loading of the set-up kernels, the
        # plan's first session), which is not time-to-solution
        for s in (sub64, sub32):
            full_solve(s, cfg["heights"], dict(cfg, **workloads.tolerances(1e-6, n, n)), maxiter=3)
        for label, g in (("gtol_1e-6", 1e-6), ("gtol_1e-10", 1e-10)):
            c = dict(cfg, **workloads.tolerances(g, n, n))
            if g < 1e-6:
                c["thermotol"] = 1e-10
            r64, t64 = full_solve(sub64, cfg["heights"], c)
            r32, t32 = full_solve(sub32, cfg["heights"], c)
            err = rel_errors(r32.jac, r64.jac)
            solves[label] = dict(
                fp64=dict(nit=int(r64.nit), seconds=round(t64, 4), success=bool(r64.success)),
                fp32=dict(nit=int(r32.nit), seconds=round(t32, 4), success=bool(r32.success)),
                fp32_vs_fp64=err)
            print(f"# solve {label}: {solves[label]}", file=sys.stderr, flush=True)
        extra["solve"] = solves
        tight = solves["gtol_1e-10"]
        verified = (tight["fp32"]["success"] and tight["fp64"]["success"] and tight["fp32_vs_fp64"]["l2"] <= 1e-6
                    and tight["fp32_vs_fp64"]["linf"] <= 1e-6)

    # headline: the FP32-transform path when this run verified its pressures
    # against the FP64 path to <= 1e-6 (L2 and Linf, gtol 1e-10), else FP64
    head = "fp32" if verified else "fp64"
    ku = tp[head]["kernel_us"]
    dom = max((k for k in ku if k != "bisect"), key=lambda k: ku[k])
    b = kernel_bytes(dom, M, Mt, head)
    achieved = b / (ku[dom] * 1e-6) / 1e9
    it_s = tp[head]["seconds"] / args.steps
    iter_bytes = sum(kernel_bytes(k, M, Mt, head) for k in ku)

    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(cfg, args.cpu_iters, sub64)
        cpu = {"value": M * cb["iterations"] / cb["seconds"], "unit": "grid-pt-iters/s", "cores": 1,
               "kind": "port",
               "sample": f"{cb['iterations']} iterations of the FP64 NumPy oracle on the same {n}^2 "
                         f"workload (numpy pocketfft, single thread) in {cb['seconds']:.2f}s",
               "gpu_fp64_vs_oracle_linf_after_same_iterations": cb["gpu_vs_oracle_linf"],
               "host": platform.processor() or platform.machine()}
        print(f"# cpu: {cb}", file=sys.stderr, flush=True)
        # multi-core variant: the same oracle with its FFTs on every core this
        # process may use (capped at 16, the host share of one GPU)
        nw = min(16, len(os.sched_getaffinity(0)))
        if nw > 1:
            cbt = cpu_baseline(cfg, args.cpu_iters, sub64, workers=nw)
            cpu["threaded"] = {"value": M * cbt["iterations"] / cbt["seconds"], "cores": nw,
                               "sample": f"same {cbt['iterations']} iterations, scipy.fft with {nw} workers, "
                                         f"element-wise NumPy on one thread, {cbt['seconds']:.2f}s"}
            print(f"# cpu threaded ({nw}): {cbt}", file=sys.stderr, flush=True)

    out = {
        "metric": "CCG grid-pt-iters/s at rel. pressure error <= 1e-6 vs FP64",
        "value": rate[head],
        "unit": "grid-pt-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": it_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32 transforms, f64 forces/gaps/reductions",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "grid": [n, n], "seed": 2048, "external_force": 0.01,
                   "path": head, "precision_verified_le_1e-6": verified,
                   "timed_region": "resident session, exact --steps iterations, tolerances unreachable"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": measured_traffic(dom, head),
                     "bytes_per_launch": b, "avg_us": ku[dom]},
        "cpu_baseline": cpu,
        "paths": {p: {"value": rate[p], "ms_per_step": tp[p]["seconds"] / args.steps * 1e3,
                      "kernel_us": {k: round(v, 2) for k, v in tp[p]["kernel_us"].items()},
                      "iteration_gbps_algorithmic": sum(kernel_bytes(k, M, Mt, p) for k in tp[p]["kernel_us"])
                      / (tp[p]["seconds"] / args.steps) / 1e9}
                  for p in tp},
        "fp32_speedup_vs_fp64": rate["fp32"] / rate["fp64"],
        "iteration_bytes_algorithmic": iter_bytes,
        "iteration_gbps_canonical": canonical_iteration_bytes(M, Mt, 8 if head == "fp64" else 4, False) / it_s / 1e9,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
