"""The benchmark workloads at their full size on the GPU.

* free 1024^2 topography (2048^2 computational grid), H = 0.7, RMS slope 0.2,
  hardness 0.05 E*, mean pressures 1e-4 .. 1e-2 E* (8 loads, log spaced),
  each load warm-started from the previous displacements with the plastic
  displacement carried over (reference ``test/test_plastic.py:159-171``):
  - the first load follows the oracle's iterate trajectory for a fixed
    number of iterations (mixing steps with the bearing-area bisection of
    ``ConstrainedConjugateGradients.py:300-330`` included);
  - the whole sweep converges, and every converged load satisfies the
    contact conditions: pressure at most the hardness, total force equal to
    the load, zero gap where 0 < p < H, no penetration anywhere;
* 256 independent instances of uneven size (128^2 .. 1024^2, 25 % free) in
  one batched run: every instance converges, one instance of each size group
  gives bit-identical forces solved alone, and the 128^2 instances match the
  oracle;
* a served Poisson stream of 64 requests (sides 256 .. 4096, half free up to
  8192^2 grids, 20 % with a hardness): every request converges, a sample of
  requests gives bit-identical forces alone and in the mix, p50/p99 latency
  are reported.
"""

import time

import numpy as np
import pytest

from contactmechanics_amd import (FreeFFTElasticHalfSpace, PlasticTopography, Topography, make_plastic_system,
                                  workloads)
from contactmechanics_amd import constrained_conjugate_gradients as ccg_gpu
from oracle.ccg import constrained_conjugate_gradients as ccg_cpu
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _c3():
    cfg = workloads.config3()
    n = cfg["n"]
    h = cfg["heights"]
    area_per_pt = 1.0 / (n * n)
    H_force = cfg["hardness_pressure"] * area_per_pt
    # start from contact at the highest point only (the reference plastic
    # tests start from the heights plus a small penetration for that reason)
    u0 = np.zeros((2 * n, 2 * n))
    u0[:n, :n] = np.maximum(h - h.max() + 1e-4, 0.0)
    return cfg, n, h, H_force, u0


def test_c3_first_load_follows_oracle(gpu):
    cfg, n, h, H_force, u0 = _c3()
    kw = dict(external_force=cfg["loads"][0], hardness=H_force, initial_displacements=u0, pentol=0.0,
              forcetol=0.0, thermotol=0.0, maxiter=8)
    trace = []
    t0 = time.perf_counter()
    rc = ccg_cpu(FreeHalfSpace((n, n), 1.0, (1.0, 1.0)), h, trace=trace, **kw)
    t_cpu = time.perf_counter() - t0
    rg = ccg_gpu(FreeFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0)), Topography(h, (1.0, 1.0)), **kw)
    modes = [t["status"] for t in trace]
    print(f"c3 first load: oracle {t_cpu:.1f}s, modes {modes}, gpu offset {rg.offset:.12e} "
          f"oracle {rc.offset:.12e}", flush=True)
    assert "mix" in modes  # the hardness path with its bisection ran
    assert rg.nit == rc.nit == 8
    # free grids: the Love-kernel spectrum carries ~1e-11 relative FFT
    # round-off; the bisection on the bearing area is a count, so the
    # trajectories agree to that round-off as long as no gap value sits
    # within it of the bisection point
    assert rel(rg.jac, rc.jac) < 1e-8
    assert rel(rg.x, rc.x) < 1e-8
    np.testing.assert_allclose(rg.offset, rc.offset, rtol=1e-8)


def test_c3_warm_started_plastic_sweep_converges(gpu):
    cfg, n, h, H_force, u0 = _c3()
    topo = Topography(h, (1.0, 1.0))
    system = make_plastic_system(FreeFFTElasticHalfSpace((n, n), 1.0, (1.0, 1.0)),
                                 PlasticTopography(topo, cfg["hardness_pressure"]))
    pentol = cfg["pentol"]
    disp = u0
    t0 = time.perf_counter()
    for k, F in enumerate(cfg["loads"]):
        res = system.minimize_proxy(external_force=F, initial_displacements=disp, pentol=pentol,
                                    forcetol=cfg["forcetol"], thermotol=1e-10, maxiter=20000)
        assert res.success, (k, res.message)
        disp = system.disp
        p = system.force / system.area_per_pt  # pressures, compressive positive
        gap = system.compute_gap(system.disp, system.offset)  # with the plastic update applied
        plastic_px = int(np.count_nonzero(system.surface.plastic_displ))
        print(f"c3 load {k}: F={F:.3e} nit={res.nit} area={np.mean(p > 0):.4f} plastic px={plastic_px} "
              f"t={time.perf_counter() - t0:.1f}s", flush=True)
        assert np.max(p) <= cfg["hardness_pressure"] * (1 + 1e-12)
        np.testing.assert_allclose(np.sum(system.force), F, rtol=1e-9)
        elastic = (p > 0) & (p < cfg["hardness_pressure"] * (1 - 1e-9))
        assert np.max(np.abs(gap[elastic])) < 10 * pentol  # contact: no gap
        assert np.min(gap) > -10 * pentol  # nowhere a penetration
    assert np.count_nonzero(system.surface.plastic_displ) > 0


@pytest.fixture(scope="module")
def c4_problems():
    recs = workloads.config4(256)
    return [workloads.materialize(r, 1e-6) for r in recs]


def test_c4_batch_full_size(gpu, c4_problems):
    from contactmechanics_amd.batch import PlanCache, solve_batch
    probs = c4_problems
    cache = PlanCache("fp64")
    t0 = time.perf_counter()
    res, info = solve_batch(probs, precision="fp64", cache=cache)
    print(f"c4 fp64: {len(probs)} instances, {info['groups']} groups, {info['seconds']:.2f}s, "
          f"{info['grid_pt_iters'] / info['seconds']:.3e} grid-pt-iters/s (wall incl. set-up "
          f"{time.perf_counter() - t0:.2f}s)", flush=True)
    assert all(r["stats"]["status"] == 0 for r in res)
    # one instance of every (size, bc) group: bit-identical alone
    seen = set()
    for i, p in enumerate(probs):
        key = (p["n"], p["bc"])
        if key in seen:
            continue
        seen.add(key)
        alone, _ = solve_batch([p], precision="fp64", cache=cache)
        assert alone[0]["stats"]["nit"] == res[i]["stats"]["nit"]
        assert bool((alone[0]["jac"] == res[i]["jac"]).all()), key
    # 128^2 instances against the oracle
    small = [i for i, p in enumerate(probs) if p["n"] == 128][:3]
    for i in small:
        p = probs[i]
        cls = PeriodicHalfSpace if p["bc"] == "periodic" else FreeHalfSpace
        ref = ccg_cpu(cls((128, 128), 1.0, (1.0, 1.0)), p["heights"], external_force=p["external_force"],
                      pentol=p["pentol"], forcetol=p["forcetol"], thermotol=p["thermotol"], maxiter=p["maxiter"])
        assert ref.success
        tol = 1e-10 if p["bc"] == "periodic" else 1e-7
        assert rel(res[i]["jac"].cpu().numpy(), ref.jac) < tol, (i, p["bc"])


def test_c5_served_stream(gpu):
    from contactmechanics_amd.serve import ContactServer
    recs = workloads.config5(64, 4.0, seed=5)
    reqs = [workloads.materialize(r, 1e-6) for r in recs]
    sample = {r["index"] for r in reqs if r["n"] <= 512}
    got = {}

    def keep(rec, s):
        if rec["index"] in sample:
            got[rec["index"]] = s.pressures(0).cpu().numpy().copy()

    srv = ContactServer(precision="fp32", streams=4)
    done, summary = srv.run(reqs, on_done=keep)
    print(f"c5: {summary}", flush=True)
    assert summary["requests"] == 64 and summary["rejected"] == 0
    assert all(d["status"] == 0 for d in done)
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0
    for idx in sorted(sample)[:6]:
        r = reqs[idx]
        alone = {}

        def keep1(rec, s):
            alone["f"] = s.pressures(0).cpu().numpy().copy()

        ContactServer(precision="fp32", streams=1).run([dict(r, arrival=0.0)], on_done=keep1)
        assert np.array_equal(alone["f"], got[idx]), idx
