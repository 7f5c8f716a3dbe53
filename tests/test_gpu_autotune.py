"""Every launch shape the autotuner can choose computes the same operator.

For a few grids, each valid (columns per block, rows per block, register
elements row/column) shape is applied to the plan and evaluate_disp /
evaluate_force are compared with the oracle; then the tuner's pick is cached
and re-applied.
"""

import itertools
import json

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace
from contactmechanics_amd import _native
from contactmechanics_amd.autotune import tune
from oracle.halfspace import FreeHalfSpace, PeriodicHalfSpace

pytestmark = pytest.mark.gpu


# Explicit launch shapes on FP32 plans: a 2048^2 FP32 solve run over
# non-automatic shapes ended in a GPU memory fault whose cause was not found;
# the library refuses explicit shapes for FP32 plans (CM_E_UNSUPPORTED), so
# this case is kept, skipped, until the cause is shown and fixed.
FP32_SHAPES_SKIP = pytest.mark.skip(reason="explicit FP32 launch shapes are refused by the library after an "
                                           "unexplained GPU memory fault in a 2048^2 FP32 solve over them")


@pytest.mark.parametrize("cls,ocls,nx,ny,prec", [
    (PeriodicFFTElasticHalfSpace, PeriodicHalfSpace, 128, 96, "fp64"),
    (PeriodicFFTElasticHalfSpace, PeriodicHalfSpace, 256, 256, "fp64"),
    (PeriodicFFTElasticHalfSpace, PeriodicHalfSpace, 2048, 64, "fp64"),
    (FreeFFTElasticHalfSpace, FreeHalfSpace, 48, 80, "fp64"),
    pytest.param(PeriodicFFTElasticHalfSpace, PeriodicHalfSpace, 256, 256, "fp32", marks=FP32_SHAPES_SKIP),
])
def test_all_shapes_agree(gpu, cls, ocls, nx, ny, prec):
    g = cls((nx, ny), 1.7, (1.0, 2.0), precision=prec)
    c = ocls((nx, ny), 1.7, (1.0, 2.0))
    f = np.random.default_rng(nx + ny).standard_normal(g.nb_domain_grid_pts)
    ud, fd = c.evaluate_disp(f), c.evaluate_force(f)
    tol = 1e-9 if prec == "fp64" else 3e-6
    n_ok = 0
    for Cb, R, er, ec in itertools.product([1, 2, 4, 8], [1, 2, 4, 8], [8, 16, 32], [8, 16, 32]):
        try:
            g.plan.set_shape(Cb, R, er, ec)
        except _native.CmError:
            continue
        n_ok += 1
        assert np.max(np.abs(g.evaluate_disp(f) - ud)) / np.max(np.abs(ud)) < tol, (Cb, R, er, ec)
        assert np.max(np.abs(g.evaluate_force(f) - fd)) / np.max(np.abs(fd)) < tol * 10, (Cb, R, er, ec)
    assert n_ok >= 4


def test_tune_and_cache(gpu, tmp_path):
    cache = str(tmp_path / "tune.json")
    g = PeriodicFFTElasticHalfSpace((256, 256), 1.0, (1.0, 1.0))
    d = tune(g, cache_path=cache, reps=3)
    assert not d["cached"] and d["candidates"] >= 1
    g2 = PeriodicFFTElasticHalfSpace((256, 256), 1.0, (1.0, 1.0))
    d2 = tune(g2, cache_path=cache)
    assert d2["cached"] and g2.plan.shape()["Cb"] == d["Cb"] and g2.plan.shape()["R"] == d["R"]
    with open(cache) as fh:
        assert len(json.load(fh)) == 1
    f = np.random.default_rng(0).standard_normal((256, 256))
    c = PeriodicHalfSpace((256, 256), 1.0, (1.0, 1.0))
    assert np.max(np.abs(g2.evaluate_disp(f) - c.evaluate_disp(f))) / np.max(np.abs(c.evaluate_disp(f))) < 1e-12


def test_fp32_plans_keep_automatic_shape(gpu):
    g = PeriodicFFTElasticHalfSpace((256, 256), 1.0, (1.0, 1.0), precision="fp32")
    with pytest.raises(_native.CmError):
        g.plan.set_shape(1, 1, 16, 16)
    d = tune(g, cache_path=None)
    assert not d["tuned"] and d["Cb"] == g.plan.shape()["Cb"]


def test_32_element_shapes_are_refused(gpu):
    # a solve at a 32-element launch shape faulted the GPU (see the skipped
    # case below): explicit 32-element shapes are refused and the tuner does
    # not try them; the automatic shape stays usable
    g = PeriodicFFTElasticHalfSpace((256, 256), 1.0, (1.0, 1.0))
    auto = g.plan.shape()
    for shp in [(2, 2, 32, 8), (2, 2, 8, 32), (1, 1, 32, 32)]:
        with pytest.raises(_native.CmError):
            g.plan.set_shape(*shp)
    assert g.plan.shape() == auto
    d = tune(g, cache_path=None, reps=2)
    assert d["erow"] != 32 and d["ecol"] != 32


# A 2048^2 FP64 solve at the explicit shape (Cb 2, R 2, 32-element rows,
# 8-element columns) ended in a GPU memory fault; the host build of the same
# kernels ran that case (set-up and one iteration, and every shape at 256^2)
# with no out-of-bounds access, so the cause is not found.  Kept, skipped,
# until it is shown and fixed; the library refuses such solves meanwhile.
@pytest.mark.skip(reason="a 2048^2 FP64 solve with 32-element row transforms faulted the GPU; cause not found, "
                         "the library refuses 32-element shapes for solves")
def test_solve_with_32_element_rows(gpu):
    from contactmechanics_amd import constrained_conjugate_gradients as ccg
    from contactmechanics_amd import workloads
    cfg = workloads.config2(2048)
    g = PeriodicFFTElasticHalfSpace((2048, 2048), 1.0, (1.0, 1.0))
    g.plan.set_shape(2, 2, 32, 8)
    r = ccg(g, cfg["heights"], external_force=cfg["external_force"], pentol=0.0, forcetol=0.0, thermotol=0.0,
            maxiter=20)
    assert r.nit == 20
