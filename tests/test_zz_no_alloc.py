"""The allocation-free solve entry point (``cm_ccg_solve_noalloc``) against
``cm_ccg_solve`` through the C ABI on the GPU.

Two calls in a row on the same GPU (the static scratch keeps the first
call's contents), the second on a smaller grid, in both precisions and both
boundary conditions: status, iteration count and the returned displacement,
force and contact-mask fields are bitwise equal to the allocating entry
point's (same kernels, same order).  A solve whose workspace exceeds the
static scratch is refused with CM_E_OOM, and nothing is allocated by the
refused call (the context's device-memory count does not change).
"""

import ctypes

import numpy as np
import pytest

from contactmechanics_amd import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace, _native, workloads

pytestmark = pytest.mark.gpu


def _params(F):
    return _native.CcgParams(load_control=1, external_force=F, offset=float("nan"), has_hardness=0,
                             hardness=0.0, pentol=1e-9, forcetol=1e-14, thermotol=1e-8, mixfac=0.1,
                             mixdecfac=0.95, minmixsteps=10, maxiter=5000, init_mode=0, iters_per_sync=0,
                             use_graph=1, refresh_every=0)


def _solve(sub, h, entry):
    import torch
    plan = sub.plan
    dev = sub._device_str()
    shape = tuple(sub.nb_domain_grid_pts)
    u = torch.zeros(shape, dtype=sub.torch_dtype, device=dev)
    f = torch.zeros(shape, dtype=sub.torch_dtype, device=dev)
    cm = torch.zeros(shape, dtype=torch.uint8, device=dev)
    hd = torch.from_numpy(np.ascontiguousarray(h)).to(dev)
    st = _native.CcgStats()
    prm = _params(3e-3)
    torch.cuda.synchronize()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(plan.ctx.stream_handle)
    if entry == "solve":
        rc = plan.lib.cm_ccg_solve(plan.handle, ctypes.byref(prm), vp(hd), None, vp(u), vp(f), vp(cm),
                                   ctypes.byref(st), _native.PROGRESS_FN(), None, stream)
    else:
        rc = plan.lib.cm_ccg_solve_noalloc(plan.handle, ctypes.byref(prm), vp(hd), None, vp(u), vp(f), vp(cm),
                                           ctypes.byref(st), stream)
    torch.cuda.synchronize()
    return rc, st.nit, u.cpu(), f.cpu(), cm.cpu()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("bc", ["periodic", "free"])
def test_noalloc_equals_solve_twice(gpu, precision, bc):
    cls = PeriodicFFTElasticHalfSpace if bc == "periodic" else FreeFFTElasticHalfSpace
    for n in (256, 96):  # second call smaller
        sub = cls((n, n), 1.0, (1.0, 1.0), precision=precision)
        h = workloads.self_affine(n, 0.8, 0.1, 7 + n)
        a = _solve(sub, h, "solve")
        b = _solve(sub, h, "noalloc")
        print(f"noalloc {precision} {bc} {n}: rc {a[0]}/{b[0]} nit {a[1]}/{b[1]}", flush=True)
        assert a[0] == b[0] == 0 and a[1] == b[1] > 0
        for x, y in zip(a[2:], b[2:]):
            assert torch_equal(x, y)


def torch_equal(x, y):
    import torch
    return bool(torch.equal(x, y))


def test_noalloc_refuses_beyond_scratch(gpu):
    import torch
    # 4096^2 FP64: conjugate direction and spectrum alone need 256 MiB
    sub = PeriodicFFTElasticHalfSpace((4096, 4096), 1.0, (1.0, 1.0), precision="fp64")
    plan = sub.plan
    ctx = plan.ctx
    used = ctx.mem_used()
    u = torch.zeros((4096, 4096), dtype=torch.float64, device="cuda")
    f = torch.zeros_like(u)
    cm = torch.zeros((4096, 4096), dtype=torch.uint8, device="cuda")
    h = torch.zeros((4096, 4096), dtype=torch.float64, device="cuda")
    st = _native.CcgStats()
    prm = _params(1e-3)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rc = plan.lib.cm_ccg_solve_noalloc(plan.handle, ctypes.byref(prm), vp(h), None, vp(u), vp(f), vp(cm),
                                       ctypes.byref(st), None)
    assert rc == _native.CM_E_OOM
    assert "static scratch" in plan.lib.cm_last_error(ctx.handle).decode()
    assert ctx.mem_used() == used
