"""The gfx950 kernels run on the CPU: cm_kernels.hip is compiled for the host
against a small HIP emulation (tests/emu: each GPU thread a fiber, blocks one
at a time, device memory under AddressSanitizer) and driven through the same
C ABI.  Each case runs in a subprocess with AddressSanitizer preloaded, so an
out-of-bounds access in any kernel aborts it; results are compared with the
oracle.  No GPU is needed: this checks kernel indexing and logic before a
kernel ever runs on a device.
"""

import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))


@pytest.fixture(scope="module")
def emu_lib():
    import build_emu
    return build_emu.build()


def run_case(case):
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    proc = subprocess.run([sys.executable, os.path.join(HERE, "emu", "run_case.py"), json.dumps(case)],
                          capture_output=True, text=True, env=env, timeout=600)
    assert proc.returncode == 0, proc.stderr[-3000:]
    assert "AddressSanitizer" not in proc.stderr
    return json.loads(proc.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("nx,ny,bc", [(32, 24, 0), (24, 20, 1)])
def test_emulated_fp64_solve_follows_oracle(emu_lib, nx, ny, bc):
    r = run_case(dict(kind="solve", nx=nx, ny=ny, bc=bc, prec=0, seed=nx + ny, force=3e-3,
                      pentol=1e-10, forcetol=1e-14))
    assert r["rc"] == 0 and r["success_oracle"]
    assert r["nit"] == r["nit_oracle"]
    assert r["rel"] < 1e-10


def test_emulated_plastic_free_solve_follows_oracle(emu_lib):
    r = run_case(dict(kind="solve", nx=32, ny=24, bc=1, prec=0, seed=7, force=3e-3, pentol=1e-10,
                      forcetol=1e-14, hardness=0.02 / (32 * 24)))
    assert r["rc"] == 0 and r["nit"] == r["nit_oracle"] and r["rel"] < 1e-8


@pytest.mark.parametrize("hscale", [1.0, 1e5])
def test_emulated_plastic_mixing_steps_follow_oracle(emu_lib, hscale):
    # a hardness low enough for the plastic mixing step (10 of the first 40
    # iterations), so the bisection runs; the emulated build counts the gaps
    # in blocks of 160, so its rounds run on 5 blocks.  hscale 1e5: the same
    # problem with a bracket about 2^51 xtol wide, past the 48 batched steps
    # into the serial continuation
    r = run_case(dict(kind="solve", nx=32, ny=24, bc=1, prec=0, seed=7, force=3e-3, pentol=1e-10,
                      forcetol=1e-14, hardness=1e-5, hscale=hscale, maxiter=40))
    assert r["mix"] == 10
    assert r["nit"] == r["nit_oracle"] == 40
    assert r["rel"] < 1e-10


@pytest.mark.parametrize("bc,hardness", [(0, None), (1, 6e-5)])
def test_emulated_fp32_solve_reaches_oracle_at_tight_tolerance(emu_lib, bc, hardness):
    # FP32 transforms with FP64 forces and gaps (incremental displacement
    # updates, FP64 refresh every 8 iterations): converges at pentol 1e-10 in
    # the oracle's iteration count; the forces come back rounded to FP32
    case = dict(kind="solve", nx=32 if bc == 0 else 24, ny=32 if bc == 0 else 20, bc=bc, prec=1,
                seed=3 if bc == 0 else 5, force=3e-3, pentol=1e-10, forcetol=1e-14)
    if hardness:
        case["hardness"] = hardness
    r = run_case(case)
    assert r["rc"] == 0 and r["success_oracle"]
    assert abs(r["nit"] - r["nit_oracle"]) <= 1
    assert r["rel"] < 1e-7


@pytest.mark.parametrize("nx,ny,bc", [(48, 40, 0), (24, 30, 1)])
def test_emulated_operators_all_shapes(emu_lib, nx, ny, bc):
    r = run_case(dict(kind="ops", nx=nx, ny=ny, bc=bc))
    assert r["shapes"] >= 4
    assert r["worst"] < 1e-9


@pytest.mark.parametrize("bc,prec", [(0, 0), (1, 1)])
def test_emulated_noalloc_equals_solve(emu_lib, bc, prec):
    # the allocation-free entry point (workspace in the static scratch) gives
    # bitwise the results of cm_ccg_solve, on two calls in a row with the
    # second grid smaller
    grids = [[32, 32], [24, 16]] if bc == 0 else [[24, 20], [16, 12]]
    r = run_case(dict(kind="noalloc", bc=bc, prec=prec, grids=grids))
    for c in r["calls"]:
        assert c["rc"] == [0, 0] and c["nit"][0] == c["nit"][1] > 0 and c["equal"], c


@pytest.mark.parametrize("nx,ny,bc", [(12, 10, 0), (15, 8, 0), (6, 5, 1)])
def test_emulated_kspace_helpers(emu_lib, nx, ny, bc):
    # the fourier-layout unpack (incl. odd Mx and the packed DC/Nyquist
    # column), k-space force and the compensated reductions, against the oracle
    r = run_case(dict(kind="kspace", nx=nx, ny=ny, bc=bc))
    tol = 1e-13 if bc == 0 else 1e-12
    for k, v in r.items():
        assert v < tol, (k, v)


@pytest.mark.parametrize("case", [
    dict(nx=21, ny=22, bc=0),                     # Mx = 3*7, My/2 = 11: generic radices on both axes
    dict(nx=8, ny=14, bc=0),                      # My/2 = 7 (16-element build for a short generic line)
    dict(nx=7, ny=11, bc=1),                      # free 14 x 22
    dict(nx=6, ny=8, bc=2, per=[0, 1]),           # semi-periodic 11 x 8
    dict(nx=5, ny=4, bc=2, per=[1, 1]),           # semi-periodic, images along both axes
    dict(nx=8, ny=7, bc=0, tile=2),               # odd periodic y: two periods, one-period kernel
    dict(nx=6, ny=7, bc=2, per=[0, 1], tile=2),   # semi-periodic 11 x 7 on two periods
    dict(nx=6, ny=7, bc=2, per=[0, 0], tile=2),   # semi-periodic 11 x 13: operators only
])
def test_emulated_operators_any_grid(emu_lib, case):
    # grids with prime factors above 5 (generic Stockham passes between
    # identity end passes), the semi-periodic kernel, and plans holding two
    # periods of an odd y, against the oracle at every launch shape
    r = run_case(dict(kind="ops", **case))
    assert r["shapes"] >= 1
    assert r["worst"] < 1e-11


@pytest.mark.parametrize("case", [
    dict(nx=21, ny=22, bc=0, prec=0, seed=3),
    dict(nx=8, ny=7, bc=0, tile=2, prec=0, seed=4),
    dict(nx=12, ny=10, bc=2, per=[0, 1], prec=0, seed=5),
    dict(nx=10, ny=9, bc=2, per=[0, 1], tile=2, prec=1, seed=6),
])
def test_emulated_solve_any_grid(emu_lib, case):
    r = run_case(dict(kind="solve", force=3e-3, pentol=1e-10, forcetol=1e-14, **case))
    assert r["rc"] == 0 and r["success_oracle"]
    if case["prec"] == 0:
        assert r["nit"] == r["nit_oracle"] and r["rel"] < 1e-11
    else:
        assert abs(r["nit"] - r["nit_oracle"]) <= 1 and r["rel"] < 1e-7
