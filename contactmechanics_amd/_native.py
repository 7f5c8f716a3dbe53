"""ctypes binding of ``libcmhip.so`` (the C ABI declared in ``include/cmhip.h``).

The library is built in-tree by ``contactmechanics_amd.build.build_library()``
(``hipcc --offload-arch=gfx950``).  There is no CPU fallback: if the library or
a HIP device is missing every entry point raises ``NativeUnavailable``.

Device memory is owned by torch-ROCm tensors; only their data pointers cross
the boundary, together with torch's current HIP stream.
"""

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcmhip.so")

CM_CONVERGED = 0
CM_MAXITER = 1
CM_RUNNING = 2
CM_E_ARG = -1
CM_E_SHAPE = -2
CM_E_NAN = -3
CM_E_OOM = -4
CM_E_HIP = -5
CM_E_UNSUPPORTED = -6
CM_E_BISECT = -7
CM_PERIODIC = 0
CM_FREE = 1
CM_SEMI = 2
CM_FP64 = 0
CM_FP32 = 1

DELTA_STR = {0: "reset", 1: "sd", 2: "cg", 3: "mix"}


class NativeUnavailable(RuntimeError):
    """The HIP library or a HIP device is not available."""


class CmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cmhip error {code}: {msg}")
        self.code = code


class SubstrateDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("bc", ctypes.c_int),
                ("precision", ctypes.c_int), ("sx", ctypes.c_double), ("sy", ctypes.c_double),
                ("young", ctypes.c_double), ("poisson", ctypes.c_double), ("q0_mode", ctypes.c_int),
                ("stiffness_q0", ctypes.c_double), ("thickness", ctypes.c_double),
                ("periodic_x", ctypes.c_int), ("periodic_y", ctypes.c_int), ("n_images", ctypes.c_int),
                ("tile_y", ctypes.c_int)]


class CcgParams(ctypes.Structure):
    _fields_ = [("load_control", ctypes.c_int), ("external_force", ctypes.c_double),
                ("offset", ctypes.c_double), ("has_hardness", ctypes.c_int),
                ("hardness", ctypes.c_double), ("pentol", ctypes.c_double),
                ("forcetol", ctypes.c_double), ("thermotol", ctypes.c_double),
                ("mixfac", ctypes.c_double), ("mixdecfac", ctypes.c_double),
                ("minmixsteps", ctypes.c_int), ("maxiter", ctypes.c_int),
                ("init_mode", ctypes.c_int), ("iters_per_sync", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("refresh_every", ctypes.c_int)]


class CcgStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("nit", ctypes.c_int), ("nfev", ctypes.c_int),
                ("last_mode", ctypes.c_int), ("offset", ctypes.c_double), ("fun", ctypes.c_double),
                ("area", ctypes.c_double), ("cg_area", ctypes.c_double),
                ("plastic_area", ctypes.c_double), ("total_force", ctypes.c_double),
                ("rms_pen", ctypes.c_double), ("max_pen", ctypes.c_double),
                ("max_pres", ctypes.c_double), ("pad_pres", ctypes.c_double),
                ("maxdu", ctypes.c_double), ("tau", ctypes.c_double), ("mixfac", ctypes.c_double),
                ("nc_count", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(CcgStats))

# Every symbol declared in include/cmhip.h with its ctypes signature.
_VP = ctypes.c_void_p
SIGNATURES = {
    "cm_version": ([], ctypes.c_char_p),
    "cm_last_error": ([_VP], ctypes.c_char_p),
    "cm_ctx_create": ([ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_VP)], ctypes.c_int),
    "cm_ctx_destroy": ([_VP], ctypes.c_int),
    "cm_ctx_mem_used": ([_VP], ctypes.c_size_t),
    "cm_ctx_stream": ([_VP, ctypes.POINTER(_VP)], ctypes.c_int),
    "cm_plan_create": ([_VP, ctypes.POINTER(SubstrateDesc), ctypes.POINTER(_VP)], ctypes.c_int),
    "cm_plan_destroy": ([_VP], ctypes.c_int),
    "cm_plan_grid": ([_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cm_plan_green_host": ([_VP, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cm_plan_autotune": ([_VP, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
    "cm_plan_set_shape": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "cm_plan_shape": ([_VP] + [ctypes.POINTER(ctypes.c_int)] * 4, ctypes.c_int),
    "cm_evaluate_disp": ([_VP, _VP, _VP, _VP], ctypes.c_int),
    "cm_evaluate_force": ([_VP, _VP, _VP, _VP], ctypes.c_int),
    "cm_fourier": ([_VP, _VP, _VP, ctypes.c_int, _VP], ctypes.c_int),
    "cm_k_force_k": ([_VP, _VP, _VP, _VP], ctypes.c_int),
    "cm_scalar_product_k": ([_VP, _VP, _VP, ctypes.POINTER(ctypes.c_double), _VP], ctypes.c_int),
    "cm_dot": ([_VP, _VP, _VP, ctypes.POINTER(ctypes.c_double), _VP], ctypes.c_int),
    "cm_ccg_solve": ([_VP, ctypes.POINTER(CcgParams), _VP, _VP, _VP, _VP, _VP,
                      ctypes.POINTER(CcgStats), PROGRESS_FN, _VP, _VP], ctypes.c_int),
    "cm_ccg_solve_noalloc": ([_VP, ctypes.POINTER(CcgParams), _VP, _VP, _VP, _VP, _VP,
                              ctypes.POINTER(CcgStats), _VP], ctypes.c_int),
    "cm_ccg_solve_batch": ([_VP, ctypes.c_int, ctypes.POINTER(CcgParams), ctypes.POINTER(_VP),
                            ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                            ctypes.POINTER(_VP), ctypes.POINTER(CcgStats), _VP], ctypes.c_int),
    "cm_ccg_session_bytes": ([_VP, ctypes.c_int], ctypes.c_size_t),
    "cm_ccg_begin": ([_VP, ctypes.c_int, ctypes.POINTER(CcgParams), ctypes.POINTER(_VP),
                      ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                      _VP, ctypes.POINTER(_VP)], ctypes.c_int),
    "cm_ccg_step": ([_VP, ctypes.c_int], ctypes.c_int),
    "cm_ccg_poll": ([_VP, ctypes.POINTER(CcgStats), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cm_ccg_finish": ([_VP, ctypes.POINTER(CcgStats)], ctypes.c_int),
    "cm_ccg_end": ([_VP], ctypes.c_int),
    "cm_ccg_profile": ([_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
}

KERNEL_NAMES = ["phase_a", "bisect", "dir_r2c", "col_conv1", "tau_c2r", "phase_b", "force_r2c",
                "col_conv2", "disp_c2r"]

_lib = None
_lock = threading.Lock()


def load_library(path=None):
    """Load libcmhip.so and bind every C-ABI symbol (no GPU needed).

    ``CMHIP_LIB`` may name another build of the same library (A/B timing)."""
    global _lib
    path = path or os.environ.get("CMHIP_LIB", LIB_PATH)
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise NativeUnavailable(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        lib = ctypes.CDLL(path)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
        return lib


def check(rc, ctx=None):
    """Raise CmError for a negative status code, return rc otherwise."""
    if rc < 0:
        lib = load_library()
        msg = lib.cm_last_error(ctx).decode() if lib else ""
        raise CmError(rc, msg)
    return rc


class Context:
    """One cmhip context (stream + memory accounting) per HIP device."""

    _instances = {}

    def __init__(self, device, mem_budget=0):
        lib = load_library()
        h = ctypes.c_void_p()
        rc = lib.cm_ctx_create(int(device), int(mem_budget), ctypes.byref(h))
        if rc != 0:
            raise NativeUnavailable(f"cm_ctx_create failed: {lib.cm_last_error(None).decode()}")
        self.lib = lib
        self.handle = h
        self.device = device
        sp = ctypes.c_void_p()
        check(lib.cm_ctx_stream(h, ctypes.byref(sp)), h)
        self.stream_handle = sp.value
        import torch
        # torch view of the context stream: torch work that feeds or consumes
        # the solver is ordered against it with events, never host syncs
        self.torch_stream = torch.cuda.ExternalStream(sp.value, device=f"cuda:{device}")

    def run_begin(self):
        """Make the context stream wait for torch's current stream."""
        import torch
        self.torch_stream.wait_stream(torch.cuda.current_stream())
        return ctypes.c_void_p(self.stream_handle)

    def run_end(self):
        """Make torch's current stream wait for the context stream."""
        import torch
        torch.cuda.current_stream().wait_stream(self.torch_stream)

    @classmethod
    def get(cls, device=None):
        import torch
        if not torch.cuda.is_available():
            raise NativeUnavailable("no HIP device visible to torch; the solver has no CPU fallback")
        if device is None:
            device = torch.cuda.current_device()
        with _lock:
            ctx = cls._instances.get(device)
        if ctx is None:
            ctx = Context(device)
            with _lock:
                cls._instances[device] = ctx
        return ctx

    def mem_used(self):
        return int(self.lib.cm_ctx_mem_used(self.handle))


class Plan:
    """Owns a ``cm_plan`` (Green's function, twiddles, launch shapes)."""

    def __init__(self, ctx, nx, ny, bc, precision, sx, sy, young, poisson=0.0,
                 q0_mode=0, stiffness_q0=0.0, thickness=None, periodic_x=0, periodic_y=1, n_images=10,
                 tile_y=1):
        self.ctx = ctx
        self.lib = ctx.lib
        d = SubstrateDesc(nx=int(nx), ny=int(ny), bc=int(bc), precision=int(precision),
                          sx=float(sx), sy=float(sy), young=float(young), poisson=float(poisson),
                          q0_mode=int(q0_mode), stiffness_q0=float(stiffness_q0),
                          thickness=float("nan") if thickness is None else float(thickness),
                          periodic_x=int(bool(periodic_x)), periodic_y=int(bool(periodic_y)),
                          n_images=int(n_images), tile_y=int(tile_y))
        h = ctypes.c_void_p()
        check(self.lib.cm_plan_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)), ctx.handle)
        self.handle = h
        mx, my = ctypes.c_int(), ctypes.c_int()
        check(self.lib.cm_plan_grid(h, ctypes.byref(mx), ctypes.byref(my)), ctx.handle)
        self.grid = (mx.value, my.value)
        self.precision = precision
        self.desc = d

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self.lib.cm_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def green_host(self):
        mx, my = self.grid
        out = np.empty((my // 2 + 1, mx), dtype=np.float64)
        check(self.lib.cm_plan_green_host(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              self.ctx.handle)
        return out

    def autotune(self, reps=10):
        buf = ctypes.create_string_buffer(512)
        check(self.lib.cm_plan_autotune(self.handle, int(reps), buf, 512), self.ctx.handle)
        import json
        return json.loads(buf.value.decode())

    def set_shape(self, Cb, R, erow, ecol):
        check(self.lib.cm_plan_set_shape(self.handle, int(Cb), int(R), int(erow), int(ecol)), self.ctx.handle)

    def shape(self):
        v = [ctypes.c_int() for _ in range(4)]
        check(self.lib.cm_plan_shape(self.handle, *[ctypes.byref(x) for x in v]), self.ctx.handle)
        return dict(zip(["Cb", "R", "erow", "ecol"], [x.value for x in v]))

    def evaluate_disp(self, f_ptr, u_ptr):
        st = self.ctx.run_begin()
        try:
            check(self.lib.cm_evaluate_disp(self.handle, f_ptr, u_ptr, st), self.ctx.handle)
        finally:
            self.ctx.run_end()

    def evaluate_force(self, u_ptr, f_ptr):
        st = self.ctx.run_begin()
        try:
            check(self.lib.cm_evaluate_force(self.handle, u_ptr, f_ptr, st), self.ctx.handle)
        finally:
            self.ctx.run_end()

    def fourier(self, src_ptr, out_ptr, kind):
        """Forward transform into the fourier layout (kind 0 plain, 1 k-disp, 2 k-force)."""
        st = self.ctx.run_begin()
        try:
            check(self.lib.cm_fourier(self.handle, src_ptr, out_ptr, int(kind), st), self.ctx.handle)
        finally:
            self.ctx.run_end()

    def k_force_k(self, in_ptr, out_ptr):
        st = self.ctx.run_begin()
        try:
            check(self.lib.cm_k_force_k(self.handle, in_ptr, out_ptr, st), self.ctx.handle)
        finally:
            self.ctx.run_end()

    def scalar_product_k(self, a_ptr, b_ptr):
        out = ctypes.c_double()
        st = self.ctx.run_begin()
        try:
            check(self.lib.cm_scalar_product_k(self.handle, a_ptr, b_ptr, ctypes.byref(out), st), self.ctx.handle)
        finally:
            self.ctx.run_end()
        return out.value

    def dot(self, a_ptr, b_ptr):
        out = ctypes.c_double()
        st = self.ctx.run_begin()
        try:
            check(self.lib.cm_dot(self.handle, a_ptr, b_ptr, ctypes.byref(out), st), self.ctx.handle)
        finally:
            self.ctx.run_end()
        return out.value
