"""FFT elastic half-spaces backed by the gfx950 HIP library.

Drop-in for ``ContactMechanics/FFTElasticHalfSpace.py``:

* ``PeriodicFFTElasticHalfSpace`` (reference :151-885; construction :169-306,
  operators :573-611);
* ``FreeFFTElasticHalfSpace`` (reference :888-1283; Hockney 2N padding
  :950-951, real-space Love kernel :1033-1139, padded/unpadded
  ``evaluate_disp`` :1141-1205, boundary check :1207-1283).

The Green's function and every convolution live on the GPU (``cm_plan_create``
/ ``cm_evaluate_disp`` / ``cm_evaluate_force``).  ``evaluate_disp`` and
``evaluate_force`` accept NumPy arrays (returned as NumPy FP64, like the
reference) or torch tensors on the HIP device (returned as device tensors in
the plan precision, no host copies).

Build-only keyword: ``precision="fp64"`` (default, every field and transform in
FP64), ``"fp32"`` (FP32 fields and transforms, FP64 reductions and scalars) or
``"mixed"`` (operators in FP64; the solver runs its bulk iterations on an FP32
twin of the plan and finishes with FP64 iterations warm-started from the FP32
forces, see ``Optimization/ConstrainedConjugateGradients.py``).
"""

import numpy as np

from . import _native
from .gridmap import GridMap
from .Substrates import ElasticSubstrate


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _tuple2(v, dim):
    if not hasattr(v, "__iter__"):
        v = (v,)
    v = tuple(v)
    return tuple(v[min(i, len(v) - 1)] for i in range(dim))


class PeriodicFFTElasticHalfSpace(ElasticSubstrate):
    """Periodic elastic half-space (Stanley & Kato).  See module docstring."""

    name = "periodic_fft_elastic_halfspace"
    _periodic = True
    _bc = _native.CM_PERIODIC
    periodicity = (True, True)

    def __init__(self, nb_grid_pts, young, physical_sizes=2 * np.pi, stiffness_q0=None,
                 thickness=None, poisson=0.0, superclass=True, fft="serial", communicator=None,
                 fftengine=None, precision="fp64", device=None):
        super().__init__()
        if not hasattr(nb_grid_pts, "__iter__"):
            nb_grid_pts = (nb_grid_pts,)
        self.__dim = len(nb_grid_pts)
        if self.__dim != 2:
            raise self.Error(
                "Dimension of this problem is {}. Only 2-dimensional problems are supported "
                "by the GPU solver".format(self.__dim))
        if stiffness_q0 is not None and thickness is not None:
            raise self.Error("Please specify either stiffness_q0 or thickness or neither.")
        if communicator is not None and getattr(communicator, "size", 1) > 1:
            raise self.Error("MPI domain decomposition is not supported; one GPU holds the whole grid")
        self._nb_grid_pts = tuple(int(n) for n in nb_grid_pts)
        self._physical_sizes = tuple(float(s) for s in _tuple2(physical_sizes, 2))
        self._steps = tuple(s / n for s, n in zip(self._physical_sizes, self._nb_grid_pts))
        self.young = young
        self.poisson = poisson
        self.contact_modulus = young / (1 - poisson ** 2)
        self.stiffness_q0 = stiffness_q0
        self.thickness = thickness
        self._communicator = communicator
        if precision not in ("fp64", "fp32", "mixed"):
            raise self.Error("precision must be 'fp64', 'fp32' or 'mixed'")
        self.precision = precision
        self._device = device
        self._plan = None
        self._plan_bytes = 0
        self._twin = None
        self._gmap = None

    # ---- grid properties (reference :308-430)
    @property
    def dim(self):
        return self.__dim

    @property
    def nb_grid_pts(self):
        return self._nb_grid_pts

    @property
    def physical_sizes(self):
        return self._physical_sizes

    @property
    def area_per_pt(self):
        return np.prod(self.physical_sizes) / np.prod(self.nb_grid_pts)

    @property
    def nb_domain_grid_pts(self):
        return self.nb_grid_pts

    @property
    def nb_subdomain_grid_pts(self):
        return self.nb_domain_grid_pts

    @property
    def topography_nb_subdomain_grid_pts(self):
        return self.nb_subdomain_grid_pts

    @property
    def subdomain_locations(self):
        return (0,) * self.dim

    @property
    def topography_subdomain_locations(self):
        return self.subdomain_locations

    @property
    def subdomain_slices(self):
        return tuple(slice(s, s + n) for s, n in zip(self.subdomain_locations, self.nb_subdomain_grid_pts))

    @property
    def topography_subdomain_slices(self):
        return tuple(slice(s, s + n) for s, n in zip(self.topography_subdomain_locations,
                                                      self.topography_nb_subdomain_grid_pts))

    @property
    def local_topography_subdomain_slices(self):
        return tuple(slice(0, n) for n in self.topography_nb_subdomain_grid_pts)

    @property
    def nb_fourier_grid_pts(self):
        mx, my = self.nb_domain_grid_pts
        return (mx // 2 + 1, my)

    @property
    def fourier_locations(self):
        return (0, 0)

    @property
    def fourier_slices(self):
        return tuple(slice(0, n) for n in self.nb_fourier_grid_pts)

    @property
    def communicator(self):
        return self._communicator

    def __repr__(self):
        size_str = ", ".join(f"{d}: {s}({n})" for d, s, n in zip("xy", self.physical_sizes, self.nb_grid_pts))
        return f"{self.dim}-dimensional halfspace '{self.name}', physical_sizes(nb_grid_pts) in {size_str}, " \
               f"E' = {self.young}"

    # ---- device plan
    def _plan_kwargs(self):
        if self.stiffness_q0 is None:
            q0_mode, q0 = 0, 0.0
        elif self.stiffness_q0 == 0.0:
            q0_mode, q0 = 1, 0.0
        else:
            q0_mode, q0 = 2, float(self.stiffness_q0)
        return dict(poisson=self.poisson, q0_mode=q0_mode, stiffness_q0=q0, thickness=self.thickness)

    @property
    def gridmap(self):
        """How the computational grid maps to the plan's (identity unless My
        is odd; see ``gridmap.py``)."""
        if self._gmap is None:
            self._gmap = GridMap(*self.nb_domain_grid_pts, *self.periodicity)
        return self._gmap

    @property
    def plan(self):
        """The device plan (built on first use: needs a HIP device)."""
        if self._plan is None:
            ctx = _native.Context.get(self._device)
            prec = _native.CM_FP32 if self.precision == "fp32" else _native.CM_FP64
            gm = self.gridmap
            nx, ny, sx, sy, px, py = gm.plan_args(*self.nb_grid_pts, *self.physical_sizes, *self.periodicity)
            used = ctx.mem_used()
            try:
                self._plan = _native.Plan(ctx, nx, ny, self._bc, prec, sx, sy, self.young, periodic_x=px,
                                          periodic_y=py, tile_y=gm.tile, **self._plan_kwargs())
            except _native.CmError as err:
                raise self.Error(str(err)) from err
            self._plan_bytes = ctx.mem_used() - used
        return self._plan

    def plan_bytes(self):
        """Device bytes the plan holds (Green's functions, twiddles, spectrum
        buffer); builds the plan if needed."""
        self.plan  # noqa: B018
        return self._plan_bytes

    @property
    def torch_dtype(self):
        import torch
        return torch.float32 if self.precision == "fp32" else torch.float64

    def fp32_twin(self):
        """The same substrate with an FP32 plan (bulk iterations of "mixed")."""
        if self._twin is None:
            import copy
            twin = copy.copy(self)
            twin.precision = "fp32"
            twin._plan = None
            twin._twin = None
            twin._gmap = None
            self._twin = twin
        return self._twin

    def _device_str(self):
        import torch
        dev = self._device if self._device is not None else torch.cuda.current_device()
        return f"cuda:{dev}"

    def _kernel_ref_layout(self):
        """Green's function (unscaled, real) in the reference's half-spectrum
        layout ``(Mx//2+1, My)``: r2c along axis 0 (reference :119-128)."""
        mx, my = self.gridmap.plan_shape
        g = self.plan.green_host()  # (my//2+1, mx), includes -1/(area*M)
        g = g / (-1.0 / (self.area_per_pt * mx * my))
        j = np.arange(my)
        g = g[np.minimum(j, my - j)][:, : mx // 2 + 1].T.copy()
        if self.gridmap.identity:
            return g
        import torch
        return self.gridmap.green_inv(torch.from_numpy(g)).numpy()

    @property
    def greens_function(self):
        return self._kernel_ref_layout()

    @property
    def surface_stiffness(self):
        g = self._kernel_ref_layout()
        s = np.zeros(g.shape, dtype=complex)
        s[g != 0] = 1.0 / g[g != 0]
        return s

    # ---- operators
    def _apply(self, arr, which):
        import torch
        expected = self.nb_subdomain_grid_pts
        if tuple(arr.shape) != tuple(expected):
            what = "force" if which == "disp" else "displacements"
            raise self.Error(
                f"{what} array has a different shape ({tuple(arr.shape)}) than this halfspace's "
                f"nb_grid_pts ({expected})")
        plan = self.plan
        dev = self._device_str()
        if _is_torch(arr):
            src = arr.to(device=dev, dtype=self.torch_dtype).contiguous()
            ret_torch = True
        else:
            src = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(
                device=dev, dtype=self.torch_dtype)
            ret_torch = False
        gm = self.gridmap
        src = gm.fwd(src)
        out = torch.empty_like(src)
        if which == "disp":
            plan.evaluate_disp(src.data_ptr(), out.data_ptr())
        else:
            plan.evaluate_force(src.data_ptr(), out.data_ptr())
        out = gm.inv(out)
        if ret_torch:
            return out
        return out.to(dtype=torch.float64).cpu().numpy()

    def evaluate_disp(self, forces):
        """Displacements due to pixel forces (reference :573-589)."""
        return self._apply(forces, "disp")

    def evaluate_force(self, disp):
        """Pixel forces due to displacements (reference :591-611)."""
        return self._apply(disp, "force")

    # ---- k-space helpers (reference :613-885)
    @property
    def torch_complex_dtype(self):
        import torch
        return torch.complex64 if self.precision == "fp32" else torch.complex128

    def _grid_tensor(self, arr, what):
        """arr on the computational grid as a device tensor of the plan dtype."""
        import torch
        if tuple(arr.shape) != tuple(self.nb_subdomain_grid_pts):
            raise self.Error(
                f"{what} array has a different shape ({tuple(arr.shape)}) than this halfspace's "
                f"nb_grid_pts ({self.nb_subdomain_grid_pts})")
        if _is_torch(arr):
            return arr.to(device=self._device_str(), dtype=self.torch_dtype).contiguous()
        return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(
            device=self._device_str(), dtype=self.torch_dtype)

    def _fourier_tensor(self, arr):
        import torch
        if tuple(arr.shape) != tuple(self.nb_fourier_grid_pts):
            raise self.Error(f"fourier array has shape {tuple(arr.shape)}, expected {self.nb_fourier_grid_pts}")
        if _is_torch(arr):
            return arr.to(device=self._device_str(), dtype=self.torch_complex_dtype).contiguous()
        return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.complex128)).to(
            device=self._device_str(), dtype=self.torch_complex_dtype)

    @staticmethod
    def _out(t, like):
        import torch
        if _is_torch(like):
            return t
        return t.to(torch.complex128 if t.is_complex() else torch.float64).cpu().numpy()

    def _fourier(self, field, kind, what):
        import torch
        gm = self.gridmap
        src = gm.fwd(self._grid_tensor(field, what))
        pmx, pmy = gm.plan_shape
        out = torch.empty((pmx // 2 + 1, pmy), dtype=self.torch_complex_dtype, device=self._device_str())
        self.plan.fourier(src.data_ptr(), out.data_ptr(), kind)
        return self._out(gm.fourier_inv(out), field)

    def fft(self, field):
        """Unnormalised forward DFT of a real field on the computational grid,
        in the fourier layout ``(Mx//2+1, My)`` (r2c along x, reference
        :119-128)."""
        return self._fourier(field, 0, "field")

    def evaluate_k_disp(self, forces):
        """k-space displacements ``G DFT(-f) / area_per_pt`` (reference :613-633)."""
        return self._fourier(forces, 1, "force")

    def evaluate_k_force(self, disp):
        """k-space forces ``-K DFT(u) area_per_pt`` (reference :635-648)."""
        return self._fourier(disp, 2, "displacements")

    def evaluate_k_force_k(self, disp_k):
        """k-space forces from k-space displacements, ``-K u_k area_per_pt``
        (reference :650-666)."""
        import torch
        gm = self.gridmap
        src = gm.fourier_fwd(self._fourier_tensor(disp_k))
        out = torch.empty_like(src)
        self.plan.k_force_k(src.data_ptr(), out.data_ptr())
        return self._out(gm.fourier_inv(out), disp_k)

    def evaluate_elastic_energy(self, forces, disp):
        """``0.5 sum(u (-f))`` (reference :676-684), compensated FP64 on the
        device; fields on the computational or the topography grid."""
        gm = self.gridmap
        f, u = gm.fwd(self._as_grid(forces)), gm.fwd(self._as_grid(disp))
        return -0.5 * gm.sum_inv(self.plan.dot(u.data_ptr(), f.data_ptr()))

    def _as_grid(self, arr):
        """A field on the computational grid (zero-padded from the topography
        grid if needed) as a device tensor."""
        import torch
        if tuple(np.shape(arr)) == tuple(self.nb_subdomain_grid_pts):
            return self._grid_tensor(arr, "field")
        nx, ny = self.nb_grid_pts
        if tuple(np.shape(arr)) != (nx, ny):
            raise self.Error(f"field has shape {tuple(np.shape(arr))}, expected {self.nb_subdomain_grid_pts} "
                             f"or {(nx, ny)}")
        out = torch.zeros(self.nb_subdomain_grid_pts, dtype=self.torch_dtype, device=self._device_str())
        out[:nx, :ny] = arr if _is_torch(arr) else torch.from_numpy(np.asarray(arr, dtype=np.float64))
        return out

    def evaluate_scalar_product_k_space(self, ka, kb):
        """``sum_ij a_ij b_ij`` from the fourier representations of a and b
        (power theorem over the half spectrum, reference :686-805)."""
        gm = self.gridmap
        a, b = gm.fourier_fwd(self._fourier_tensor(ka)), gm.fourier_fwd(self._fourier_tensor(kb))
        return gm.sum_inv(self.plan.scalar_product_k(a.data_ptr(), b.data_ptr()))

    def evaluate_elastic_energy_k_space(self, kforces, kdisp):
        """``-0.5 sum u f`` from the fourier representations (reference :807-837)."""
        return -0.5 * self.evaluate_scalar_product_k_space(kdisp, kforces)

    def evaluate(self, disp, pot=True, forces=False):
        """Elastic energy and/or forces of a displacement field (reference
        :839-862): with ``forces`` the energy is the real-space
        ``0.5 sum u (-f)``, otherwise it is computed in k-space."""
        force = potential = None
        if forces:
            force = self.evaluate_force(disp)
            if pot:
                potential = self.evaluate_elastic_energy(force, disp)
        elif pot:
            dispk = self.fft(disp)
            potential = self.evaluate_elastic_energy_k_space(self.evaluate_k_force_k(dispk), dispk)
        return potential, force

    def evaluate_k(self, disp_k, pot=True, forces=False):
        """Energy and/or k-space forces of k-space displacements (reference
        :864-885)."""
        force_k = potential = None
        if forces or pot:
            force_k = self.evaluate_k_force_k(disp_k)
            if pot:
                potential = self.evaluate_elastic_energy_k_space(force_k, disp_k)
        return potential, force_k


class FreeFFTElasticHalfSpace(PeriodicFFTElasticHalfSpace):
    """Non-periodic half-space via Hockney's 2N zero padding.  See module docstring."""

    name = "free_fft_elastic_halfspace"
    _periodic = False
    _bc = _native.CM_FREE
    periodicity = (False, False)

    class FreeBoundaryError(Exception):
        """Forces overlap into the padding region (reference :1207-1215)."""

    def __init__(self, nb_grid_pts, young, physical_sizes=2 * np.pi, fft="serial",
                 communicator=None, check_boundaries=False, fftengine=None, precision="fp64",
                 device=None):
        super().__init__(nb_grid_pts, young, physical_sizes, fft=fft, communicator=communicator,
                         precision=precision, device=device)
        self._comp_nb_grid_pts = tuple(2 * r for r in self.nb_grid_pts)
        self._check_boundaries = check_boundaries

    def _plan_kwargs(self):
        return dict(poisson=0.0, q0_mode=0, stiffness_q0=0.0, thickness=None)

    def spawn_child(self, nb_grid_pts):
        size = tuple(nb_grid_pts[i] / float(self.nb_grid_pts[i]) * self.physical_sizes[i]
                     for i in range(self.dim))
        return type(self)(nb_grid_pts, self.young, size, precision=self.precision, device=self._device)

    @property
    def nb_domain_grid_pts(self):
        return self._comp_nb_grid_pts

    @property
    def topography_nb_subdomain_grid_pts(self):
        return self.nb_grid_pts

    @property
    def domain_boundary_mask(self):
        nx, ny = self.nb_grid_pts
        mask = np.zeros((nx, ny), dtype=bool)
        mask[:, 0] = True
        if ny - 1 > 0:
            mask[:, ny - 1] = True
        mask[0, :] = True
        if nx - 1 > 0:
            mask[nx - 1, :] = True
        return mask

    def _kernel_ref_layout(self):
        # the reference stores the complex FFT of the real even kernel; its
        # imaginary part is round-off, the GPU keeps the real part
        return super()._kernel_ref_layout()

    def evaluate_disp(self, forces, bIncludePadding=False):
        """Padded input -> padded output; unpadded input is zero-padded and the
        output cropped unless ``bIncludePadding`` (reference :1141-1205)."""
        shape = tuple(forces.shape)
        if shape == self.nb_subdomain_grid_pts:
            return super().evaluate_disp(forces)
        if shape == self.nb_grid_pts:
            nx, ny = self.nb_grid_pts
            if _is_torch(forces):
                import torch
                padded = torch.zeros(self.nb_domain_grid_pts, dtype=self.torch_dtype,
                                     device=self._device_str())
                padded[:nx, :ny] = forces
            else:
                padded = np.zeros(self.nb_domain_grid_pts)
                padded[:nx, :ny] = forces
            u = super().evaluate_disp(padded)
            return u if bIncludePadding else u[:nx, :ny]
        raise self.Error(
            "force array has a different shape ({0}) than the subdomain nb_grid_pts ({1}), this "
            "halfspace's nb_grid_pts ({2}) or half of it.".format(
                shape, self.nb_subdomain_grid_pts, self.nb_domain_grid_pts))

    def check_boundaries(self, force=None, tol=0):
        """Raise FreeBoundaryError if the force on the outer ring of the
        topography exceeds ``tol`` (reference :1217-1269).  For a masked
        force array, masked pixels are ignored and an edge whose pixels are
        all masked passes."""
        if force is None:
            force = self.force
        if _is_torch(force):
            force = force.detach().to("cpu").double().numpy()
        masked = np.ma.is_masked(force)
        force = np.ma.asarray(force) if masked else np.asarray(force)
        nx, ny = self.topography_nb_subdomain_grid_pts

        def edge_ok(vals):
            small = np.abs(vals) <= tol
            if masked:
                return bool(np.ma.getmaskarray(vals).all() or np.ma.filled(small, True).all())
            return bool(small.all())

        edges = [force[:, 0], force[0, :]]
        if 0 < ny - 1 < force.shape[1]:
            edges.append(force[:, ny - 1])
        if 0 < nx - 1 < force.shape[0]:
            edges.append(force[nx - 1, :])
        if not all(edge_ok(e) for e in edges):
            raise self.FreeBoundaryError(
                "The forces not zero at the boundary of the active domain. This is typically an "
                "indication that the contact geometry exceeds the bounds of the domain. Since this is "
                "a nonperiodic calculation, you may want to increase the size of your domain. If you "
                "are sure that the calculation is correct, set check_boundary to False")

    def check(self, force=None):
        if self._check_boundaries:
            self.check_boundaries(force)


class SemiPeriodicFFTElasticHalfSpace(FreeFFTElasticHalfSpace):
    """Half-space periodic along some directions and free along the others
    (reference ``FFTElasticHalfSpace.py:1286-1615``).

    The computational grid has ``n`` points along a periodic direction and
    ``2n - 1`` (Hockney padding) along a free one (reference :1344-1352).  The
    Green's function is the Love kernel at numpy-fftfreq-wrapped coordinates
    summed over ``n_images`` periodic images per side (reference
    :1354-1497), computed and transformed on the device (``CM_SEMI``).  Odd
    computational sides are handled by the grid map (``gridmap.py``); with no
    periodic direction and both sides odd, the operators work and the solver
    refuses the grid.
    """

    name = "semi_periodic_fft_elastic_halfspace"
    _bc = _native.CM_SEMI

    def __init__(self, nb_grid_pts, young, physical_sizes=(2 * np.pi, 2 * np.pi), periodicity=(False, True),
                 n_images=10, fft="serial", communicator=None, check_boundaries=False, fftengine=None,
                 precision="fp64", device=None):
        super().__init__(nb_grid_pts, young, physical_sizes, fft=fft, communicator=communicator,
                         check_boundaries=check_boundaries, precision=precision, device=device)
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.n_images = int(n_images)
        self._comp_nb_grid_pts = tuple(n if p else 2 * n - 1 for n, p in zip(self.nb_grid_pts, self.periodicity))

    def _plan_kwargs(self):
        return dict(poisson=0.0, q0_mode=0, stiffness_q0=0.0, thickness=None, n_images=self.n_images)

    def spawn_child(self, nb_grid_pts):
        size = tuple(nb_grid_pts[i] / float(self.nb_grid_pts[i]) * self.physical_sizes[i]
                     for i in range(self.dim))
        return type(self)(nb_grid_pts, self.young, size, periodicity=self.periodicity, n_images=self.n_images,
                          precision=self.precision, device=self._device)

    def evaluate_force(self, disp):
        """Forces from displacements; a smaller displacement array is
        zero-padded to the computational grid first (reference :1522-1542)."""
        if tuple(disp.shape) != tuple(self.nb_subdomain_grid_pts):
            if _is_torch(disp):
                import torch
                padded = torch.zeros(self.nb_subdomain_grid_pts, dtype=self.torch_dtype, device=self._device_str())
            else:
                padded = np.zeros(self.nb_subdomain_grid_pts)
            padded[:disp.shape[0], :disp.shape[1]] = disp
            disp = padded
        return super().evaluate_force(disp)

    @property
    def G_real(self):
        """The real-space kernel (sum over images) on the computational grid,
        from the device Green's function (reference :1379-1381 keeps it)."""
        mx, my = self.nb_domain_grid_pts
        return np.fft.irfft2(self.greens_function, s=(my, mx), axes=(1, 0))

    def get_G_real(self):
        """The real-space kernel with G(0, 0) at the array centre (reference
        :1544-1579)."""
        mx, my = self.nb_domain_grid_pts
        m = np.round(np.fft.fftfreq(mx) * mx).astype(int)
        n = np.round(np.fft.fftfreq(my) * my).astype(int)
        return self.G_real[np.ix_(np.argsort(m), np.argsort(n))]

    def get_G_real_slices(self):
        """Middle slices of ``get_G_real`` along x and y (reference :1581-1604)."""
        g = self.get_G_real()
        return g[:, g.shape[1] // 2], g[g.shape[0] // 2, :]
