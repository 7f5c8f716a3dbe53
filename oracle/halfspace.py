"""Oracle FFT elastic half-space operators (TEST INFRASTRUCTURE ONLY).

Restates, in NumPy FP64:

* the DFT convention of the reference (``FFTElasticHalfSpace.py:29-134``):
  forward transform unnormalised, inverse divided by the number of grid
  points, real-to-complex along axis 0 so the half spectrum has shape
  ``(nx//2+1, ny)``.  ``numpy.fft.rfft2(a, axes=(1, 0))`` is exactly that
  layout (pinned by reference test ``test/test_fft_elastic_half_space.py:356-371``
  which asserts ``fftengine.fft == rfftn(a.T).T``);
* the periodic Green's function and stiffness
  (``FFTElasticHalfSpace.py:445-571``) and ``evaluate_disp`` /
  ``evaluate_force`` (``FFTElasticHalfSpace.py:573-611``);
* the free-boundary (Hockney 2N padding) operator
  (``FFTElasticHalfSpace.py:912-964, 1033-1205``) with its real-space Love
  kernel for a uniformly loaded rectangular pixel.

What pins each function (see ``tests/test_oracle_halfspace.py``):

* periodic operators: closed-form sine-wave laws ``p = -E*/2 q u`` and
  ``u = -2 p/(E* q)`` (reference ``test_mpi_periodic_elastic_half_space.py``
  76-193, 253-299), reversibility, ``stiffness_q0`` law;
* free operator: displacement under a uniform pressure patch vs. Love's closed
  form (reference ``test_mpi_free_elastic_half_space.py:128-186``) and a
  brute-force real-space convolution on tiny grids;
* FFT layout: brute-force DFT sums.

``fft_dtype=np.float32`` switches the transforms to single precision
(scipy.fft complex64); that mode exists only to emulate the GPU's FP32 path.
"""

import numpy as np
import scipy.fft


class SubstrateError(Exception):
    """Shape/argument errors (reference ``Substrates.py:39-41``)."""


def _tuple2(v, dim):
    if not hasattr(v, "__iter__"):
        v = (v,)
    v = tuple(v)
    return tuple(v[min(i, len(v) - 1)] for i in range(dim))


# FP64 transforms use numpy's pocketfft (one thread); set_fft_workers(n)
# switches them to scipy.fft with n threads (used to time a multi-core CPU
# baseline; same transforms up to round-off).
_FFT_WORKERS = None


def set_fft_workers(n):
    global _FFT_WORKERS
    _FFT_WORKERS = n


def _rfft2_axis0(a, dtype):
    """Unnormalised r2c along axis 0 -> shape (nx//2+1, ny)."""
    if dtype == np.float32:
        return scipy.fft.rfft2(np.asarray(a, dtype=np.float32), axes=(1, 0))
    if _FFT_WORKERS:
        return scipy.fft.rfft2(a, axes=(1, 0), workers=_FFT_WORKERS)
    return np.fft.rfft2(a, axes=(1, 0))


def _irfft2_axis0(a, shape, dtype):
    """Normalised c2r inverse of :func:`_rfft2_axis0`."""
    if dtype == np.float32:
        return scipy.fft.irfft2(np.asarray(a, dtype=np.complex64), s=shape[::-1],
                                axes=(1, 0))
    if _FFT_WORKERS:
        return scipy.fft.irfft2(a, s=shape[::-1], axes=(1, 0), workers=_FFT_WORKERS)
    return np.fft.irfft2(a, s=shape[::-1], axes=(1, 0))


def periodic_surface_stiffness(nb_grid_pts, physical_sizes, contact_modulus,
                               stiffness_q0=None, thickness=None, young=None,
                               poisson=0.0):
    """Periodic stiffness ``pi E* |q|`` on the half spectrum (nx//2+1, ny).

    Follows ``FFTElasticHalfSpace.py:494-556``: q in cycles per length, the
    r2c axis 0 uses ``qx = i/sx``, axis 1 wraps ``qy = min(j, ny-j)/sy``;
    the q=0 entry follows the ``stiffness_q0`` rule (:546-555) or, for a finite
    thickness, the uniaxial-strain modulus (:524-544).
    """
    nx, ny = nb_grid_pts
    sx, sy = physical_sizes
    ix = np.arange(nx // 2 + 1, dtype=np.float64)
    qx = np.where(ix <= nx // 2, ix / sx, (nx - ix) / sx)
    iy = np.arange(ny, dtype=np.float64)
    qy = np.where(iy <= ny // 2, iy / sy, (ny - iy) / sy)
    q = np.sqrt(qx[:, None] ** 2 + qy[None, :] ** 2)
    q[0, 0] = np.nan
    stiff = np.pi * contact_modulus * q
    if thickness is not None:
        qt = q * (2 * np.pi * thickness)
        fac = 3 - 4 * poisson
        off = 4 * poisson * (2 * poisson - 3) + 5
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            corr = (fac * np.cosh(2 * qt) + 2 * qt ** 2 + off) / (
                fac * np.sinh(2 * qt) - 2 * qt)
        corr[np.isnan(corr)] = 1.0
        stiff = stiff * corr
        stiff[0, 0] = young / thickness * (1 - poisson) / (
            (1 - 2 * poisson) * (1 + poisson))
    else:
        if stiffness_q0 is None:
            stiff[0, 0] = (stiff[1, 0] + stiff[0, 1]) / 2
        elif stiffness_q0 == 0.0:
            stiff[0, 0] = 1.0
        else:
            stiff[0, 0] = stiffness_q0
    return stiff


def love_rectangle_kernel(x, y, a, b, young):
    """Surface displacement at (x, y) under unit pressure on the rectangle
    ``|x|<a, |y|<b`` of an elastic half-space with contact modulus ``young``
    (Johnson 1985, p. 54; reference ``FFTElasticHalfSpace.py:1064-1137``)."""
    def term(p, q, r, s):
        # p*log((q + sqrt(q^2+p^2)) / (r + sqrt(r^2+s^2)))
        return p * np.log((q + np.sqrt(q * q + p * p)) / (r + np.sqrt(r * r + s * s)))

    xp, xm, yp, ym = x + a, x - a, y + b, y - b
    return (term(xp, yp, ym, xp)
            + term(yp, xp, xm, yp)
            + term(xm, ym, yp, xm)
            + term(ym, xm, xp, ym)) / (np.pi * young)


class PeriodicHalfSpace:
    """Oracle ``PeriodicFFTElasticHalfSpace`` (2D, serial).

    ``FFTElasticHalfSpace.py:169-306`` (construction), 573-611 (operators).
    """

    _periodic = True
    Error = SubstrateError

    def __init__(self, nb_grid_pts, young, physical_sizes=2 * np.pi,
                 stiffness_q0=None, thickness=None, poisson=0.0,
                 fft_dtype=np.float64, _green=True):
        self.dim = len(nb_grid_pts)
        if self.dim != 2:
            raise SubstrateError("oracle supports 2D substrates only")
        if stiffness_q0 is not None and thickness is not None:
            raise SubstrateError("Please specify either stiffness_q0 or thickness or neither.")
        self.nb_grid_pts = tuple(int(n) for n in nb_grid_pts)
        self.physical_sizes = tuple(float(s) for s in _tuple2(physical_sizes, 2))
        self._steps = tuple(s / n for s, n in zip(self.physical_sizes, self.nb_grid_pts))
        self.young = young
        self.poisson = poisson
        self.contact_modulus = young / (1 - poisson ** 2)
        self.stiffness_q0 = stiffness_q0
        self.thickness = thickness
        self.fft_dtype = fft_dtype
        self.communicator = None
        self.force = None
        if _green:
            stiff = periodic_surface_stiffness(
                self.nb_grid_pts, self.physical_sizes, self.contact_modulus,
                stiffness_q0, thickness, young, poisson)
            self.greens_function = 1.0 / stiff
            if thickness is None and stiffness_q0 == 0.0:
                self.greens_function[0, 0] = 0.0
            self.surface_stiffness = self._stiffness_from_green()

    def _stiffness_from_green(self):
        # FFTElasticHalfSpace.py:563-571: 1/G where G != 0, else 0.
        g = self.greens_function
        s = np.zeros(g.shape, dtype=complex)
        nz = g != 0
        s[nz] = 1.0 / g[nz]
        return s

    @property
    def nb_domain_grid_pts(self):
        return self.nb_grid_pts

    @property
    def nb_subdomain_grid_pts(self):
        return self.nb_domain_grid_pts

    @property
    def topography_nb_subdomain_grid_pts(self):
        return self.nb_grid_pts

    @property
    def subdomain_locations(self):
        return (0, 0)

    @property
    def area_per_pt(self):
        # FFTElasticHalfSpace.py:319-321 (topography grid, not padded grid)
        return np.prod(self.physical_sizes) / np.prod(self.nb_grid_pts)

    def _convolve(self, field, kernel):
        shape = self.nb_domain_grid_pts
        spec = _rfft2_axis0(field, self.fft_dtype)
        return _irfft2_axis0(spec * kernel, shape, self.fft_dtype)

    def evaluate_disp(self, forces):
        """``u = IDFT(G DFT(-f)) / area_per_pt`` (FFTElasticHalfSpace.py:573-589)."""
        if forces.shape != self.nb_subdomain_grid_pts:
            raise SubstrateError(
                "force array has a different shape ({0}) than this halfspace's "
                "nb_grid_pts ({1})".format(forces.shape, self.nb_subdomain_grid_pts))
        out = self._convolve(-forces, self.greens_function) / self.area_per_pt
        return np.asarray(out, dtype=np.float64)

    def evaluate_force(self, disp):
        """``f = -IDFT(K DFT(u)) * area_per_pt`` (FFTElasticHalfSpace.py:591-611)."""
        if disp.shape != self.nb_subdomain_grid_pts:
            raise SubstrateError(
                "displacements array has a different shape ({0}) than this "
                "halfspace's nb_grid_pts ({1})".format(disp.shape, self.nb_subdomain_grid_pts))
        out = -self._convolve(disp, self.surface_stiffness) * self.area_per_pt
        return np.asarray(out, dtype=np.float64)

    def check(self, force=None):
        pass

    # ---- k-space helpers (FFTElasticHalfSpace.py:613-885), fourier layout
    # (Mx//2+1, My) with the r2c axis along x, unnormalised forward DFT

    @property
    def nb_fourier_grid_pts(self):
        mx, my = self.nb_domain_grid_pts
        return (mx // 2 + 1, my)

    def fft(self, field):
        return _rfft2_axis0(np.asarray(field, dtype=np.float64), np.float64)

    def evaluate_k_disp(self, forces):
        """``G DFT(-f) / area_per_pt`` (:613-633)."""
        if forces.shape != self.nb_subdomain_grid_pts:
            raise SubstrateError("force array has a different shape")
        return self.greens_function * self.fft(-forces) / self.area_per_pt

    def evaluate_k_force(self, disp):
        """``-K DFT(u) area_per_pt`` (:635-648)."""
        if disp.shape != self.nb_subdomain_grid_pts:
            raise SubstrateError("displacements array has a different shape")
        return -self.surface_stiffness * self.fft(disp) * self.area_per_pt

    def evaluate_k_force_k(self, disp_k):
        """``-K u_k area_per_pt`` (:650-666)."""
        return -self.surface_stiffness * disp_k * self.area_per_pt

    def evaluate_elastic_energy(self, forces, disp):
        """``0.5 u . (-f)`` (:676-684)."""
        return 0.5 * np.dot(np.ravel(disp), np.ravel(-forces))

    def evaluate_scalar_product_k_space(self, ka, kb):
        """Power theorem over the half spectrum (:686-805, serial case): the
        rows kx of the r2c axis that stand for two spectrum rows (kx and -kx)
        count twice; kx = 0 once, and the last row kx = Mx//2 once when Mx is
        even (it is its own mirror) and twice when Mx is odd."""
        mx = self.nb_domain_grid_pts[0]
        rows = ka.shape[0]
        fact0 = 1
        factend = 0 if rows == 1 else (2 if mx % 2 == 1 else 1)
        factmiddle = 2 if rows > 2 else 0
        s = (factmiddle * np.vdot(ka[1:-1], kb[1:-1]).real + fact0 * np.vdot(ka[0], kb[0]).real
             + factend * np.vdot(ka[-1], kb[-1]).real)
        return s / np.prod(self.nb_domain_grid_pts)

    def evaluate_elastic_energy_k_space(self, kforces, kdisp):
        """``-0.5 sum u f`` from fourier representations (:807-837)."""
        return -0.5 * self.evaluate_scalar_product_k_space(kdisp, kforces)

    def evaluate(self, disp, pot=True, forces=False):
        """(:839-862) real-space energy with forces, k-space energy without."""
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
        """(:864-885)."""
        force_k = potential = None
        if forces or pot:
            force_k = self.evaluate_k_force_k(disp_k)
            if pot:
                potential = self.evaluate_elastic_energy_k_space(force_k, disp_k)
        return potential, force_k


class FreeHalfSpace(PeriodicHalfSpace):
    """Oracle ``FreeFFTElasticHalfSpace`` (Hockney 2N padding).

    ``FFTElasticHalfSpace.py:912-964`` (grid doubling), 1033-1139 (kernel),
    1141-1205 (padded/unpadded ``evaluate_disp``), 1217-1283 (boundary check).
    """

    _periodic = False

    class FreeBoundaryError(Exception):
        pass

    def __init__(self, nb_grid_pts, young, physical_sizes=2 * np.pi,
                 check_boundaries=False, fft_dtype=np.float64):
        super().__init__(nb_grid_pts, young, physical_sizes, fft_dtype=fft_dtype,
                         _green=False)
        self._check_boundaries = check_boundaries
        self.greens_function = self._free_green()
        self.surface_stiffness = self._stiffness_from_green()

    @property
    def nb_domain_grid_pts(self):
        return tuple(2 * n for n in self.nb_grid_pts)

    def _free_green(self):
        nx, ny = self.nb_grid_pts
        dx, dy = self._steps
        ix = np.arange(2 * nx)
        iy = np.arange(2 * ny)
        # wrapped coordinates: i <= n -> i*dx, else (i - 2n)*dx (:1046-1063)
        x = (np.where(ix <= nx, ix, ix - 2 * nx) * dx)[:, None]
        y = (np.where(iy <= ny, iy, iy - 2 * ny) * dy)[None, :]
        kernel = love_rectangle_kernel(x, y, dx / 2, dy / 2, self.young)
        # forward FFT of the kernel, always in FP64 (construction, :1138)
        return np.fft.rfft2(kernel, axes=(1, 0))

    def evaluate_disp(self, forces, bIncludePadding=False):
        if forces.shape == self.nb_subdomain_grid_pts:
            return super().evaluate_disp(forces)
        if forces.shape == self.nb_grid_pts:
            padded = np.zeros(self.nb_domain_grid_pts)
            padded[:forces.shape[0], :forces.shape[1]] = forces
            u = super().evaluate_disp(padded)
            if bIncludePadding:
                return u
            return u[:self.nb_grid_pts[0], :self.nb_grid_pts[1]]
        raise SubstrateError(
            "force array has a different shape ({0}) than the subdomain nb_grid_pts "
            "({1}), this halfspace's nb_grid_pts ({2}) or half of it.".format(
                forces.shape, self.nb_subdomain_grid_pts, self.nb_domain_grid_pts))

    def check_boundaries(self, force=None, tol=0):
        """Reference ``FFTElasticHalfSpace.py:1217-1269`` (serial): the four
        edges of the topography block must carry |force| <= tol; with a
        masked force, an edge passes if its unmasked values do or if it is
        masked entirely (:1236-1243)."""
        if force is None:
            force = self.force
        nx, ny = self.nb_grid_pts
        if np.ma.is_masked(force):
            def edge(v):
                return bool((abs(v) <= tol).all() or v.mask.all())
        else:
            def edge(v):
                return bool((abs(v) <= tol).all())
        ok = edge(force[:, 0])
        if 0 < ny - 1 < force.shape[1]:
            ok &= edge(force[:, ny - 1])
        ok &= edge(force[0, :])
        if 0 < nx - 1 < force.shape[0]:
            ok &= edge(force[nx - 1, :])
        if not ok:
            raise self.FreeBoundaryError(
                "The forces not zero at the boundary of the active domain.")

    def check(self, force=None):
        if self._check_boundaries:
            self.check_boundaries(force)


class SemiPeriodicHalfSpace(FreeHalfSpace):
    """Oracle ``SemiPeriodicFFTElasticHalfSpace`` (``FFTElasticHalfSpace.py:1286-1497``).

    * computational grid: ``n`` points along a periodic direction, ``2n - 1``
      along a free one (:1344-1352);
    * kernel: the free Love kernel at numpy-fftfreq-wrapped coordinates
      (``fftfreq(M) * M / n * size``, :1422-1433), summed over the periodic
      images shifted by ``i * size`` for ``|i| <= n_images`` along each
      periodic direction, x-images outer (:1363-1376), then forward FFT
      (:1379-1383, kept complex as in the reference);
    * ``evaluate_disp`` as the free half-space (pads an unpadded force array,
      :1499-1520); ``evaluate_force`` zero-pads a smaller displacement array
      (:1522-1542).
    """

    def __init__(self, nb_grid_pts, young, physical_sizes=2 * np.pi, periodicity=(False, True),
                 n_images=10, check_boundaries=False, fft_dtype=np.float64):
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.n_images = int(n_images)
        PeriodicHalfSpace.__init__(self, nb_grid_pts, young, physical_sizes, fft_dtype=fft_dtype,
                                   _green=False)
        self._check_boundaries = check_boundaries
        self.greens_function = self._semi_green()
        self.surface_stiffness = self._stiffness_from_green()

    @property
    def nb_domain_grid_pts(self):
        return tuple(n if p else 2 * n - 1 for n, p in zip(self.nb_grid_pts, self.periodicity))

    def _semi_green(self):
        mx, my = self.nb_domain_grid_pts
        nx, ny = self.nb_grid_pts
        sx, sy = self.physical_sizes
        dx, dy = self._steps
        x0 = (np.fft.fftfreq(mx) * mx / nx * sx)[:, None]
        y0 = (np.fft.fftfreq(my) * my / ny * sy)[None, :]
        kernel = np.zeros((mx, my))
        images_x = range(-self.n_images, self.n_images + 1) if self.periodicity[0] else [0]
        images_y = range(-self.n_images, self.n_images + 1) if self.periodicity[1] else [0]
        for ix in images_x:
            for iy in images_y:
                kernel += love_rectangle_kernel(x0 + float(ix) * sx, y0 + float(iy) * sy, dx / 2, dy / 2,
                                                self.young)
        self.G_real = kernel
        return np.fft.rfft2(kernel, axes=(1, 0))

    def evaluate_force(self, disp):
        if disp.shape != self.nb_subdomain_grid_pts:
            padded = np.zeros(self.nb_subdomain_grid_pts)
            padded[:disp.shape[0], :disp.shape[1]] = disp
            return super().evaluate_force(padded)
        return super().evaluate_force(disp)
