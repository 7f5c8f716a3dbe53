"""Grids with an odd contiguous dimension.

The row transforms of the HIP kernels pack pairs of neighbouring points of a
row into one complex value, so the computational grid they take has an even
``My``.  The reference's FFT engine takes any size
(``FFTElasticHalfSpace.py:278-296``); a grid ``(Mx, My)`` with odd ``My`` is
computed here on a related grid with even ``My``:

* ``Mx`` even: on the transposed grid ``(My, Mx)`` (the half-space is
  symmetric under exchanging the axes together with their sizes and
  periodicities);
* ``y`` periodic: on two periods ``(Mx, 2 My)``.  A field repeated twice along
  y has a spectrum that vanishes at odd ``ky`` and equals twice the one-period
  spectrum at ``ky = 2k``; the plan holds the one-period Green's function at
  even ``ky`` and zero at odd ``ky`` (``tile_y = 2``), so the convolution of a
  repeated field is the one-period convolution, repeated;
* both odd and ``x`` periodic: transposed, then two periods along the new y;
* both odd and neither direction periodic (a semi-periodic half-space with no
  periodic direction): two periods along y.  The identity above is the
  algebra of a circular convolution of length My and holds for any field, so
  the operators are exact; the solver's contact set is the top-left block of
  the grid, which a repeated field breaks, so such grids have no solve
  (``solvable`` is False).

The maps work on torch tensors on any device; they are pure index algebra
(no transform of their own).
"""


class GridMap:
    def __init__(self, Mx, My, periodic_x, periodic_y):
        self.shape = (int(Mx), int(My))
        self.transpose = False
        self.tile = 1
        if My % 2 == 0:
            pass
        elif Mx % 2 == 0:
            self.transpose = True
        elif periodic_y:
            self.tile = 2
        elif periodic_x:
            self.transpose = True
            self.tile = 2
        else:
            self.tile = 2
        # a solve needs the repeated direction to be periodic (its contact set
        # then covers every period)
        self.solvable = self.tile == 1 or bool(periodic_x if self.transpose else periodic_y)

    @property
    def identity(self):
        return not self.transpose and self.tile == 1

    @property
    def plan_shape(self):
        mx, my = self.shape
        if self.transpose:
            mx, my = my, mx
        return (mx, my * self.tile)

    def plan_args(self, nx, ny, sx, sy, periodic_x, periodic_y):
        """Topography grid, sizes and periodicities of the plan."""
        if self.transpose:
            nx, ny, sx, sy, periodic_x, periodic_y = ny, nx, sy, sx, periodic_y, periodic_x
        return nx, ny * self.tile, sx, sy * self.tile, periodic_x, periodic_y

    # ---- real-space fields (computational or topography grid)
    def fwd(self, t):
        if self.identity:
            return t
        import torch
        if self.transpose:
            t = t.transpose(0, 1)
        if self.tile > 1:
            t = torch.cat([t] * self.tile, dim=1)
        return t.contiguous()

    def inv(self, t):
        if self.identity:
            return t
        if self.tile > 1:
            t = t[:, : t.shape[1] // self.tile]
        if self.transpose:
            t = t.transpose(0, 1)
        return t.contiguous()

    # ---- fourier layout (r2c along axis 0: (Mx//2+1, My)) of the DFT of a
    # real field, or any quantity linear in it
    @staticmethod
    def _full_from_half(F, mx):
        """Full spectrum X[kx, ky] (mx x my) from the half layout
        F = X[:mx//2+1, :] by Hermitian symmetry X[kx, ky] = conj X[-kx, -ky]."""
        import torch
        h, my = F.shape
        kx = torch.arange(h, mx, device=F.device)
        ky = torch.arange(my, device=F.device)
        rest = F[(mx - kx)[:, None], ((my - ky) % my)[None, :]].conj()
        return torch.cat([F, rest], dim=0)

    @classmethod
    def _transpose_half(cls, F, mx, my):
        """Half layout of the transposed field: F is (mx//2+1, my) for a field
        of shape (mx, my); returns (my//2+1, mx)."""
        full = cls._full_from_half(F, mx)  # mx x my
        return full.transpose(0, 1)[: my // 2 + 1].contiguous()

    def fourier_fwd(self, F):
        """Caller's fourier layout -> the plan's."""
        if self.identity:
            return F
        import torch
        mx, my = self.shape
        if self.transpose:
            F = self._transpose_half(F, mx, my)
            mx, my = my, mx
        if self.tile > 1:
            out = torch.zeros((F.shape[0], my * self.tile), dtype=F.dtype, device=F.device)
            out[:, :: self.tile] = F * self.tile
            F = out
        return F.contiguous()

    def fourier_inv(self, Fp):
        """The plan's fourier layout -> the caller's."""
        if self.identity:
            return Fp
        mx, my = self.shape
        F = Fp
        if self.tile > 1:
            F = F[:, :: self.tile] / self.tile
        if self.transpose:
            # the plan grid (before tiling) is (my, mx)
            F = self._transpose_half(F, my, mx)
        return F.contiguous()

    def green_inv(self, Gp_half):
        """The plan's Green's function (real, even in each axis) in its half
        layout (plan Mx//2+1, plan My) -> the caller's half layout; no
        rescaling (G is not a field spectrum)."""
        if self.identity:
            return Gp_half
        mx, my = self.shape
        G = Gp_half
        if self.tile > 1:
            G = G[:, :: self.tile]
        if self.transpose:
            G = self._transpose_half(G.to(_complex_of(G)), my, mx).real
        return G.contiguous()

    # ---- sums over the grid (real-space or power-theorem sums of fields)
    def sum_inv(self, value):
        return value / self.tile


def _complex_of(t):
    import torch
    return torch.complex128 if t.dtype == torch.float64 else torch.complex64
