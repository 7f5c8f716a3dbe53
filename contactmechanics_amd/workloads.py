"""Synthetic inputs for the benchmark configurations.

C1  periodic 256^2 paraboloid (Hertz), L = 1, E* = 1, R = 1, F = 1e-3 E* L^2.
C2  periodic 2048^2 self-affine surface, H = 0.8, RMS slope 0.1, p/E* = 0.01.
C3  free 1024^2 (2048^2 padded), H = 0.7, RMS slope 0.2, hardness 0.05 E*,
    8 loads log-spaced 1e-4..1e-2 E*, warm-started.
C4  256 instances of uneven sizes {128..1024}^2, 25% free.
C5  Poisson stream of requests, side log-uniform 256..4096 (FFT-friendly).

Self-affine heights: seeded standard-normal white noise, Fourier-filtered to
|q|^-(1+H) between q_long = 2 pi/L and q_short = 2 pi/(4 dx) (zero elsewhere,
including q = 0), then scaled so that the RMS slope sqrt(<|grad h|^2>) computed
spectrally equals the target.

Tolerance mapping for "gtol" (the reference CCG has none): forcetol =
gtol * area_per_pt (a pressure in units of E*, converted to pixel force like
PlasticSystemSpecialisations.py:65), pentol = gtol * L (E* = L = 1),
thermotol = 1e-6 (the reference default).
"""

import numpy as np


def paraboloid(n, L=1.0, R=1.0):
    """C1 heights: -((x-L/2)^2 + (y-L/2)^2)/(2R), x_i = i L / n."""
    x = np.arange(n) * L / n
    return -((x[:, None] - L / 2) ** 2 + (x[None, :] - L / 2) ** 2) / (2 * R)


def self_affine(n, H, slope, seed, L=1.0, short_px=4):
    """Seeded isotropic self-affine surface with the given RMS slope."""
    rng = np.random.default_rng(seed)
    noise = np.fft.rfft2(rng.standard_normal((n, n)))
    qx = 2 * np.pi * np.fft.fftfreq(n, d=L / n)[:, None]
    qy = 2 * np.pi * np.fft.rfftfreq(n, d=L / n)[None, :]
    q = np.hypot(qx, qy)
    q_long, q_short = 2 * np.pi / L, 2 * np.pi / (short_px * L / n)
    band = (q >= q_long * (1 - 1e-12)) & (q <= q_short * (1 + 1e-12))
    amp = np.zeros_like(q)
    amp[band] = q[band] ** (-(1 + H))
    spec = noise * amp
    h = np.fft.irfft2(spec, s=(n, n))
    gx = np.fft.irfft2(1j * qx * spec, s=(n, n))
    gy = np.fft.irfft2(1j * qy * spec, s=(n, n))
    rms_slope = np.sqrt(np.mean(gx ** 2 + gy ** 2))
    return h * (slope / rms_slope)


def tolerances(gtol, nx, ny, L=1.0):
    area_per_pt = L * L / (nx * ny)
    return dict(pentol=gtol * L, forcetol=gtol * area_per_pt, thermotol=1e-6)


def config1():
    n = 256
    return dict(name="C1 periodic 256^2 Hertz paraboloid R=1, F=1e-3 E*L^2", n=n, bc="periodic",
                heights=paraboloid(n), external_force=1e-3, **tolerances(1e-8, n, n))


def config2(n=2048, seed=2048):
    return dict(name=f"C2 periodic {n}^2 self-affine H=0.8 rms-slope=0.1 p/E*=0.01", n=n,
                bc="periodic", heights=self_affine(n, 0.8, 0.1, seed), external_force=0.01,
                **tolerances(1e-6, n, n))


def config3(n=1024, seed=1024):
    return dict(name=f"C3 free {n}^2 (padded {2 * n}^2) self-affine H=0.7 rms-slope=0.2 hardness=0.05E*, "
                     "8 loads 1e-4..1e-2",
                n=n, bc="free", heights=self_affine(n, 0.7, 0.2, seed), hardness_pressure=0.05,
                loads=list(np.logspace(-4, -2, 8)), **tolerances(1e-6, n, n))


C4_SIDES = (128, 192, 256, 384, 512, 768, 1024)


def config4(count=256, seed=4):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(C4_SIDES))
        free = bool(rng.random() < 0.25)
        H = float(rng.uniform(0.6, 0.9))
        slope = float(rng.uniform(0.05, 0.3))
        p = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        out.append(dict(index=i, n=n, bc="free" if free else "periodic", H=H, slope=slope, pressure=p,
                        seed=1000 + i))
    return out


def fft_friendly(n):
    """Smallest 2^a 3^b 5^c >= n."""
    m = n
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1 and m % 2 == 0:
            return m
        m += 1


def config5(count, rate_hz, seed=5, side_min=256, side_max=4096):
    """Poisson arrivals; sides log-uniform, rounded up to 2^a 3^b 5^c."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate_hz)
        side = fft_friendly(int(np.exp(rng.uniform(np.log(side_min), np.log(side_max)))))
        free = bool(rng.random() < 0.5)  # free requests run on the padded 2n grid
        hard = bool(rng.random() < 0.2)
        out.append(dict(index=i, arrival=t, n=side, bc="free" if free else "periodic",
                        H=float(rng.uniform(0.6, 0.9)), slope=float(rng.uniform(0.05, 0.3)),
                        pressure=float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2)))),
                        hardness=float(rng.uniform(0.02, 0.1)) if hard else None, seed=5000 + i))
    return out


def materialize(rec, gtol=1e-6, maxiter=20000):
    """Solver inputs for one C4/C5 record: heights, mean pressure -> total
    force (pressure x L^2, L = 1), hardness pressure -> pixel hardness
    (pressure x area_per_pt), mapped tolerances."""
    n = rec["n"]
    out = dict(rec)
    out["heights"] = self_affine(n, rec["H"], rec["slope"], rec["seed"])
    out["external_force"] = rec["pressure"]
    hp = rec.get("hardness")
    out["hardness"] = hp / (n * n) if hp is not None else None
    out["hardness_force"] = out["hardness"]
    out["maxiter"] = maxiter
    out.update(tolerances(gtol, n, n))
    return out
