"""Synthetic initial conditions for the benchmark and parity configurations.

All profiles are energy densities in GeV/fm^3 on the reference grid
(cell index is = (DIM_Y*DIM_ETA)*ix + DIM_ETA*iy + ieta, coordinates
x = ix*DX + xmin with xmin = -(DIM_X-1)/2*DX, src/FreeStream.cpp:16-65), fed
through the vector path (IC_ENERGY = 5, src/FreestreamMilne.cpp:251-271).
Random draws use numpy's PCG64 with the seed recorded in each config.

  C0  2-D smooth Gaussian, 101^2, DX = 0.1, N_phi = 51, DTAU = 0.5 (the
      reference's built-in defaults, src/FreestreamMilne.cpp:149-171)
  C1  2-D lumpy: 261^2, DX = 0.06, 200 Gaussian hotspots (width 0.5 fm,
      amplitude Exponential(mean 2 GeV/fm^3)), centres with radius drawn from
      p(r) ~ r / (1 + exp((r - 6.4)/0.54)) and uniform angle; N_phi = 200,
      DTAU = 1.0
  C2  3-D: 131^2 x 101, DX = 0.1, DETA = 0.1, transverse profile as C1 with 100
      hotspots times a flat-top envelope in eta (flat for |eta| < 2, Gaussian
      tails of width 0.5); SIGMA = 1, DIM_RAP = 21, N_phi = 48, DTAU = 0.8
  C3  batch of 2-D events of mixed sizes (see ``batch_mix``)
  C4  serving mix: 80% C3-style 2-D events, 20% C2-style 3-D events with
      DIM_ETA uniform in 41..101; 25% of requests also carry n_B
      (BARYON = 1, SIGMA_B = 0.8, n_B = 0.1 x normalised energy profile)
      (see ``serving_mix``)
"""
from __future__ import annotations

import numpy as np

from . import Params, params_default


def grid_coords(n: int, d: float) -> np.ndarray:
    """Float32 cell coordinates exactly as getX/getY compute them."""
    xmin = np.float32(-1.0 * (float(np.float32(n - 1)) / 2.0) * float(np.float32(d)))
    return (np.arange(n, dtype=np.float32) * np.float32(d) + xmin).astype(np.float32)


def base_params(**kw) -> Params:
    p = params_default()
    p.OUTPUTFORMAT = 2
    p.IC_ENERGY = 5
    p.BARYON = 0
    p.EOS_TYPE = 1
    p.VISCOUS_MATCHING = 1
    p.E_DEP_FS = 0
    p.update(**kw)
    p.DIM = p.DIM_X * p.DIM_Y * p.DIM_ETA
    p.TAU = np.float32(p.TAU0) + np.float32(p.DTAU)
    return p


def gaussian_2d(n: int = 101, d: float = 0.1, amp: float = 15.0, width: float = 1.0) -> np.ndarray:
    x = grid_coords(n, d).astype(np.float64)
    r2 = x[:, None] ** 2 + x[None, :] ** 2
    return (amp * np.exp(-r2 / (2.0 * width ** 2))).astype(np.float32).ravel()


def woods_saxon_radii(rng, n, R=6.4, a=0.54):
    """Rejection sampling of p(r) ~ r / (1 + exp((r - R)/a)) on [0, R + 10 a]."""
    rmax = R + 10.0 * a
    out = []
    fmax = rmax  # r * f(r) <= r <= rmax
    while len(out) < n:
        r = rng.uniform(0.0, rmax, size=4 * n)
        u = rng.uniform(0.0, fmax, size=4 * n)
        keep = r[u < r / (1.0 + np.exp((r - R) / a))]
        out.extend(keep.tolist())
    return np.array(out[:n])


def hotspots_2d(nx: int, ny: int, dx: float, dy: float, nspots: int, seed: int, width: float = 0.5,
                mean_amp: float = 2.0) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    r = woods_saxon_radii(rng, nspots)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=nspots)
    amp = rng.gamma(1.0, mean_amp, size=nspots)
    cx, cy = r * np.cos(phi), r * np.sin(phi)
    x = grid_coords(nx, dx).astype(np.float64)
    y = grid_coords(ny, dy).astype(np.float64)
    e = np.zeros((nx, ny))
    for k in range(nspots):
        gx = np.exp(-((x - cx[k]) ** 2) / (2 * width ** 2))
        gy = np.exp(-((y - cy[k]) ** 2) / (2 * width ** 2))
        e += amp[k] * gx[:, None] * gy[None, :]
    return e.astype(np.float32)


def eta_envelope(neta: int, deta: float, flat: float = 2.0, tail: float = 0.5) -> np.ndarray:
    eta = grid_coords(neta, deta).astype(np.float64)
    a = np.abs(eta)
    return np.where(a < flat, 1.0, np.exp(-((a - flat) ** 2) / (2 * tail ** 2)))


def config(name: str, seed: int | None = None):
    """Returns (params, e) for configuration C0, C1 or C2 (and small test variants)."""
    if name == "C0":
        p = base_params(DIM_X=101, DIM_Y=101, DIM_ETA=1, DIM_RAP=1, DIM_PHIP=51, DX=0.1, DY=0.1,
                        TAU0=0.1, DTAU=0.5, SIGMA=1.0)
        return p, gaussian_2d(101, 0.1, 15.0, 1.0)
    if name == "C1":
        p = base_params(DIM_X=261, DIM_Y=261, DIM_ETA=1, DIM_RAP=1, DIM_PHIP=200, DX=0.06, DY=0.06,
                        TAU0=0.1, DTAU=1.0)
        e = hotspots_2d(261, 261, 0.06, 0.06, 200, 1 if seed is None else seed)
        return p, e.ravel()
    if name == "C2":
        p = base_params(DIM_X=131, DIM_Y=131, DIM_ETA=101, DIM_RAP=21, DIM_PHIP=48, DX=0.1, DY=0.1,
                        DETA=0.1, TAU0=0.1, DTAU=0.8, SIGMA=1.0)
        e2 = hotspots_2d(131, 131, 0.1, 0.1, 100, 2 if seed is None else seed)
        env = eta_envelope(101, 0.1)
        e = (e2[:, :, None] * env[None, None, :]).astype(np.float32)
        return p, e.ravel()
    raise KeyError(name)


def c2_like(neta: int = 101, seed: int = 2, nx: int = 131):
    """A C2-style 3-D event with DIM_ETA = neta (the 3-D requests of C4);
    nx < 131 keeps the central nx^2 window of the same kind of profile."""
    p = base_params(DIM_X=nx, DIM_Y=nx, DIM_ETA=neta, DIM_RAP=21, DIM_PHIP=48, DX=0.1, DY=0.1,
                    DETA=0.1, TAU0=0.1, DTAU=0.8, SIGMA=1.0)
    e2 = hotspots_2d(nx, nx, 0.1, 0.1, 100, seed)
    env = eta_envelope(neta, 0.1)
    return p, (e2[:, :, None] * env[None, None, :]).astype(np.float32).ravel()


def small_3d(nx: int = 41, neta: int = 21, nrap: int = 9, nphi: int = 24, dtau: float = 0.5, seed: int = 5):
    """A reduced 3-D event for parity tests (seconds on the CPU oracle)."""
    p = base_params(DIM_X=nx, DIM_Y=nx, DIM_ETA=neta, DIM_RAP=nrap, DIM_PHIP=nphi, DX=0.1, DY=0.1,
                    DETA=0.1, TAU0=0.1, DTAU=dtau, SIGMA=1.0)
    e2 = hotspots_2d(nx, nx, 0.1, 0.1, 12, seed, width=0.5)
    env = eta_envelope(neta, 0.1, flat=0.5, tail=0.5)
    return p, (e2[:, :, None] * env[None, None, :]).astype(np.float32).ravel()


def batch_mix(n_events: int, seed: int = 7):
    """C3: independent 2-D events, side in {101,141,181,221,261,301}, DX in
    [0.06, 0.1], 20-300 hotspots, N_phi in {48,100,200}, DTAU in [0.4, 1.2]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    events = []
    for k in range(n_events):
        n = int(rng.choice([101, 141, 181, 221, 261, 301]))
        d = float(np.float32(rng.uniform(0.06, 0.1)))
        ns = int(rng.integers(20, 301))
        nphi = int(rng.choice([48, 100, 200]))
        dtau = float(np.float32(rng.uniform(0.4, 1.2)))
        p = base_params(DIM_X=n, DIM_Y=n, DIM_ETA=1, DIM_RAP=1, DIM_PHIP=nphi, DX=d, DY=d, TAU0=0.1, DTAU=dtau)
        e = hotspots_2d(n, n, d, d, ns, seed * 100003 + k)
        events.append((p, e.ravel()))
    return events


def baryon_density(e: np.ndarray, scale: float = 0.1) -> np.ndarray:
    """Initial n_B (fm^-3) as `scale` times the normalised energy profile (the
    baryon-carrying events of the serving mix)."""
    e = np.asarray(e, dtype=np.float64)
    return (scale * e / max(float(e.max()), 1e-30)).astype(np.float32)


def with_baryon(p: Params, sigma_b: float = 0.8) -> Params:
    """Copy of p with BARYON = 1 and SIGMA_B = sigma_b."""
    q = p.copy()
    q.BARYON = 1
    q.SIGMA_B = sigma_b
    return q


def serving_mix(n_events: int, seed: int = 11):
    """C4 request pool: list of (params, e, nB or None).  80% C3-style 2-D
    events, 20% C2-style 3-D events with DIM_ETA uniform in 41..101; 25% of
    requests carry a baryon density (BARYON = 1, SIGMA_B = 0.8)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    twod = batch_mix(n_events, seed=seed)
    # exact proportions in the pool (rounded), in shuffled order
    is3d = np.zeros(n_events, dtype=bool)
    is3d[: int(round(0.2 * n_events))] = True
    isb = np.zeros(n_events, dtype=bool)
    isb[: int(round(0.25 * n_events))] = True
    rng.shuffle(is3d)
    rng.shuffle(isb)
    out = []
    for k in range(n_events):
        if is3d[k]:
            p, e = c2_like(int(rng.integers(41, 102)), seed=seed * 1009 + k)
        else:
            p, e = twod[k]
        nb = None
        if isb[k]:
            p = with_baryon(p)
            nb = baryon_density(e)
        out.append((p, e, nb))
    return out
