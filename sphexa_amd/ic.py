"""Seeded synthetic initial conditions for the hydro-path workloads (host numpy arrays).

* sedov_lattice  -- jittered cubic lattice in the periodic box [-0.5,0.5]^3, equal masses
  (total 1), v = 0, u = 1e-5 plus a Gaussian energy deposit of total 1 (sigma = 2 h0); the
  Sedov set-up of main/src/init/sedov_init.hpp:48-95 with the lattice of
  main/src/init/grid.hpp:101-130 and h0 = 1/2 cbrt(3 ng0 V / (4 pi N)) (sedov_init.hpp:55).
* turbulence_box -- jittered lattice in the periodic unit box, rho = 1, u = 1, and a seeded
  solenoidal Gaussian velocity field with P(k) ~ k^-4 on integer modes 1 <= |k| <= 2, scaled
  to rms Mach 0.3 (the subsonic-turbulence workload; cf. main/src/init/turbulence_init.hpp).
* evrard_sphere  -- N particles in the unit sphere with rho ~ 1/r (r = U^(1/2)), isotropic
  angles, u = 0.05, v = 0, open box [-1,1]^3, h from the local density
  (cf. main/src/init/evrard_init.hpp:55-106).

Particles are returned unsorted; the SFC sort happens in the domain sync.
"""
import math

import numpy as np

GAMMA = 5.0 / 3.0


def h_lattice(n_total, volume, ng0=100):
    return 0.5 * (3.0 * ng0 * volume / (4.0 * math.pi * n_total)) ** (1.0 / 3.0)


def jittered_lattice(side, lo, length, rng, jitter=0.1):
    dx = length / side
    g = (np.arange(side, dtype=np.float64) + 0.5) * dx + lo
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    out = []
    for c in (x, y, z):
        c = c.ravel()
        c = c + rng.uniform(-jitter, jitter, c.size) * dx
        out.append(np.ascontiguousarray(c))
    return out


def sedov_lattice(side, seed=42, ng0=100, jitter=0.1):
    rng = np.random.default_rng(seed)
    n = side ** 3
    x, y, z = jittered_lattice(side, -0.5, 1.0, rng, jitter)
    h0 = h_lattice(n, 1.0, ng0)
    m = np.full(n, 1.0 / n)
    sigma = 2.0 * h0
    r2 = x * x + y * y + z * z
    g = np.exp(-r2 / (2.0 * sigma * sigma))
    u = 1e-5 + g / np.sum(m * g)  # deposit normalised to total energy 1
    zeros = np.zeros(n)
    return dict(x=x, y=y, z=z, h=np.full(n, h0), m=m, u=u, vx=zeros.copy(), vy=zeros.copy(), vz=zeros.copy(),
                alpha=np.full(n, 0.05)), dict(lo=-0.5, hi=0.5, periodic=True)


def _solenoidal_field(x, y, z, rng, kmin=1.0, kmax=2.0, slope=-4.0):
    modes = []
    for kx in range(-2, 3):
        for ky in range(-2, 3):
            for kz in range(-2, 3):
                kk = math.sqrt(kx * kx + ky * ky + kz * kz)
                if kmin <= kk <= kmax and (kx, ky, kz) > (0, 0, 0):
                    modes.append((kx, ky, kz, kk))
    v = [np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)]
    for kx, ky, kz, kk in modes:
        khat = np.array([kx, ky, kz]) / kk
        amp = math.sqrt(kk ** slope)
        a = rng.normal(size=3)
        b = rng.normal(size=3)
        a -= khat * np.dot(a, khat)  # transverse => divergence free
        b -= khat * np.dot(b, khat)
        phase = 2.0 * math.pi * (kx * x + ky * y + kz * z)
        cs, sn = np.cos(phase), np.sin(phase)
        for d in range(3):
            v[d] += amp * (a[d] * cs + b[d] * sn)
    return v


def turbulence_box(side, seed=42, ng0=100, mach=0.3, jitter=0.1):
    rng = np.random.default_rng(seed)
    n = side ** 3
    x, y, z = jittered_lattice(side, -0.5, 1.0, rng, jitter)
    u0 = 1.0
    cs = math.sqrt(GAMMA * (GAMMA - 1.0) * u0)
    vx, vy, vz = _solenoidal_field(x, y, z, rng)
    vrms = math.sqrt(np.mean(vx * vx + vy * vy + vz * vz))
    s = mach * cs / vrms
    h0 = h_lattice(n, 1.0, ng0)
    return dict(x=x, y=y, z=z, h=np.full(n, h0), m=np.full(n, 1.0 / n), u=np.full(n, u0), vx=vx * s, vy=vy * s,
                vz=vz * s, alpha=np.full(n, 0.05)), dict(lo=-0.5, hi=0.5, periodic=True)


def evrard_sphere(n, seed=42, ng0=100, radius=1.0):
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.uniform(0.0, 1.0, n))
    cos_t = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sin_t = np.sqrt(1.0 - cos_t * cos_t)
    x, y, z = r * sin_t * np.cos(phi), r * sin_t * np.sin(phi), r * cos_t
    m = np.full(n, 1.0 / n)
    rho = 1.0 / (2.0 * math.pi * radius * radius * np.maximum(r, 1e-6 * radius))
    h = 0.5 * np.cbrt(3.0 * ng0 * m / (4.0 * math.pi * rho))
    zeros = np.zeros(n)
    return dict(x=x, y=y, z=z, h=h, m=m, u=np.full(n, 0.05), vx=zeros.copy(), vy=zeros.copy(), vz=zeros.copy(),
                alpha=np.full(n, 0.05)), dict(lo=-radius, hi=radius, periodic=False)
