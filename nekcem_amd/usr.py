"""User-hook library (the .usr side): initial conditions and exact solutions.

The analytic modes are those of the reference's own test cases:
3dboxper (tests/3dboxper/3dboxper.usr:45-86, periodic (1,1,1) mode, omega =
sqrt(3)) and 3dboxpec (tests/3dboxpec/3dboxpec.usr:91-113, PEC cavity mode,
omega = pi*sqrt(3)).
"""
import numpy as np


def boxper_solution(t, x, y, z):
    om = np.sqrt(3.0)
    th, te = np.sin(om * t) / om, np.cos(om * t)
    H = np.stack([2 * np.cos(x) * np.sin(y) * np.cos(z) * th,
                  -np.sin(x) * np.cos(y) * np.cos(z) * th,
                  np.sin(x) * np.sin(y) * np.sin(z) * th])
    E = np.stack([0.0 * x,
                  np.cos(x) * np.sin(y) * np.sin(z) * te,
                  np.cos(x) * np.cos(y) * np.cos(z) * te])
    return H, E


def boxpec_solution(t, x, y, z):
    ww = np.pi
    th = np.cos(ww * np.sqrt(3.0) * t) / np.sqrt(6.0)
    te = np.sin(ww * np.sqrt(3.0) * t) / np.sqrt(2.0)
    H = np.stack([-np.sin(ww * x) * np.cos(ww * y) * np.cos(ww * z) * th,
                  -np.cos(ww * x) * np.sin(ww * y) * np.cos(ww * z) * th,
                  2 * np.cos(ww * x) * np.cos(ww * y) * np.sin(ww * z) * th])
    E = np.stack([-np.cos(ww * x) * np.sin(ww * y) * np.sin(ww * z) * te,
                  np.sin(ww * x) * np.cos(ww * y) * np.sin(ww * z) * te,
                  0.0 * x])
    return H, E


def plane_waves(x, y, z, seed=2, nwaves=8, kmax=4):
    """Superposition of nwaves plane waves with integer wavevectors
    0 < |k| <= kmax (periodic on [0, 2pi]^3), unit polarisation p orthogonal
    to k, seeded phase: E = sum p cos(k.x + phi), H = sum (khat x p) cos(...)."""
    rng = np.random.default_rng(seed)
    E = np.zeros((3,) + x.shape)
    H = np.zeros((3,) + x.shape)
    n = 0
    while n < nwaves:
        k = rng.integers(-kmax, kmax + 1, size=3).astype(float)
        kn = np.linalg.norm(k)
        if kn == 0 or kn > kmax:
            continue
        khat = k / kn
        v = rng.normal(size=3)
        p = v - np.dot(v, khat) * khat
        p /= np.linalg.norm(p)
        phi = rng.uniform(0, 2 * np.pi)
        ph = np.cos(k[0] * x + k[1] * y + k[2] * z + phi)
        q = np.cross(khat, p)
        for c in range(3):
            E[c] += p[c] * ph
            H[c] += q[c] * ph
        n += 1
    return H, E


def element_eps(nelt, seed=1, lo=1.0, hi=4.0):
    return np.random.default_rng(seed).uniform(lo, hi, nelt)


def cavity_batch(ncav=256, seed=3, nmin=2, nmax=16, Nmin=5, Nmax=9):
    """The batched workload: ncav independent PEC box cavities on [0,1]^3,
    each n^3 elements with n ~ U{nmin..nmax} and polynomial order
    N ~ U{Nmin..Nmax}, per-element eps ~ U[1,4], mu = 1 (one seed)."""
    rng = np.random.default_rng(seed)
    cavs = []
    for _ in range(ncav):
        n = int(rng.integers(nmin, nmax + 1))
        N = int(rng.integers(Nmin, Nmax + 1))
        cavs.append(dict(nel=(n, n, n), N=N, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                         bc=("PEC", "PEC", "PEC"), eps_elem=rng.uniform(1.0, 4.0, n ** 3)))
    return cavs


def gaussian_ez(x, y, z, center=(0.5, 0.5, 0.5), width=0.2):
    """Initial state with Ez = exp(-|x-c|^2 / (2 w^2)), all other fields 0."""
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    E = np.zeros((3,) + x.shape)
    E[2] = np.exp(-r2 / (2.0 * width * width))
    return np.zeros_like(E), E


def smooth_field(x, y, z, seed, nmodes=4, kmax=2):
    """Smooth field on [0,1]^3: a few low-wavenumber sinusoids with seeded
    integer wavevectors and phases, scaled to max |s| = 1."""
    rng = np.random.default_rng(seed)
    s = np.zeros_like(x)
    for _ in range(nmodes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not np.any(k):
            k[0] = 1
        s += rng.normal() * np.cos(np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
    return s / np.max(np.abs(s))


def gaussian_source(x, y, z, bm1, center=(0.4, 0.55, 0.45), width=0.1):
    """bm1-weighted spatial amplitude of a Gaussian pulse source on Ez."""
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    return np.exp(-r2 / (2.0 * width * width)) * bm1
