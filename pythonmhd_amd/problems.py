"""Initial states of the benchmark configurations, in the reference layout
(cells (8, ny, nx[, nz]) or (8, nx), faces faceBx/faceBy/faceBz), as float64 NumPy
arrays: (prim, cons, faces).

  brio_wu        1D Brio-Wu shock tube (gamma = 2)
  orszag_tang    2D Orszag-Tang vortex, face B from the vector potential
                 (InitializationFiles/SampleSimulations/2D_MHD/OrszagTang_InitializationFile.py:149-173)
  turbulence     3D random-phase solenoidal velocity, uniform B_x
  kh_boxes       2D Kelvin-Helmholtz boxes of uneven sizes
  adjoint_state  2D smooth density perturbation (the adjoint configuration)

Conserved variables follow Source/PrimCons_HelperModule.py:314-340:
E = p/(gamma-1) + rho v^2/2 + B^2/2.
"""
import numpy as np


def cons_of(prim, gamma):
    """Conserved variables of primitive ones (8 variables, any grid shape)."""
    cons = prim.copy()
    cons[1:4] = prim[1:4] * prim[0]
    cons[7] = prim[7] / (gamma - 1.0) + 0.5 * prim[0] * (prim[1] ** 2 + prim[2] ** 2 + prim[3] ** 2) \
        + 0.5 * (prim[4] ** 2 + prim[5] ** 2 + prim[6] ** 2)
    return cons


def brio_wu(n=512):
    """Brio-Wu on [0, 1], gamma = 2: rho = 1, p = 1, B_y = 1 for x <= 0.5;
    rho = 0.125, p = 0.1, B_y = -1 beyond; B_x = 0.75, v = 0, B_z = 0."""
    gamma = 2.0
    x = (np.arange(n) + 0.5) / n
    left = x <= 0.5
    prim = np.zeros((8, n))
    prim[0] = np.where(left, 1.0, 0.125)
    prim[7] = np.where(left, 1.0, 0.1)
    prim[4] = 0.75
    prim[5] = np.where(left, 1.0, -1.0)
    return prim, cons_of(prim, gamma), ()


def orszag_tang(n=512):
    """Orszag-Tang on [0, 1]^2 (gamma = 5/3): rho = 25/(36 pi), p = 5/(12 pi),
    v = (-sin 2 pi y, sin 2 pi x, 0), B0 = 1/sqrt(4 pi); the face field is the
    discrete curl of A_z = B0/(4 pi) cos 4 pi x + B0/(2 pi) cos 2 pi y sampled at the
    cell corners (x - dx/2, y - dy/2), periodic, and cell B is the face average."""
    gamma = 5.0 / 3.0
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    y, x = np.meshgrid(c, c, indexing='ij')
    b0 = 1.0 / np.sqrt(4.0 * np.pi)
    a = b0 / (4.0 * np.pi) * np.cos(4.0 * np.pi * (x - h / 2)) + b0 / (2.0 * np.pi) * np.cos(2.0 * np.pi * (y - h / 2))
    bx_left = (np.roll(a, -1, axis=0) - a) / h          # on the face at x - dx/2
    by_top = -(np.roll(a, -1, axis=1) - a) / h          # on the face at y - dy/2
    fbx = np.concatenate([bx_left, bx_left[:, :1]], axis=1)
    fby = np.concatenate([by_top, by_top[:1, :]], axis=0)
    prim = np.zeros((8, n, n))
    prim[0] = 25.0 / (36.0 * np.pi)
    prim[7] = 5.0 / (12.0 * np.pi)
    prim[1] = -np.sin(2.0 * np.pi * y)
    prim[2] = np.sin(2.0 * np.pi * x)
    prim[4] = 0.5 * (fbx[:, :-1] + fbx[:, 1:])
    prim[5] = 0.5 * (fby[:-1, :] + fby[1:, :])
    return prim, cons_of(prim, gamma), (fbx, fby)


def turbulence(n=256, seed=2024, kmax=8, mach=0.5, gamma=5.0 / 3.0):
    """3D box (y, x, z): rho = p = 1, uniform B_x = 0.5, and a random-phase
    solenoidal velocity with per-mode power |k|^-2 over the integer modes
    1 <= |k| <= kmax (seeded), rescaled to rms Mach `mach`."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, 1.0 / n)
    ky, kx, kz = np.meshgrid(k1, k1, k1, indexing='ij')
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    sel = (kk >= 1.0) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[sel] = kk[sel] ** -1.0          # per-mode power |k|^-2
    vh = []
    for _ in range(3):
        ph = rng.uniform(0, 2 * np.pi, kk.shape)
        g = rng.standard_normal(kk.shape)
        vh.append(amp * g * np.exp(1j * ph))
    kd = np.where(kk > 0, kk * kk, 1.0)
    kdotv = (kx * vh[0] + ky * vh[1] + kz * vh[2]) / kd
    vx = np.real(np.fft.ifftn(vh[0] - kx * kdotv))
    vy = np.real(np.fft.ifftn(vh[1] - ky * kdotv))
    vz = np.real(np.fft.ifftn(vh[2] - kz * kdotv))
    del vh, kdotv, kx, ky, kz, kk, amp
    rms = np.sqrt(np.mean(vx * vx + vy * vy + vz * vz))
    s = mach * np.sqrt(gamma) / rms
    prim = np.zeros((8, n, n, n))
    prim[0] = 1.0
    prim[1], prim[2], prim[3] = vx * s, vy * s, vz * s
    prim[4] = 0.5
    prim[7] = 1.0
    faces = (np.full((n, n + 1, n), 0.5), np.zeros((n + 1, n, n)), np.zeros((n, n, n + 1)))
    return prim, cons_of(prim, gamma), faces


def kh_boxes(nbox=256, seed=7, lo=64, hi=256, gamma=5.0 / 3.0):
    """Sides n ~ U{lo..hi} and (n, prim, cons, faces) of each Kelvin-Helmholtz box:
    rho = 1 and v_x = +0.5 for |y - 0.5| < 0.25, rho = 2 and v_x = -0.5 elsewhere,
    p = 2.5, B = (0.1, 0, 0), v_y = 0.01 * uniform white noise in [-1, 1]."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, nbox)
    out = []
    for n in sides:
        n = int(n)
        y = (np.arange(n) + 0.5) / n
        inner = (np.abs(y - 0.5) < 0.25)[:, None] * np.ones((1, n), dtype=bool)
        prim = np.zeros((8, n, n))
        prim[0] = np.where(inner, 1.0, 2.0)
        prim[1] = np.where(inner, 0.5, -0.5)
        prim[2] = 0.01 * rng.uniform(-1.0, 1.0, (n, n))
        prim[4] = 0.1
        prim[7] = 2.5
        faces = (np.full((n, n + 1), 0.1), np.zeros((n + 1, n)))
        out.append((n, prim, cons_of(prim, gamma), faces))
    return out


def grf(n, corr=0.1, seed=11):
    """Zero-mean, unit-variance periodic Gaussian random field on n^2 cells with a
    Gaussian correlation of length corr (spectral filter exp(-k^2 corr^2 / 2))."""
    rng = np.random.default_rng(seed)
    k1 = 2 * np.pi * np.fft.fftfreq(n, 1.0 / n)
    ky, kx = np.meshgrid(k1, k1, indexing='ij')
    f = np.real(np.fft.ifft2(np.fft.fft2(rng.standard_normal((n, n))) * np.exp(-0.5 * (kx * kx + ky * ky) * corr ** 2)))
    f = f - f.mean()
    return f / f.std()


def adjoint_state(n, velocity=0.0, seed=11, gamma=5.0 / 3.0):
    """rho = 1 + 0.05 * GRF, p = 1, B = (0.3, 0, 0); v = 0 (the configured entropy
    mode) or, for the gradient check, a smooth velocity field of amplitude velocity."""
    prim = np.zeros((8, n, n))
    prim[0] = 1.0 + 0.05 * grf(n, seed=seed)
    if velocity:
        prim[1] = velocity * grf(n, corr=0.15, seed=seed + 1)
        prim[2] = velocity * grf(n, corr=0.15, seed=seed + 2)
    prim[4] = 0.3
    prim[7] = 1.0
    faces = (np.full((n, n + 1), 0.3), np.zeros((n + 1, n)))
    return prim, cons_of(prim, gamma), faces
