"""Seeded synthetic initial conditions of the benchmark workloads.

The reference has no seeded Fourier initial condition (inicond_ACM.f90 sets
analytic flows only); these fields are the inputs the workloads are defined
on and are injected the way the reference's read_from_files=1 does, straight
into hvy_block.

Fields are built on the uniform periodic grid of the finest level with an
inverse real FFT (exactly the sum of the chosen modes at the grid points) and
sampled into blocks; a block at a coarser level takes every 2^(J_fine-J)-th
point, since its points coincide with fine-grid points.

u_rms is the per-component rms, sqrt(<|u|^2>/dim), over the grid.
"""
from __future__ import annotations

import numpy as np


def _half_modes(dim, kmin, kmax):
    """One wavevector of each +-k pair with kmin <= |k| <= kmax, kx >= 0."""
    r = range(-kmax, kmax + 1)
    out = []
    for kx in range(0, kmax + 1):
        for ky in r:
            for kz in (r if dim == 3 else (0,)):
                n2 = kx * kx + ky * ky + kz * kz
                if not (kmin * kmin <= n2 <= kmax * kmax):
                    continue
                if kx == 0 and (ky, kz) < (0, 0) or (kx == 0 and ky == 0 and kz < 0):
                    continue
                if kx == 0 and ky < 0:
                    continue
                out.append((kx, ky, kz))
    return np.array(out, np.int64)


def _synth(dim, N, ks, coef):
    """Real field sum_k |coef| cos(2 pi k.n/N + arg coef) on an N^dim grid,
    axes ordered (z, y, x)."""
    shape = (N,) * dim
    F = np.zeros(shape[:-1] + (N // 2 + 1,), np.complex128)
    for (kx, ky, kz), c in zip(ks, coef):
        c = 0.5 * c
        idx = ((kz % N, ky % N, kx) if dim == 3 else (ky % N, kx))
        F[idx] += c
        if kx == 0:
            idx2 = (((-kz) % N, (-ky) % N, 0) if dim == 3 else ((-ky) % N, 0))
            F[idx2] += np.conj(c)
    return np.fft.irfftn(F, s=shape, axes=tuple(range(dim))) * float(N ** dim)


def fourier_velocity_grid(dim, N, seed, kmax=8, kmin=1, amp_exponent=-2.0,
                          energy_exponent=None, solenoidal=False, urms=1.0):
    """Velocity components on the N^dim grid (list of dim arrays, axes z,y,x).

    Mode amplitude |k|^amp_exponent, or with energy_exponent e an energy
    envelope E(k) ~ k^e (per-mode amplitude sqrt(E(k)/k^(dim-1))).  With
    solenoidal=True each mode's amplitude vector is projected normal to k."""
    rng = np.random.default_rng(seed)
    ks = _half_modes(dim, kmin, kmax)
    kn = np.linalg.norm(ks, axis=1)
    if energy_exponent is not None:
        amp = np.sqrt(kn ** energy_exponent / kn ** (dim - 1))
    else:
        amp = kn ** amp_exponent
    vec = rng.normal(size=(len(ks), dim))
    if solenoidal:
        kk = ks[:, :dim] / kn[:, None]
        vec = vec - (vec * kk).sum(axis=1)[:, None] * kk
    vec /= np.maximum(np.linalg.norm(vec, axis=1)[:, None], 1e-300)
    vec *= amp[:, None]
    phase = np.exp(1j * rng.uniform(0, 2 * np.pi, size=len(ks)))
    comps = [_synth(dim, N, ks, vec[:, d] * phase) for d in range(dim)]
    ms = sum(np.mean(c * c) for c in comps) / dim
    s = urms / np.sqrt(ms)
    return [c * s for c in comps]


def sample_blocks(forest, grid, level_fine):
    """Sample a global periodic grid (axes z,y,x; N = 2^level_fine * Bs per
    direction) into every block incl. ghosts: (n, nz, ny, nx)."""
    dim, g, bs = forest.dim, forest.g, forest.bs
    N = grid.shape[-1]
    nx = bs[0] + 2 * g
    ny = bs[1] + 2 * g
    nz = bs[2] + 2 * g if dim == 3 else 1
    out = np.empty((forest.n, nz, ny, nx))
    for k in range(forest.n):
        lev = forest.levels[k] if hasattr(forest, "levels") else forest.level
        st = 2 ** (level_fine - lev)
        ijk = forest.ijk[k]
        ix = ((ijk[0] * bs[0] + np.arange(nx) - g) * st) % N
        iy = ((ijk[1] * bs[1] + np.arange(ny) - g) * st) % N
        if dim == 2:
            out[k, 0] = grid[np.ix_(iy, ix)]
        else:
            iz = ((ijk[2] * bs[2] + np.arange(nz) - g) * st) % N
            out[k] = grid[np.ix_(iz, iy, ix)]
    return out


def gaussian_pulse(forest, seed, sigma=0.05, amp=0.1):
    """p = amp * exp(-r^2/(2 sigma^2)) around a seeded centre; r is the
    periodic (minimum image) distance.  Returns (n, nz, ny, nx)."""
    rng = np.random.default_rng(seed + 7919)
    L, dim, g, bs = forest.domain, forest.dim, forest.g, forest.bs
    ctr = [rng.uniform(0, L[d]) for d in range(dim)]
    nx = bs[0] + 2 * g
    ny = bs[1] + 2 * g
    nz = bs[2] + 2 * g if dim == 3 else 1
    out = np.empty((forest.n, nz, ny, nx))
    for k in range(forest.n):
        x0, dx = forest.x0dx[k, :3], forest.x0dx[k, 3:]
        ax = [x0[0] + (np.arange(nx) - g) * dx[0], x0[1] + (np.arange(ny) - g) * dx[1],
              (x0[2] + (np.arange(nz) - g) * dx[2]) if dim == 3 else np.zeros(1)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        r2 = np.zeros_like(X)
        for d, A in enumerate((X, Y, Z)[:dim]):
            dd = np.abs(A - ctr[d]) % L[d]
            dd = np.minimum(dd, L[d] - dd)
            r2 += dd * dd
        out[k] = amp * np.exp(-r2 / (2 * sigma * sigma))
    return out


def acm_state(forest, seed, level_fine=None, pressure=True, **kw):
    """(n, dim+1, nz, ny, nx): seeded Fourier velocity + Gaussian pressure
    pulse (or p = 0)."""
    lf = forest.level if level_fine is None else level_fine
    N = 2 ** lf * forest.bs[0]
    comps = fourier_velocity_grid(forest.dim, N, seed, **kw)
    u = np.stack([sample_blocks(forest, c, lf) for c in comps], axis=1)
    if pressure:
        p = gaussian_pulse(forest, seed)[:, None]
    else:
        p = np.zeros_like(u[:, :1])
    return np.concatenate([u, p], axis=1)
