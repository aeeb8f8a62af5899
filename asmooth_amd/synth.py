"""Seeded synthetic particle sets for the smoothing workloads.

All generators return an (n, 6) float32 array of {x, y, z, vx, vy, vz} with
positions in [0, L) - the content of a cubep3m xv file (file_reader2.cpp:64-78)
- so the same bytes feed the device path, the CPU oracle and the reference.
numpy's PCG64 streams and pocketfft are deterministic, so every machine with
this numpy regenerates identical inputs.

    blobs_uniform      64^3 particles, L=128: 70% uniform + 30% in 200 Gaussian blobs
    zeldovich          lattice displaced by a Gaussian field with P(k) ~ k^-2
    batched_cubes      64 independent periodic cubes of lognormal particle count
    precision_study    512^3 Zel'dovich in L=1024 plus 5000 Gaussian blobs
"""
from __future__ import annotations

import numpy as np

__all__ = ["blobs_uniform", "zeldovich", "batched_cubes", "precision_study", "wrap", "n_buffered_estimate"]


def wrap(x: np.ndarray, L: float) -> np.ndarray:
    """Wrap float64 positions into [0, L) and return float32 values < L."""
    y = np.mod(np.asarray(x, np.float64), L).astype(np.float32)
    y[y >= np.float32(L)] = np.float32(0.0)
    y[y < 0] = np.float32(0.0)
    return y


def blobs_uniform(n_side: int = 64, L: float = 128.0, seed: int = 0, n_blobs: int = 200,
                  blob_frac: float = 0.3, sigma_median: float = 1.5, sigma_dex: float = 0.3) -> np.ndarray:
    """n_side^3 particles in a periodic cube: (1-blob_frac) uniform, the rest in
    isotropic Gaussian blobs (centres uniform, sigma lognormal with the given
    median and dex, counts multinomial with equal weights)."""
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    n_uni = int(round((1.0 - blob_frac) * n))
    n_blob = n - n_uni
    uni = rng.uniform(0.0, L, size=(n_uni, 3))
    centres = rng.uniform(0.0, L, size=(n_blobs, 3))
    sigma = sigma_median * 10.0 ** (sigma_dex * rng.standard_normal(n_blobs))
    counts = rng.multinomial(n_blob, np.full(n_blobs, 1.0 / n_blobs))
    owner = np.repeat(np.arange(n_blobs), counts)
    blob = centres[owner] + sigma[owner, None] * rng.standard_normal((n_blob, 3))
    pos = wrap(np.concatenate([uni, blob]), L)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([pos, vel], axis=1), dtype=np.float32)


def _zeldovich_psi(n: int, L: float, rng: np.random.Generator, pk_index: float = -2.0):
    """Displacement field psi (3 arrays (n,n,n) in (z, y, x) order) for a
    Gaussian density field with P(k) ~ k^pk_index: psi_hat = i k / k^2 delta_hat."""
    w = rng.standard_normal((n, n, n))
    W = np.fft.rfftn(w)
    del w
    kf = 2.0 * np.pi / L
    kz = (np.fft.fftfreq(n, d=1.0 / n) * kf)[:, None, None]
    ky = (np.fft.fftfreq(n, d=1.0 / n) * kf)[None, :, None]
    kx = (np.fft.rfftfreq(n, d=1.0 / n) * kf)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = k2 ** (pk_index / 4.0)          # sqrt(P) = k^(pk_index/2)
    base = W * (amp / k2)
    base[0, 0, 0] = 0.0
    del W, amp
    psi = []
    for kc in (kx, ky, kz):
        psi.append(np.fft.irfftn(1j * kc * base, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32))
    return psi


def zeldovich(n: int = 256, L: float = 512.0, rms: float = 2.5, seed: int = 1,
              pk_index: float = -2.0) -> np.ndarray:
    """n^3 particles on a lattice of spacing L/n (cell centres), displaced by
    the Zel'dovich field scaled to an rms |psi| of `rms`, wrapped periodically.
    Velocities are the displacement (Zel'dovich: v is proportional to psi).
    Particle order is the lattice order with x fastest."""
    rng = np.random.default_rng(seed)
    px, py, pz = _zeldovich_psi(n, L, rng, pk_index)
    s = np.sqrt(np.mean(px.astype(np.float64) ** 2 + py.astype(np.float64) ** 2 + pz.astype(np.float64) ** 2))
    f = np.float32(rms / s) if s > 0 else np.float32(0.0)
    px *= f
    py *= f
    pz *= f
    d = L / n
    q = (np.arange(n, dtype=np.float64) + 0.5) * d
    out = np.empty((n ** 3, 6), np.float32)
    out[:, 0] = wrap(q[None, None, :] + px, L).reshape(-1)
    out[:, 1] = wrap(q[None, :, None] + py, L).reshape(-1)
    out[:, 2] = wrap(q[:, None, None] + pz, L).reshape(-1)
    out[:, 3] = px.reshape(-1)
    out[:, 4] = py.reshape(-1)
    out[:, 5] = pz.reshape(-1)
    return out


def batched_cubes(n_cubes: int = 64, L: float = 128.0, seed0: int = 1000, log2_median: float = 18.0,
                  lo: int = 2 ** 15, hi: int = 2 ** 21):
    """Independent periodic cubes: cube i (seed seed0+i) has N_i = n_i^3 with
    ln N_i ~ Normal(ln 2^log2_median, (ln 2)^2) clipped to [lo, hi], and a
    Zel'dovich displacement of rms sigma_i ~ U[1, 4] on its own lattice."""
    cubes = []
    for i in range(n_cubes):
        rng = np.random.default_rng(seed0 + i)
        lnN = rng.normal(log2_median * np.log(2.0), np.log(2.0))
        N = float(np.clip(np.exp(lnN), lo, hi))
        ni = max(4, int(round(N ** (1.0 / 3.0))))
        sigma = float(rng.uniform(1.0, 4.0))
        px, py, pz = _zeldovich_psi(ni, L, rng)
        s = np.sqrt(np.mean(px.astype(np.float64) ** 2 + py ** 2 + pz ** 2))
        f = np.float32(sigma / s)
        px *= f
        py *= f
        pz *= f
        d = L / ni
        q = (np.arange(ni, dtype=np.float64) + 0.5) * d
        xv = np.empty((ni ** 3, 6), np.float32)
        xv[:, 0] = wrap(q[None, None, :] + px, L).reshape(-1)
        xv[:, 1] = wrap(q[None, :, None] + py, L).reshape(-1)
        xv[:, 2] = wrap(q[:, None, None] + pz, L).reshape(-1)
        xv[:, 3] = px.reshape(-1)
        xv[:, 4] = py.reshape(-1)
        xv[:, 5] = pz.reshape(-1)
        cubes.append(xv)
    return cubes


def precision_study(n: int = 512, L: float = 1024.0, seed: int = 2, n_blobs: int = 5000,
                    move_frac: float = 0.1) -> np.ndarray:
    """Zel'dovich lattice (rms 2.5) with move_frac of the particles moved into
    n_blobs Gaussian blobs of sigma log-uniform in [0.3, 3]."""
    xv = zeldovich(n, L, 2.5, seed)
    rng = np.random.default_rng(seed + 7919)
    m = int(move_frac * len(xv))
    idx = rng.choice(len(xv), size=m, replace=False)
    centres = rng.uniform(0.0, L, size=(n_blobs, 3))
    sigma = np.exp(rng.uniform(np.log(0.3), np.log(3.0), n_blobs))
    owner = rng.integers(0, n_blobs, size=m)
    pos = centres[owner] + sigma[owner, None] * rng.standard_normal((m, 3))
    xv[idx, :3] = wrap(pos, L)
    return xv


def n_buffered_estimate(n_local: int, bf: float) -> int:
    """Expected buffered count for a uniform node share (exchange_buffer.cpp:166-167)."""
    return int(n_local * (1.0 + 2.0 * bf) ** 3)
