"""Synthetic scientific fields for the benchmark workloads.

Random numbers come from SplitMix64-seeded xoshiro256** streams; normals use
Box-Muller.  Parameters are drawn on the host; the fields themselves are
evaluated on the device in float64 and cast to the storage type.

    smooth_mix    64^3-style field: 8 random-phase plane waves (k 1..6) plus
                  4 Gaussian blobs (sigma 4..10 cells) plus N(0, 1e-3)
    density       Miranda-like density: 1 + 0.5 tanh(sum of 8 plane waves,
                  k 1..8) plus N(0, 1e-4)
    grf           Gaussian random field with power ~ |k|^-5/3 (amplitude
                  |k|^-5/6), random phases, unit variance
    modes         grf-like spectrum by direct synthesis of 64 random modes
                  (any shape, cheap), unit variance
"""
from __future__ import annotations

import math

import numpy as np
import torch

M64 = (1 << 64) - 1


def splitmix64(x: int):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


class Xoshiro:
    """Scalar xoshiro256** for parameters."""

    def __init__(self, seed: int):
        s, self.s = seed & M64, []
        for _ in range(4):
            s, v = splitmix64(s)
            self.s.append(v)

    def next(self) -> int:
        s0, s1, s2, s3 = self.s
        r = (((s1 * 5) & M64) << 7 | ((s1 * 5) & M64) >> 57) & M64
        r = (r * 9) & M64
        t = (s1 << 17) & M64
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        s3 = ((s3 << 45) | (s3 >> 19)) & M64
        self.s = [s0, s1, s2, s3]
        return r

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:  # inclusive
        return lo + int(self.uniform() * (hi - lo + 1))


def _rotl(x, k):
    return (x << np.uint64(k)) | (x >> np.uint64(64 - k))


def normals(n: int, seed: int, streams: int = 1 << 16) -> np.ndarray:
    """n standard normals from `streams` independent xoshiro256** streams
    (stream i seeded by SplitMix64 from seed*2^20 + i), Box-Muller."""
    streams = max(1, min(streams, (n + 1) // 2))
    st = np.empty((4, streams), np.uint64)
    for i in range(streams):
        s = (seed * (1 << 20) + i) & M64
        for k in range(4):
            s, v = splitmix64(s)
            st[k, i] = v
    steps = (n + 2 * streams - 1) // (2 * streams)
    out = np.empty((steps, 2, streams), np.float64)
    with np.errstate(over="ignore"):
        for t in range(steps):
            u = []
            for _ in range(2):
                s0, s1, s2, s3 = st
                r = _rotl(s1 * np.uint64(5), 7) * np.uint64(9)
                tt = s1 << np.uint64(17)
                s2 ^= s0
                s3 ^= s1
                s1 ^= s2
                s0 ^= s3
                s2 ^= tt
                st[3] = _rotl(s3, 45)
                u.append((r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53)))
            r1 = np.sqrt(-2.0 * np.log1p(-u[0]))
            out[t, 0] = r1 * np.cos(2 * math.pi * u[1])
            out[t, 1] = r1 * np.sin(2 * math.pi * u[1])
    return out.reshape(-1)[:n]


def _grid(nx, ny, nz, dev):
    z = torch.arange(nz, dtype=torch.float64, device=dev).view(nz, 1, 1)
    y = torch.arange(ny, dtype=torch.float64, device=dev).view(1, ny, 1)
    x = torch.arange(nx, dtype=torch.float64, device=dev).view(1, 1, nx)
    return x, y, z


def _noise(shape, sigma, seed, dev):
    n = int(np.prod(shape))
    return torch.from_numpy(normals(n, seed).reshape(shape)).to(dev) * sigma


def smooth_mix(nx=64, ny=64, nz=64, seed=0, dtype=torch.float32, device="cuda"):
    g = Xoshiro(seed)
    x, y, z = _grid(nx, ny, nz, device)
    f = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for _ in range(8):
        kx, ky, kz = g.randint(1, 6), g.randint(1, 6), g.randint(1, 6)
        ph = 2 * math.pi * g.uniform()
        f += torch.sin(2 * math.pi * (kx * x / nx + ky * y / ny + kz * z / nz) + ph)
    for _ in range(4):
        a = 0.5 + 1.5 * g.uniform()
        c = (g.uniform() * nx, g.uniform() * ny, g.uniform() * nz)
        s = 4 + 6 * g.uniform()
        f += a * torch.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    f += _noise((nz, ny, nx), 1e-3, seed, device)
    return f.to(dtype)


def density(nx=384, ny=384, nz=256, seed=2, dtype=torch.float64, device="cuda"):
    g = Xoshiro(seed)
    x, y, z = _grid(nx, ny, nz, device)
    s = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for _ in range(8):
        kx, ky, kz = g.randint(1, 8), g.randint(1, 8), g.randint(1, 8)
        ph = 2 * math.pi * g.uniform()
        s += torch.sin(2 * math.pi * (kx * x / nx + ky * y / ny + kz * z / nz) + ph)
    f = 1 + 0.5 * torch.tanh(s)
    del s
    f += _noise((nz, ny, nx), 1e-4, seed, device)
    return f.to(dtype)


def grf(nx=512, ny=512, nz=512, seed=1, dtype=torch.float32, device="cuda"):
    """Gaussian random field, power ~ |k|^-5/3, by an inverse real FFT."""
    kz = torch.fft.fftfreq(nz, device=device, dtype=torch.float64).view(nz, 1, 1) * nz
    ky = torch.fft.fftfreq(ny, device=device, dtype=torch.float64).view(1, ny, 1) * ny
    kx = torch.fft.rfftfreq(nx, device=device, dtype=torch.float64).view(1, 1, nx // 2 + 1) * nx
    k = torch.sqrt(kx * kx + ky * ky + kz * kz)
    amp = torch.where(k > 0, k.clamp_min(1e-30) ** (-5.0 / 6.0), torch.zeros_like(k))
    del k
    nph = amp.numel()
    rng = np.random.Generator(np.random.PCG64(seed))  # phases only; the field is then normalised
    ph = torch.from_numpy(rng.random(nph) * 2 * math.pi).to(device).view(amp.shape)
    spec = torch.polar(amp, ph)
    del amp, ph
    f = torch.fft.irfftn(spec, s=(nz, ny, nx))
    del spec
    f = (f - f.mean()) / f.std()
    return f.to(dtype)


def modes(nx, ny, nz, seed, dtype=torch.float32, device="cuda", count=64):
    """Spectrum ~|k|^-5/6 by direct synthesis of `count` random modes."""
    g = Xoshiro(seed)
    x, y, z = _grid(nx, ny, nz, device)
    f = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for _ in range(count):
        k = [g.randint(-8, 8) if d > 1 else 0 for d in (nx, ny, nz)]
        if k == [0, 0, 0]:
            k[0] = 1
        kk = math.sqrt(sum(v * v for v in k))
        ph = 2 * math.pi * g.uniform()
        f += kk ** (-5.0 / 6.0) * torch.cos(2 * math.pi * (k[0] * x / nx + k[1] * y / ny + k[2] * z / nz) + ph)
    f = (f - f.mean()) / f.std()
    return f.to(dtype)


def batch_spec(i: int, seed: int, max3d: int = 256, max2d: int = 2048, max1d: float = 1e7):
    """Shape, dtype, generator and relative bound of field i of the mixed
    batch: float32 for even i; by i mod 4, 3-D (axes U{32..max3d}), 2-D (axes
    U{64..max2d}) or 1-D (n log-uniform in [1e5, max1d]); the density
    generator for i mod 8 < 4, else the mode spectrum; rel eb log-uniform in
    [1e-6, 1e-2].  The draws come from the field's own seed."""
    g = Xoshiro(seed)
    dtype = torch.float32 if i % 2 == 0 else torch.float64
    cls = i % 4
    if cls <= 1:
        shape = (g.randint(32, max3d), g.randint(32, max3d), g.randint(32, max3d))
    elif cls == 2:
        shape = (g.randint(64, max2d), g.randint(64, max2d), 1)
    else:
        shape = (int(round(10 ** (5 + g.uniform() * (math.log10(max1d) - 5)))), 1, 1)
    rel_eb = 10 ** (-6 + 4 * g.uniform())
    gen = "density" if i % 8 < 4 else "modes"
    return shape, dtype, gen, rel_eb


def batch_field(i: int, seed: int, device="cuda", **limits):
    """(tensor [nz][ny][nx], rel_eb) of field i of the mixed batch."""
    (nx, ny, nz), dtype, gen, rel_eb = batch_spec(i, seed, **limits)
    if gen == "density":
        t = density(nx, ny, nz, seed=seed, dtype=dtype, device=device)
    else:
        t = modes(nx, ny, nz, seed=seed, dtype=dtype, device=device)
    return t, rel_eb
