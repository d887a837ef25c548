"""Synthetic scenes of the benchmark configurations.

Generator (a build choice; the reference ships no scenes):
  * seeds: SplitMix64 of base 20261202 xor (scene index * golden ratio);
  * terrain: value noise, lattice 4 * 2^o per octave, smoothstep bilinear,
    persistence 0.5, normalised to [0, 1] (ridged: (1 - |2v - 1|)^2 per
    octave); height = base + amplitude * noise at cell centres, FP64 then FP32;
  * particles: jittered cubic lattice of spacing s, jitter U[-0.1 s, 0.1 s]
    per axis, filled x fastest, then z, then y layers, starting one spacing
    above the highest terrain under the block; ids in fill order;
  * h = 2.15 s (32 lattice neighbours); mass m = rho0 / sum over the lattice
    of the poly6 kernel, so the unjittered interior density is rho0;
  * unit box [0, 1]^3, rho0 = 1000, k = 3, visc = 3.5, surf_tens = 0.0728,
    g = -9.8; erosion kc = 0.05, ke = kd = 0.3, contact band h.
Positions are rounded to FP32 so FP32 and FP64 runs start from one state.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import EROSION_ATOMIC
from .core import default_params

BASE_SEED = 20261202
GOLDEN = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1
PI_F = float(np.float32(3.141592))  # Erosion/sphere.h:8


def splitmix64(x):
    x = (x + GOLDEN) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def scene_seed(idx, base=BASE_SEED):
    return splitmix64(base ^ ((idx * GOLDEN) & MASK))


def value_noise(T, octaves, rng, ridged=False):
    """T x T cell-centred noise in [0, 1] (row = z)."""
    c = (np.arange(T) + 0.5) / T
    total = np.zeros((T, T))
    wsum = 0.0
    w = 1.0
    for o in range(octaves):
        L = 4 * (2 ** o)
        lat = rng.random((L + 1, L + 1))
        u = c * L
        i0 = np.minimum(np.floor(u).astype(np.int64), L - 1)
        f = u - i0
        f = f * f * (3.0 - 2.0 * f)  # smoothstep
        # rows = z, cols = x
        a = lat[np.ix_(i0, i0)]
        b = lat[np.ix_(i0, i0 + 1)]
        cc = lat[np.ix_(i0 + 1, i0)]
        d = lat[np.ix_(i0 + 1, i0 + 1)]
        fx = f[None, :]
        fz = f[:, None]
        v = (1 - fz) * ((1 - fx) * a + fx * b) + fz * ((1 - fx) * cc + fx * d)
        if ridged:
            v = (1.0 - np.abs(2.0 * v - 1.0)) ** 2
        total += w * v
        wsum += w
        w *= 0.5
    total /= wsum
    lo, hi = total.min(), total.max()
    return (total - lo) / max(hi - lo, 1e-300)


def poly6_lattice_sum(s, h):
    """Sum of the poly6 kernel over a cubic lattice of spacing s (self included)."""
    r = int(math.ceil(h / s))
    g = np.arange(-r, r + 1) * s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    r2 = X * X + Y * Y + Z * Z
    h2 = h * h
    q = np.where(r2 <= h2, h2 - r2, 0.0)
    return float(np.sum(q ** 3) * 315.0 / (64.0 * PI_F * h ** 9))


@dataclass
class Scene:
    name: str
    params: object
    terrain: dict
    pos: np.ndarray
    vel: np.ndarray
    ids: np.ndarray
    steps: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.pos)


def make_terrain(T, octaves, amp, base, rng, ridged=False):
    n = value_noise(T, octaves, rng, ridged)
    hts = (base + amp * n).astype(np.float32)
    return dict(nx=T, nz=T, origin=(0.0, 0.0, 0.0), cell=1.0 / T, heights=hts)


def terrain_max_under(terrain, x0, x1, z0, z1):
    T = terrain["nx"]
    i0, i1 = max(0, int(x0 * T) - 1), min(T, int(math.ceil(x1 * T)) + 1)
    k0, k1 = max(0, int(z0 * T) - 1), min(T, int(math.ceil(z1 * T)) + 1)
    return float(terrain["heights"][k0:k1, i0:i1].max())


def lattice_block(n, s, x0, x1, z0, z1, y0, rng):
    """n particles of a jittered lattice: x fastest, then z, then y layers."""
    nxl = max(1, int(math.floor((x1 - x0) / s + 1e-9)))
    nzl = max(1, int(math.floor((z1 - z0) / s + 1e-9)))
    per = nxl * nzl
    layers = (n + per - 1) // per
    k = np.arange(n)
    ix = k % nxl
    iz = (k // nxl) % nzl
    iy = k // per
    pos = np.empty((n, 3))
    pos[:, 0] = x0 + (ix + 0.5) * s
    pos[:, 2] = z0 + (iz + 0.5) * s
    pos[:, 1] = y0 + iy * s
    pos += rng.uniform(-0.1 * s, 0.1 * s, size=(n, 3))
    return pos.astype(np.float32).astype(np.float64), dict(per_layer=per, layers=layers)


def sph_params(h, s, dt, erosion=EROSION_ATOMIC, rho0=1000.0):
    m = rho0 / poly6_lattice_sum(s, h)
    return default_params(h=h, mass=m, rest_density=rho0, stiffness=3.0, visc=3.5, surf_tens=0.0728,
                          gravity=(0.0, -9.8, 0.0), dt=dt, box_center=(0.5, 0.5, 0.5), box_half=0.5,
                          erosion=erosion, kc=0.05, ke=0.3, kd=0.3, band=h)


def cfg1(seed_idx=0, erosion=EROSION_ATOMIC, n=4096, steps=200):
    """Dam break: 4,096 particles in x < 0.3, 64^2 terrain, 4 octaves, amp 0.2, base 0.1,
    h = 0.0457, dt = 1e-3, 200 steps."""
    rng = np.random.default_rng(scene_seed(seed_idx))
    h = 0.0457
    s = h / 2.15
    ter = make_terrain(64, 4, 0.2, 0.1, rng)
    y0 = terrain_max_under(ter, 0.0, 0.3, 0.0, 1.0) + s
    pos, meta = lattice_block(n, s, 0.0, 0.3, 0.0, 1.0, y0, rng)
    prm = sph_params(h, s, 1e-3, erosion)
    return Scene("cfg1_dam_break_4096", prm, ter, pos, np.zeros_like(pos), np.arange(n, dtype=np.int32),
                 steps, dict(meta, s=s, h=h))


def cfg2(seed_idx=1, erosion=EROSION_ATOMIC, side=64, steps=500, T=512):
    """Column of side^3 = 262,144 particles (s = 1/256) on a 512^2 terrain, 6 octaves,
    amp 0.3, base 0.1, dropped from 0.05 above the terrain, dt = 5e-4, 500 steps."""
    rng = np.random.default_rng(scene_seed(seed_idx))
    s = 1.0 / 256.0
    h = 2.15 * s
    ter = make_terrain(T, 6, 0.3, 0.1, rng)
    w = side * s
    x0 = 0.5 - w / 2
    y0 = terrain_max_under(ter, x0, x0 + w, x0, x0 + w) + 0.05
    n = side ** 3
    pos, meta = lattice_block(n, s, x0, x0 + w, x0, x0 + w, y0, rng)
    prm = sph_params(h, s, 5e-4, erosion)
    return Scene(f"cfg2_column_{n}", prm, ter, pos, np.zeros_like(pos), np.arange(n, dtype=np.int32),
                 steps, dict(meta, s=s, h=h))


def dam_break_scene(idx, n, frac, T, dt=5e-4, octaves=4, amp=0.2, base=0.1, erosion=EROSION_ATOMIC,
                    steps=300):
    """One batched scene: n particles filling x < frac, 0.3 high."""
    rng = np.random.default_rng(scene_seed(idx))
    s = (frac * 0.3 / n) ** (1.0 / 3.0)
    h = 2.15 * s
    ter = make_terrain(T, octaves, amp, base, rng)
    y0 = terrain_max_under(ter, 0.0, frac, 0.0, 1.0) + s
    pos, meta = lattice_block(n, s, 0.0, frac, 0.0, 1.0, y0, rng)
    prm = sph_params(h, s, dt, erosion)
    return Scene(f"scene{idx}_{n}", prm, ter, pos, np.zeros_like(pos), np.arange(n, dtype=np.int32),
                 steps, dict(meta, s=s, h=h, frac=frac))


def cfg3(count=256, erosion=EROSION_ATOMIC, steps=300, base_idx=1000):
    """256 scenes, N log-uniform in [2048, 65536], T from {64,128,192,256}, frac in [0.2, 0.5]."""
    rng = np.random.default_rng(scene_seed(base_idx))
    out = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(2048), math.log(65536)))))
        frac = float(rng.uniform(0.2, 0.5))
        T = int(rng.choice([64, 128, 192, 256]))
        out.append(dam_break_scene(base_idx + 1 + i, n, frac, T, erosion=erosion, steps=steps))
    return out


def cfg4(seed_idx=4, erosion=EROSION_ATOMIC, n=4194304, steps=1000, T=2048):
    """Full-footprint layer 0.1 thick of 4,194,304 particles over a ridged 2048^2 terrain,
    amp 0.5, base 0.05, dt = 5e-4."""
    rng = np.random.default_rng(scene_seed(seed_idx))
    s = (0.1 / n) ** (1.0 / 3.0)
    h = 2.15 * s
    ter = make_terrain(T, 6, 0.5, 0.05, rng, ridged=True)
    y0 = float(ter["heights"].max()) + s
    pos, meta = lattice_block(n, s, 0.0, 1.0, 0.0, 1.0, y0, rng)
    prm = sph_params(h, s, 5e-4, erosion)
    return Scene(f"cfg4_layer_{n}", prm, ter, pos, np.zeros_like(pos), np.arange(n, dtype=np.int32),
                 steps, dict(meta, s=s, h=h))


def cfg5(count=1024, erosion=EROSION_ATOMIC, big_n=8388608, steps=2000, big_T=4096, base_idx=5000):
    """1,024 scenes with N log-uniform in [1024, 16384] plus one 8,388,608-particle scene on a
    4096^2 terrain (cfg2 noise)."""
    rng = np.random.default_rng(scene_seed(base_idx))
    out = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(1024), math.log(16384)))))
        frac = float(rng.uniform(0.2, 0.5))
        T = int(rng.choice([64, 128, 192, 256]))
        out.append(dam_break_scene(base_idx + 1 + i, n, frac, T, erosion=erosion, steps=steps))
    rng2 = np.random.default_rng(scene_seed(base_idx))
    s = (0.1 / big_n) ** (1.0 / 3.0)
    h = 2.15 * s
    ter = make_terrain(big_T, 6, 0.3, 0.1, rng2)
    y0 = float(ter["heights"].max()) + s
    pos, meta = lattice_block(big_n, s, 0.0, 1.0, 0.0, 1.0, y0, rng2)
    out.append(Scene(f"cfg5_big_{big_n}", sph_params(h, s, 5e-4, erosion), ter, pos, np.zeros_like(pos),
                     np.arange(big_n, dtype=np.int32), steps, dict(meta, s=s, h=h)))
    return out


def cfg5_share(gpus=8, erosion=EROSION_ATOMIC, big_n=8388608, steps=2000, big_T=4096, base_idx=5000):
    """One GPU's share of cfg5's 8,388,608-particle scene split into `gpus` x-slabs: the
    particles of the first slab (x < 1/gpus) plus a ghost layer one cell (h (1 + 2^-8)) wide
    beyond it, over the matching columns of the same 4096^2 terrain (generated from the same
    seed as cfg5's; the slab's columns plus two)."""
    rng2 = np.random.default_rng(scene_seed(base_idx))
    s = (0.1 / big_n) ** (1.0 / 3.0)
    h = 2.15 * s
    ter = make_terrain(big_T, 6, 0.3, 0.1, rng2)
    y0 = float(ter["heights"].max()) + s
    pos, meta = lattice_block(big_n, s, 0.0, 1.0, 0.0, 1.0, y0, rng2)
    edge = 1.0 / gpus + h * (1.0 + 2.0 ** -8)
    keep = pos[:, 0] < edge
    pos = pos[keep]
    cols = min(big_T, int(np.ceil(edge * big_T)) + 2)
    ter = dict(ter, nx=cols, heights=np.ascontiguousarray(ter["heights"][:, :cols]))
    n = len(pos)
    return Scene(f"cfg5_share_{n}", sph_params(h, s, 5e-4, erosion), ter, pos, np.zeros_like(pos),
                 np.arange(n, dtype=np.int32), steps, dict(meta, s=s, h=h, slab=1.0 / gpus))


def reference_lattice_scene(n=1000, dt=0.01):
    """The reference app's scene (Erosion/main.cpp:170-183): Initialize(n) at the origin with
    the default parameters and dt = 0.01 (main.cpp:481), no terrain."""
    return default_params(dt=dt)
