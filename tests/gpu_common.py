"""Shared helpers of the GPU parity tests: run a scene through libsphe and
through the oracle from the same input state."""
from __future__ import annotations

import numpy as np

import oracle as O
import sphe
from sphe import scenes


def orc_scene_from(params, terrain=None, fx_exp=60):
    """Oracle scene description with the same parameter values as a sphe scene."""
    kw = dict(h=params.h, mass=params.mass, rho0=params.rest_density, k=params.stiffness,
              visc=params.visc, surf_tens=params.surf_tens, g=tuple(params.gravity), dt=params.dt,
              box_c=tuple(params.box_center), box_e=params.box_half, kc=params.kc, ke=params.ke,
              kd=params.kd, band=params.band)
    if terrain is not None:
        kw.update(tnx=terrain["nx"], tnz=terrain["nz"], t_org=tuple(terrain["origin"]),
                  t_cell=float(np.float32(terrain["cell"])), fx_exp=fx_exp,
                  erosion=1 if params.erosion else 0)
    return O.make_scene(**kw)


def gpu_context(sc, precision=sphe.FP32, graph=True):
    ctx = sphe.Context(precision)
    sid = ctx.add_synthetic(sc)
    ctx.use_graph(graph)
    return ctx, sid


def oracle_inputs(ctx, sid, sc):
    """Initial state as the oracle takes it (terrain volumes and heights read
    back from the library so both start from identical bits)."""
    pos = np.asarray(sc.pos, dtype=np.float64)
    vel = np.asarray(sc.vel, dtype=np.float64)
    sed = np.zeros(len(pos), dtype=np.int64)
    vol = height = None
    fx = 60
    if sc.terrain is not None:
        vol = ctx.terrain_volume_fx(sid)
        height = ctx.terrain_heights(sid)
        fx = ctx.fixed_exponent(sid)
    return pos, vel, sed, vol, height, fx


def small_dam(n=4096, erosion=sphe.EROSION_ATOMIC, seed=0):
    return scenes.cfg1(seed_idx=seed, erosion=erosion, n=n)


def bilinear_heights(H, cell, x, z):
    """Bilinear height of cell-centred heights H[z, x] at (x, z) (FP64; for
    counting contacts only, not a parity reference)."""
    nz, nx = H.shape
    u = np.asarray(x) / cell - 0.5
    w = np.asarray(z) / cell - 0.5
    i0 = np.clip(np.floor(u).astype(np.int64), 0, nx - 2)
    k0 = np.clip(np.floor(w).astype(np.int64), 0, nz - 2)
    fu = np.clip(u - i0, 0.0, 1.0)
    fw = np.clip(w - k0, 0.0, 1.0)
    H = np.asarray(H, dtype=np.float64)
    a0 = (1 - fu) * H[k0, i0] + fu * H[k0, i0 + 1]
    a1 = (1 - fu) * H[k0 + 1, i0] + fu * H[k0 + 1, i0 + 1]
    return (1 - fw) * a0 + fw * a1


def contact_block(seed=77, erosion=sphe.EROSION_ATOMIC, n=4096):
    """A cfg1-size block in contact with the terrain: 4 layers of a jittered
    32 x 32 lattice (h = 0.0457, s = h / 2.15) following the 64^2 value-noise
    terrain, each column lifted by 0 .. 5.5 mm, moving down at 2 m/s with a
    random horizontal drift.  Every step some particles hit the terrain and
    about three quarters are inside the erosion band."""
    rng = np.random.default_rng(scenes.scene_seed(seed))
    h = 0.0457
    s = h / 2.15
    ter = scenes.make_terrain(64, 4, 0.2, 0.1, rng)
    cols = 32 * 32
    k = np.arange(n)
    ix, iz, iy = k % 32, (k // 32) % 32, k // cols
    x = 0.15 + (ix + 0.5) * s + rng.uniform(-0.1 * s, 0.1 * s, n)
    z = 0.15 + (iz + 0.5) * s + rng.uniform(-0.1 * s, 0.1 * s, n)
    lift = rng.uniform(0.0, 0.0055, n)[k % cols]
    y = bilinear_heights(ter["heights"], 1.0 / 64, x, z) + lift + iy * s
    pos = np.stack([x, y, z], 1).astype(np.float32).astype(np.float64)
    vel = np.zeros_like(pos)
    vel[:, 0] = rng.uniform(-0.3, 0.3, n)
    vel[:, 1] = -2.0
    vel[:, 2] = rng.uniform(-0.3, 0.3, n)
    vel = vel.astype(np.float32).astype(np.float64)
    prm = scenes.sph_params(h, s, 1e-3, erosion)
    return scenes.Scene(f"contact_block_{n}", prm, ter, pos, vel, np.arange(n, dtype=np.int32), 3,
                        dict(s=s, h=h))
