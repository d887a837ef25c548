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
