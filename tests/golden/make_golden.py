#!/usr/bin/env python3
"""Writes the golden fixtures of the reference step.

Runs the reference's own FluidSystemSPH (Erosion/fluid_system.h, compiled
from the reference checkout into oracle/_ref/libsphref.so by oracle/Makefile)
on small inputs and stores inputs and outputs as .npz.  Run in a checkout
that has the reference; the fixtures are committed, the reference is not.

  lattice.npz      Initialize(1000), Initialize(100), Initialize(125) +
                   AddParticles(125): Id, Position, Velocity, Acceleration
  f1_lattice.npz   Initialize(1000), dt = 0.01 (Erosion/main.cpp:481):
                   full 112-byte records after 1 and 10 steps
  f2_coincident.npz Initialize(125) + AddParticles(125) (coincident pairs,
                   the dist < 1e-4 branch of gradPressure), after 1 and 3 steps
  f3_jitter.npz    jittered 4096-particle block, box len 0.5, p0 = 1000,
                   g = (0,-9.8,0), dt = 1e-3: input and state after 1 step
  f4_box.npz       isolated particles crossing 1, 2 or 3 walls, on a wall, and
                   at rest outside the box: input and state after 1 step
  mesh_ref.npz     the reference's Grid(24, 24, 24) UpdateGrid on a 512 x 512
                   uint8 heightfield with smooth heights in [0, 22] in the
                   24 x 24 corner: GetSurfaceParts and GetIndices
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

INIT_FIELDS = ["Id", "Position", "Velocity", "Acceleration"]


def records(n):
    P = np.zeros(n, dtype=O.PARTICLE_DTYPE)
    P["Id"] = np.arange(n)
    P["NeighbId"] = -1
    return P


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libsphref.so is not built (needs the reference checkout)")
    out = {}
    for name, (n, add) in {"l1000": (1000, 0), "l100": (100, 0), "l125_125": (125, 125)}.items():
        L = O.refbin_lattice(n, add)
        for f in INIT_FIELDS:
            out[f"{name}_{f}"] = L[f]
    np.savez_compressed(os.path.join(HERE, "lattice.npz"), **out)

    p = O.default_params(dt=0.01)
    L = O.refbin_lattice(1000)
    s1 = O.refbin_step(L, p, 1)
    s10 = O.refbin_step(L, p, 10)
    np.savez_compressed(os.path.join(HERE, "f1_lattice.npz"), step1=s1.view(np.uint8),
                        step10=s10.view(np.uint8), input=np.array(0))

    L = O.refbin_lattice(125, 125)
    c1 = O.refbin_step(L, p, 1)
    c3 = O.refbin_step(L, p, 3)
    np.savez_compressed(os.path.join(HERE, "f2_coincident.npz"), step1=c1.view(np.uint8),
                        step3=c3.view(np.uint8))

    rng = np.random.default_rng(20261202)
    s = 0.0457 / 2.15
    g = (np.arange(16) - 7.5) * s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel() - 0.2, Z.ravel()], axis=1)
    pos += rng.uniform(-0.1 * s, 0.1 * s, size=pos.shape)
    vel = rng.uniform(-0.05, 0.05, size=pos.shape)
    P = records(len(pos))
    P["Position"] = pos.astype(np.float32)
    P["Velocity"] = vel.astype(np.float32)
    p3 = O.default_params(dt=1e-3)
    p3.len, p3.p0 = 0.5, 1000.0
    p3.g[1] = -9.8
    R = O.refbin_step(P, p3, 1)
    np.savez_compressed(os.path.join(HERE, "f3_jitter.npz"), in_pos=P["Position"], in_vel=P["Velocity"],
                        Position=R["Position"], Velocity=R["Velocity"], Acceleration=R["Acceleration"],
                        Density=R["Density"], Pressure=R["Pressure"], NeighbId=R["NeighbId"],
                        PressureForce=R["PressureForce"], SurfaceNormal=R["SurfaceNormal"])

    # isolated particles (pairwise far apart) hitting the box walls
    cases = np.array([
        # position                velocity
        [0.195, 0.0, 0.0, 1.0, 0.0, 0.0],        # +x wall
        [-0.195, 0.1, 0.0, -1.0, 0.0, 0.0],      # -x wall
        [0.0, -0.198, 0.1, 0.0, -0.5, 0.0],      # floor
        [0.1, 0.198, -0.1, 0.0, 2.0, 0.0],       # ceiling
        [0.0, 0.1, 0.199, 0.0, 0.0, 0.3],        # +z wall
        [0.199, -0.199, 0.05, 1.0, -1.0, 0.0],   # two walls, x dominant
        [-0.15, -0.199, -0.199, 0.0, -1.0, -2.0],  # two walls, z dominant
        [0.199, 0.199, 0.199, 1.0, 1.0, 1.0],    # corner, tie on all axes
        [0.2, 0.05, -0.05, 0.0, 0.0, 0.0],       # on the wall at rest (d = 0)
        [0.3, -0.05, 0.15, 0.0, 0.0, 0.0],       # outside, moving with gravity only
        [0.05, 0.05, 0.05, 0.1, 0.1, 0.1],       # inside, no collision
        [-0.1, -0.3, -0.1, 0.0, 0.0, 0.0],       # outside with v' = 0 exactly -> NaN velocity
    ], dtype=np.float32)
    # v_y = -fl(g dt): the isolated particle's velocity after the kick is 0
    cases[11, 4] = -(np.float32(-9.82) * np.float32(0.01))
    P = records(len(cases))
    P["Position"] = cases[:, :3]
    P["Velocity"] = cases[:, 3:]
    p4 = O.default_params(dt=0.01)
    R = O.refbin_step(P, p4, 1)
    np.savez_compressed(os.path.join(HERE, "f4_box.npz"), in_pos=P["Position"], in_vel=P["Velocity"],
                        Position=R["Position"], Velocity=R["Velocity"], Acceleration=R["Acceleration"],
                        Density=R["Density"], NeighbId=R["NeighbId"])
    # terrain mesh (grid.h:118-176): smooth integer heights below the clamp
    # of UpdateGrid (dimy - 1) in the grid's corner of the 512 x 512 image
    d = 24
    x = np.arange(d)
    hx = 11.0 + 7.0 * np.sin(x / 3.1)[:, None] + 4.0 * np.cos(x / 2.3)[None, :]
    h = np.clip(np.rint(hx), 0, d - 2).astype(np.uint8)  # h[x, z]
    img = np.zeros((512, 512), dtype=np.uint8)
    img[:d, :d] = h
    v, i = O.refbin_grid_mesh(img, d)
    np.savez_compressed(os.path.join(HERE, "mesh_ref.npz"), h_xz=h, verts=v, indices=i)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
