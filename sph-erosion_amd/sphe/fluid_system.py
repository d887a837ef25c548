"""FluidSystemSPH and Grid with the reference's method names, over the C ABI.

Mirrors Erosion/fluid_system.h:66-485 (class FluidSystemSPH) and the terrain
part of Erosion/grid.h (class Grid) so that code written against the
reference reads the same.  One FluidSystemSPH is one scene of a sphe context;
Run advances it on the GPU.  The C++ drop-in with the same names is
include/sphe_fluid_system.hpp.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FP32, P_GRAV, P_MASS, P_P0, P_SURF_TEN, P_VISC, lib
from .core import Context, default_params


class Grid:
    """Terrain holder with Grid's calls (grid.h:72-176): LoadHeightfield of a
    uint8 image, GetHeightfieldAt, UpdateGrid -> surface vertices/indices."""

    def __init__(self, dimX=512, dimY=512, dimZ=512):
        self.dims = (dimX, dimY, dimZ)
        self.map = np.zeros((512, 512), dtype=np.uint8)  # grid.h:78-81

    def LoadHeightfield(self, img):
        self.map = np.ascontiguousarray(img, dtype=np.uint8).reshape(self.map.shape)

    def GetHeightfieldAt(self, x, y):
        return int(self.map.reshape(-1)[self.map.shape[1] * x + y])  # grid.h:104-107


class FluidSystemSPH:
    def __init__(self, precision=FP32, params=None, ctx=None):
        self.ctx = ctx if ctx is not None else Context(precision)
        self.scene = self.ctx.add_scene(params if params is not None else default_params())

    # fluid_system.h:74-102
    def Initialize(self, nParts):
        self.ctx._chk(lib().sphe_init_lattice(self.ctx._p, self.scene, int(nParts)), "Initialize")

    # fluid_system.h:104-183 (the Grid argument is unused there too)
    def Run(self, grid=None, steps=1):
        self.ctx.step(steps)

    def SetOrigin(self, o):
        arr = (C.c_float * 3)(*[float(v) for v in o])
        self.ctx._chk(lib().sphe_set_origin(self.ctx._p, self.scene, arr), "SetOrigin")

    def GetOrigin(self):
        arr = (C.c_float * 3)()
        self.ctx._chk(lib().sphe_get_origin(self.ctx._p, self.scene, arr), "GetOrigin")
        return tuple(arr)

    def SetDeltaTime(self, dt):
        self.ctx._chk(lib().sphe_set_dt(self.ctx._p, self.scene, float(dt)), "SetDeltaTime")

    def GetDeltaTime(self):
        v = C.c_float()
        self.ctx._chk(lib().sphe_get_dt(self.ctx._p, self.scene, C.byref(v)), "GetDeltaTime")
        return v.value

    def PrintCoords(self):
        for i, p in enumerate(self.ctx.positions(self.scene)):
            print(f"[{i}] {p[0]} {p[1]} {p[2]}")

    def AddParticles(self, n):
        self.ctx._chk(lib().sphe_add_particles(self.ctx._p, self.scene, int(n)), "AddParticles")

    def Reset(self):
        self.ctx._chk(lib().sphe_reset(self.ctx._p, self.scene), "Reset")

    def _param(self, which):
        return lib().sphe_param_ptr(self.ctx._p, self.scene, which)

    # pointer getters (fluid_system.h:261-284): writable float storage
    def GetMass(self):
        return self._param(P_MASS)

    def GetVisc(self):
        return self._param(P_VISC)

    def GetSurfTen(self):
        return self._param(P_SURF_TEN)

    def Getp0(self):
        return self._param(P_P0)

    def GetGrav(self):
        return self._param(P_GRAV)

    def GetParticle(self, pid):
        return self.ctx.particle(self.scene, int(pid))

    def NumParticles(self):
        return self.ctx.num_particles(self.scene)

    # terrain coupling (Grid::LoadHeightfield / UpdateGrid / GetSurfaceParts /
    # GetIndices, grid.h:98-176), as the C++ drop-in header exposes it
    def LoadHeightfield(self, img, scale, cell, origin, erosion=None):
        """img: uint8 array (rows, columns); terrain cell (x, z) takes
        GetHeightfieldAt(x, z) = img[x, z] (grid.h:104-107)."""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        p = self.ctx.get_params(self.scene)
        if erosion is not None:
            p.erosion = erosion
            self.ctx.set_params(self.scene, p)
        self.ctx.set_terrain(self.scene, dict(nx=img.shape[0], nz=img.shape[1], origin=tuple(origin),
                                              cell=float(cell), heights=None))
        self.ctx.terrain_load_u8(self.scene, img, scale)

    def GetSurfaceParts(self):
        return self.ctx.terrain_mesh(self.scene)[0].reshape(-1)

    def GetIndices(self):
        return self.ctx.terrain_mesh(self.scene)[1]
