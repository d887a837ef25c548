"""Python mirror of the C ABI (include/sphe.h) for tests and benchmarks.

Context wraps one sphe_ctx: a batch of scenes advanced together on the GPU.
All compute runs in libsphe.so; this module only moves host arrays across the
boundary.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (BUF_DENSITY, BUF_ID, BUF_MASS, BUF_POS, BUF_PRESS, BUF_SED, BUF_VEL, FP32, FP64,
                   Params, Particle, TerrainDesc, check, lib)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_params(**kw):
    """Reference defaults (Erosion/fluid_system.h:460-478), overridden by kw."""
    p = Params()
    lib().sphe_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("gravity", "box_center"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
        else:
            setattr(p, k, v)
    return p


def params_dict(p):
    return {k: (list(getattr(p, k)) if k in ("gravity", "box_center") else getattr(p, k))
            for k, _ in Params._fields_}


class Context:
    def __init__(self, precision=FP32):
        self.precision = precision
        self._p = C.c_void_p()
        check(lib().sphe_create(precision, C.byref(self._p)), None, "sphe_create")
        self._terrains = []

    # ---- lifetime
    def close(self):
        if self._p:
            lib().sphe_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        return check(rc, self._p, what)

    # ---- scenes
    def add_scene(self, params, terrain=None):
        td = None
        if terrain is not None:
            hts = np.ascontiguousarray(terrain["heights"], dtype=np.float32)
            self._terrains.append(hts)
            td = TerrainDesc()
            td.nx, td.nz = int(terrain["nx"]), int(terrain["nz"])
            for i in range(3):
                td.origin[i] = float(terrain["origin"][i])
            td.cell = float(terrain["cell"])
            td.heights = hts.ctypes.data
        sid = C.c_int()
        self._chk(lib().sphe_add_scene(self._p, C.byref(params), C.byref(td) if td else None,
                                       C.byref(sid)), "add_scene")
        return sid.value

    def add_synthetic(self, sc):
        """Adds a scenes.Scene (params, terrain and particle state)."""
        sid = self.add_scene(sc.params, sc.terrain)
        self.set_particles(sid, sc.pos, sc.vel, sc.ids)
        return sid

    def set_params(self, scene, params):
        self._chk(lib().sphe_set_params(self._p, scene, C.byref(params)), "set_params")

    def get_params(self, scene):
        p = Params()
        self._chk(lib().sphe_get_params(self._p, scene, C.byref(p)), "get_params")
        return p

    def set_particles(self, scene, pos, vel=None, ids=None, sed=None):
        n = len(pos)
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        sed = None if sed is None else np.ascontiguousarray(sed, dtype=np.int64)
        if self.precision == FP64:
            pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
            vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
            rc = lib().sphe_set_particles_f64(self._p, scene, n, _ptr(pos), _ptr(vel), _ptr(ids), _ptr(sed))
        else:
            pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(n, 3)
            vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
            rc = lib().sphe_set_particles(self._p, scene, n, _ptr(pos), _ptr(vel), _ptr(ids), _ptr(sed))
        self._chk(rc, "set_particles")

    def num_particles(self, scene):
        n = C.c_int()
        self._chk(lib().sphe_num_particles(self._p, scene, C.byref(n)), "num_particles")
        return n.value

    # ---- stepping
    def step(self, n=1):
        self._chk(lib().sphe_step(self._p, n), "step")

    def sync(self):
        self._chk(lib().sphe_sync(self._p), "sync")

    def use_graph(self, on):
        self._chk(lib().sphe_use_graph(self._p, int(on)), "use_graph")

    # ---- reads (ascending id order)
    def read(self, scene, which):
        n = self.num_particles(scene)
        if which in (BUF_POS, BUF_VEL):
            out = np.zeros((n, 3), dtype=np.float32)
        elif which == BUF_ID:
            out = np.zeros(n, dtype=np.int32)
        else:
            out = np.zeros(n, dtype=np.float32)
        self._chk(lib().sphe_read_buffer(self._p, scene, which, _ptr(out), out.nbytes), "read_buffer")
        return out

    def view(self, scene, which):
        """sphe_buffer: the context's host view of a buffer (copied out here)."""
        v = C.c_void_p()
        n = C.c_size_t()
        self._chk(lib().sphe_buffer(self._p, scene, which, C.byref(v), C.byref(n)), "buffer")
        comps = 3 if which in (BUF_POS, BUF_VEL) else 1
        ctype = C.c_int32 if which == BUF_ID else C.c_float
        arr = np.ctypeslib.as_array(C.cast(v, C.POINTER(ctype)), shape=(max(1, n.value * comps),))
        out = arr[:n.value * comps].copy()
        return out.reshape(n.value, 3) if comps == 3 else out

    def positions(self, scene):
        return self.read(scene, BUF_POS)

    def velocities(self, scene):
        return self.read(scene, BUF_VEL)

    def ids(self, scene):
        return self.read(scene, BUF_ID)

    def density(self, scene):
        return self.read(scene, BUF_DENSITY)

    def pressure(self, scene):
        return self.read(scene, BUF_PRESS)

    def sediment(self, scene):
        return self.read(scene, BUF_SED)

    def mass_buffer(self, scene):
        return self.read(scene, BUF_MASS)

    def state_f64(self, scene):
        n = self.num_particles(scene)
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        self._chk(lib().sphe_read_state_f64(self._p, scene, _ptr(pos), _ptr(vel)), "read_state_f64")
        return pos, vel

    def sediment_fx(self, scene):
        out = np.zeros(self.num_particles(scene), dtype=np.int64)
        self._chk(lib().sphe_read_sediment_fx(self._p, scene, _ptr(out)), "read_sediment_fx")
        return out

    def particle(self, scene, pid):
        p = Particle()
        self._chk(lib().sphe_get_particle(self._p, scene, pid, C.byref(p)), "get_particle")
        return p

    # ---- terrain
    def terrain_dims(self, scene):
        nx, nz = C.c_int(), C.c_int()
        self._chk(lib().sphe_terrain_dims(self._p, scene, C.byref(nx), C.byref(nz)), "terrain_dims")
        return nx.value, nz.value

    def terrain_heights(self, scene):
        nx, nz = self.terrain_dims(scene)
        out = np.zeros((nz, nx), dtype=np.float32)
        self._chk(lib().sphe_terrain_heights(self._p, scene, _ptr(out), out.size), "terrain_heights")
        return out

    def terrain_volume_fx(self, scene):
        nx, nz = self.terrain_dims(scene)
        out = np.zeros((nz, nx), dtype=np.int64)
        self._chk(lib().sphe_terrain_volume_fx(self._p, scene, _ptr(out), out.size), "terrain_volume_fx")
        return out

    def set_terrain(self, scene, terrain):
        """Attach, replace or (None) remove a scene's heightfield; heights may be
        None for a flat terrain at the origin's y."""
        if terrain is None:
            self._chk(lib().sphe_set_terrain(self._p, scene, None), "set_terrain")
            return
        td = TerrainDesc()
        td.nx, td.nz = int(terrain["nx"]), int(terrain["nz"])
        for i in range(3):
            td.origin[i] = float(terrain["origin"][i])
        td.cell = float(terrain["cell"])
        if terrain.get("heights") is not None:
            hts = np.ascontiguousarray(terrain["heights"], dtype=np.float32)
            self._terrains.append(hts)
            td.heights = hts.ctypes.data
        self._chk(lib().sphe_set_terrain(self._p, scene, C.byref(td)), "set_terrain")

    def set_terrain_volume_fx(self, scene, vol):
        """Restore fixed-point cell volumes (a checkpoint's terrain)."""
        v = np.ascontiguousarray(vol, dtype=np.int64).ravel()
        self._chk(lib().sphe_set_terrain_volume_fx(self._p, scene, _ptr(v), v.size), "set_terrain_volume_fx")

    def fixed_exponent(self, scene):
        e = C.c_int()
        self._chk(lib().sphe_fixed_exponent(self._p, scene, C.byref(e)), "fixed_exponent")
        return e.value

    def terrain_load_u8(self, scene, img, scale):
        """Grid::LoadHeightfield of an image array of shape (rows, columns):
        terrain cell (x, z) = img[x, z] (GetHeightfieldAt, grid.h:104-107)."""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape[0], img.shape[1]
        self._chk(lib().sphe_terrain_load_u8(self._p, scene, _ptr(img), w, h, float(scale)), "terrain_load_u8")

    def terrain_mesh(self, scene):
        nv, ni = C.c_size_t(), C.c_size_t()
        self._chk(lib().sphe_terrain_mesh(self._p, scene, None, C.byref(nv), None, C.byref(ni)), "mesh sizes")
        verts = np.zeros(nv.value, dtype=np.float32)
        idx = np.zeros(ni.value, dtype=np.uint32)
        self._chk(lib().sphe_terrain_mesh(self._p, scene, _ptr(verts), C.byref(nv), _ptr(idx), C.byref(ni)),
                  "terrain_mesh")
        return verts.reshape(-1, 6), idx

    # ---- bookkeeping
    def mass(self, scene):
        t, s, u = C.c_double(), C.c_double(), C.c_uint64()
        self._chk(lib().sphe_mass(self._p, scene, C.byref(t), C.byref(s), C.byref(u)), "mass")
        return t.value, s.value, u.value

    def mass_all(self):
        """(terrain, sediment, units) arrays over every scene, one launch."""
        n = lib().sphe_scene_count
        cnt = C.c_int()
        self._chk(n(self._p, C.byref(cnt)), "scene_count")
        t = np.zeros(cnt.value)
        s = np.zeros(cnt.value)
        u = np.zeros(cnt.value, dtype=np.uint64)
        self._chk(lib().sphe_mass_all(self._p, _ptr(t), _ptr(s), _ptr(u), cnt.value), "mass_all")
        return t, s, u

    def mass_all_noalloc(self):
        """mass_all from the library's static scratch (no device allocation)."""
        cnt = C.c_int()
        self._chk(lib().sphe_scene_count(self._p, C.byref(cnt)), "scene_count")
        t = np.zeros(cnt.value)
        s = np.zeros(cnt.value)
        u = np.zeros(cnt.value, dtype=np.uint64)
        self._chk(lib().sphe_mass_all_noalloc(self._p, _ptr(t), _ptr(s), _ptr(u), cnt.value), "mass_all_noalloc")
        return t, s, u

    def status(self):
        v = C.c_uint32()
        self._chk(lib().sphe_status(self._p, C.byref(v)), "status")
        return v.value

    # ---- measurement and tuning
    def profile(self, on):
        self._chk(lib().sphe_profile(self._p, int(on)), "profile")

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_int()
        self._chk(lib().sphe_profile_read(self._p, kernel.encode(), C.byref(ms), C.byref(n)), "profile_read")
        return ms.value, n.value

    def profile_names(self):
        buf = C.create_string_buffer(1024)
        self._chk(lib().sphe_profile_names(self._p, buf, 1024), "profile_names")
        return [x for x in buf.value.decode().split(",") if x]

    def kernel_bytes(self, kernel):
        b = C.c_double()
        self._chk(lib().sphe_kernel_bytes(self._p, kernel.encode(), C.byref(b)), "kernel_bytes")
        return b.value

    def set_tuning(self, block, remap):
        self._chk(lib().sphe_set_tuning(self._p, block, int(remap)), "set_tuning")

    def get_tuning(self):
        b, r = C.c_int(), C.c_int()
        self._chk(lib().sphe_get_tuning(self._p, C.byref(b), C.byref(r)), "get_tuning")
        return b.value, r.value

    def set_force_block(self, block):
        """Block size of the force pass (0: the density pass's, as set_tuning leaves it)."""
        self._chk(lib().sphe_set_force_block(self._p, int(block)), "set_force_block")

    def get_force_block(self):
        v = C.c_int()
        self._chk(lib().sphe_get_force_block(self._p, C.byref(v)), "get_force_block")
        return v.value

    def set_lanes(self, lanes):
        self._chk(lib().sphe_set_lanes(self._p, lanes), "set_lanes")

    def get_lanes(self):
        v = C.c_int()
        self._chk(lib().sphe_get_lanes(self._p, C.byref(v)), "get_lanes")
        return v.value

    def save_state(self):
        self._chk(lib().sphe_save_state(self._p), "save_state")

    def restore_state(self):
        self._chk(lib().sphe_restore_state(self._p), "restore_state")

    def set_list_capacity(self, entries):
        self._chk(lib().sphe_set_list_capacity(self._p, entries), "set_list_capacity")

    def get_list_capacity(self):
        v = C.c_int()
        self._chk(lib().sphe_get_list_capacity(self._p, C.byref(v)), "get_list_capacity")
        return v.value

    def get_overflow_capacity(self):
        """(entries per overflow chunk, chunks in the pool)."""
        e, n = C.c_int(), C.c_int64()
        self._chk(lib().sphe_get_overflow_capacity(self._p, C.byref(e), C.byref(n)), "get_overflow_capacity")
        return e.value, n.value

    def set_subcells(self, slabs):
        self._chk(lib().sphe_set_subcells(self._p, slabs), "set_subcells")

    def get_subcells(self):
        v = C.c_int()
        self._chk(lib().sphe_get_subcells(self._p, C.byref(v)), "get_subcells")
        return v.value

    def set_yz_split(self, split):
        """Cells 1/split of the x edge along y and z (1, 2 or 3)."""
        self._chk(lib().sphe_set_yz_split(self._p, split), "set_yz_split")

    def get_yz_split(self):
        v = C.c_int()
        self._chk(lib().sphe_get_yz_split(self._p, C.byref(v)), "get_yz_split")
        return v.value

    def autotune(self, steps_per_trial=5):
        """Tunes the block sizes; returns (density block, xcd run); get_force_block() has the force pass's."""
        self._chk(lib().sphe_autotune(self._p, steps_per_trial), "autotune")
        return self.get_tuning()

    def autotune_grid(self, steps_per_trial=3):
        """Tunes the cell grid; returns (y/z split, x slabs, log2 cells per scan tile)."""
        self._chk(lib().sphe_autotune_grid(self._p, steps_per_trial), "autotune_grid")
        return self.get_yz_split(), self.get_subcells(), self.get_scan_tile()

    def set_persistent(self, on):
        self._chk(lib().sphe_set_persistent(self._p, int(on)), "set_persistent")

    def get_persistent(self):
        v = C.c_int()
        self._chk(lib().sphe_get_persistent(self._p, C.byref(v)), "get_persistent")
        return bool(v.value)

    def set_sequence_steps(self, steps):
        self._chk(lib().sphe_set_sequence_steps(self._p, steps), "set_sequence_steps")

    def get_sequence_steps(self):
        v = C.c_int()
        self._chk(lib().sphe_get_sequence_steps(self._p, C.byref(v)), "get_sequence_steps")
        return v.value

    def set_scan_tile(self, log2_cells):
        self._chk(lib().sphe_set_scan_tile(self._p, log2_cells), "set_scan_tile")

    def get_scan_tile(self):
        v = C.c_int()
        self._chk(lib().sphe_get_scan_tile(self._p, C.byref(v)), "get_scan_tile")
        return v.value

    # ---- verification hooks
    def debug_cell_keys(self, scene):
        out = np.zeros(self.num_particles(scene), dtype=np.int32)
        self._chk(lib().sphe_debug_cell_keys(self._p, scene, _ptr(out)), "debug_cell_keys")
        return out

    def debug_consts(self, scene):
        out = np.zeros(20)
        self._chk(lib().sphe_debug_consts(self._p, scene, _ptr(out), 20), "debug_consts")
        keys = ["h", "h2", "thr", "t3h2", "mc6", "pf_scale", "vf_scale", "ng", "cell", "inv_cell",
                "org_x", "org_y", "org_z", "dim_x", "dim_y", "dim_z", "hq", "kcv", "up", "dn"]
        return dict(zip(keys, out))

    def debug_neighbours(self, scene, csr=True):
        n = self.num_particles(scene)
        counts = np.zeros(n, dtype=np.int32)
        self._chk(lib().sphe_debug_neighbours(self._p, scene, _ptr(counts), None, None), "debug_neighbours")
        if not csr:
            return counts, None, None
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=off[1:])
        idx = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        self._chk(lib().sphe_debug_neighbours(self._p, scene, _ptr(counts), _ptr(off), _ptr(idx)),
                  "debug_neighbours")
        return counts, off, idx[:off[-1]]

    def debug_copy(self, nbytes, iters=20):
        ms = C.c_double()
        self._chk(lib().sphe_debug_copy(self._p, nbytes, iters, C.byref(ms)), "debug_copy")
        return ms.value

    def debug_density(self, scene):
        out = np.zeros(self.num_particles(scene))
        self._chk(lib().sphe_debug_density(self._p, scene, _ptr(out)), "debug_density")
        return out

    def debug_accel(self, scene):
        """Acceleration of every particle in the current state (id order, n x 3)."""
        out = np.zeros((self.num_particles(scene), 3))
        self._chk(lib().sphe_debug_accel(self._p, scene, _ptr(out)), "debug_accel")
        return out
