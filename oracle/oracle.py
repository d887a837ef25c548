"""ctypes binding of the CPU oracle (test infrastructure only).

Loads oracle/liboracle.so (the oracle's C restatement) and, when it was built,
oracle/_ref/libsphref.so (the reference's own FluidSystemSPH compiled from the
reference checkout).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker or the CPU
baseline.  Nothing here is part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsphref.so")

# FluidParticle (Erosion/fluid_system.h:49-64), 112 bytes
PARTICLE_DTYPE = np.dtype([
    ("Id", "<i4"), ("Position", "<f4", 3), ("Velocity", "<f4", 3),
    ("Acceleration", "<f4", 3), ("Density", "<f4"), ("Pressure", "<f4"),
    ("PressureForce", "<f4", 3), ("ViscosityForce", "<f4", 3),
    ("GravityForce", "<f4", 3), ("SurfaceForce", "<f4", 3),
    ("SurfaceNormal", "<f4", 3), ("NeighbId", "<i4"),
])
assert PARTICLE_DTYPE.itemsize == 112


class RefParams(C.Structure):
    """Reference parameters, Erosion/fluid_system.h:460-478."""
    _fields_ = [("g", C.c_float * 3), ("dt", C.c_float), ("p0", C.c_float),
                ("mass", C.c_float), ("visc", C.c_float), ("surf_tens", C.c_float),
                ("k", C.c_float), ("h", C.c_float), ("len", C.c_float)]

    def as_array(self):
        """Order used by oracle/_ref's refx_step."""
        return np.array([self.h, self.mass, self.p0, self.k, self.visc, self.surf_tens,
                         self.g[0], self.g[1], self.g[2], self.dt, self.len], dtype=np.float32)


class OrcScene(C.Structure):
    _fields_ = [("h", C.c_double), ("mass", C.c_double), ("rho0", C.c_double),
                ("k", C.c_double), ("visc", C.c_double), ("surf_tens", C.c_double),
                ("g", C.c_double * 3), ("dt", C.c_double),
                ("box_c", C.c_double * 3), ("box_e", C.c_double),
                ("tnx", C.c_int32), ("tnz", C.c_int32), ("t_org", C.c_double * 3),
                ("t_cell", C.c_double), ("fx_exp", C.c_int32), ("erosion", C.c_int32),
                ("kc", C.c_double), ("ke", C.c_double), ("kd", C.c_double), ("band", C.c_double),
                ("lanes", C.c_int32), ("slabs", C.c_int32), ("yz", C.c_int32)]


class OrcConsts(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("h", "h2", "thr", "t3h2", "eps_r", "inv_sqrt3", "mc6",
                                           "pf_scale", "vf_scale", "ng", "rho0", "k", "mass",
                                           "surf_tens")] + [
        ("g", C.c_float * 3), ("dt", C.c_float), ("box_c", C.c_float * 3), ("box_e", C.c_float),
        ("box_lo", C.c_float * 3), ("box_hi", C.c_float * 3), ("cell", C.c_float),
        ("inv_cell", C.c_float), ("grid_org", C.c_float * 3), ("dims", C.c_int32 * 3),
        ("t_org", C.c_float * 3), ("t_inv", C.c_float), ("hq", C.c_float),
        ("kcv", C.c_float), ("ke", C.c_float), ("kd", C.c_float), ("band", C.c_float),
        ("up", C.c_float), ("dn", C.c_float), ("dmax", C.c_float),
        ("dh", C.c_double), ("dh2", C.c_double), ("dthr_unused", C.c_double),
        ("inv_yz", C.c_float), ("yz_scale", C.c_float), ("yz", C.c_int32)]


class OrcDiag(C.Structure):
    _fields_ = [("rho", C.c_void_p), ("acc", C.c_void_p), ("pforce", C.c_void_p),
                ("vforce", C.c_void_p), ("sforce", C.c_void_p), ("normal", C.c_void_p),
                ("nbmax", C.c_void_p), ("nbcount", C.c_void_p)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.ref_default_params.argtypes = [C.POINTER(RefParams)]
        L.ref_lattice.argtypes = [C.c_int, P, C.c_int, P, C.c_int]
        L.ref_lattice.restype = C.c_int
        L.ref_step.argtypes = [P, C.c_int, C.POINTER(RefParams)]
        L.ref_step_sampled.argtypes = [P, C.c_int, C.POINTER(RefParams), C.c_int]
        L.ref_step_sampled.restype = C.c_int
        L.ref_num_threads.restype = C.c_int
        L.orc_derive.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcConsts)]
        L.orc_pred_threshold.argtypes = [C.c_float]
        L.orc_pred_threshold.restype = C.c_float
        L.orc_root_f32_n.argtypes = [P, C.c_int, P, P]
        L.orc_row_width_f32_n.argtypes = [P, C.c_int, P]
        L.orc_cell_keys.argtypes = [C.POINTER(OrcConsts), P, C.c_int, P]
        L.orc_neighbours_brute.argtypes = [P, C.c_int, C.c_float, P, P, P]
        L.orc_neighbours_brute.restype = C.c_int64
        L.orc_step.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_int, P, P, P, P, P, P,
                               C.POINTER(OrcDiag)]
        L.orc_step.restype = C.c_int
        L.orc_heights.argtypes = [C.POINTER(OrcConsts), C.c_int, P, P]
        L.orc_erode.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, P, P, P, P, P]
        L.orc_erode.restype = C.c_int
        L.orc_terrain_mesh.argtypes = [P, C.c_int, C.c_int, C.c_float, P, P]
        L.orc_terrain_indices.argtypes = [C.c_int, C.c_int, P]
        L.orc_terrain_indices.restype = C.c_int64
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def reflib():
    """The reference's own step (oracle/_ref); None when it was not built."""
    global _ref
    if _ref is None and ref_available():
        R = C.CDLL(REF_PATH)
        R.refx_lattice.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        R.refx_lattice.restype = C.c_int
        R.refx_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        R.refx_grid_mesh.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                     C.POINTER(C.c_int)]
        R.refx_grid_mesh.restype = C.c_int
        _ref = R
    return _ref


# ---------------------------------------------------------------- ref32

def default_params(dt=0.0):
    p = RefParams()
    lib().ref_default_params(C.byref(p))
    p.dt = dt
    return p


def lattice(n, origin=(0.0, 0.0, 0.0), first_id=0):
    org = np.asarray(origin, dtype=np.float32)
    cnt = lib().ref_lattice(n, _ptr(org), first_id, None, 0)
    out = np.zeros(cnt, dtype=PARTICLE_DTYPE)
    lib().ref_lattice(n, _ptr(org), first_id, _ptr(out), cnt)
    return out


def ref32_step(particles, params, steps=1):
    P = np.ascontiguousarray(particles).copy()
    for _ in range(steps):
        lib().ref_step(_ptr(P), len(P), C.byref(params))
    return P


def ref32_step_sampled(particles, params, stride):
    return lib().ref_step_sampled(_ptr(particles), len(particles), C.byref(params), stride)


def ref32_threads():
    return lib().ref_num_threads()


# ------------------------------------------------------ reference binary

def refbin_lattice(n, add_n=0, origin=(0.0, 0.0, 0.0)):
    R = reflib()
    org = np.asarray(origin, dtype=np.float32)
    cnt = R.refx_lattice(n, add_n, _ptr(org), None, 0)
    out = np.zeros(cnt, dtype=PARTICLE_DTYPE)
    R.refx_lattice(n, add_n, _ptr(org), _ptr(out), cnt)
    return out


def refbin_step(particles, params, steps=1):
    R = reflib()
    P = np.ascontiguousarray(particles).copy()
    prm = params.as_array()
    R.refx_step(_ptr(P), len(P), _ptr(prm), steps)
    return P


# ------------------------------------------------------------------ orc

# cell split along y and z of the product library's default grid
YZ_DEFAULT = 2


def make_scene(**kw):
    s = OrcScene()
    defaults = dict(h=0.0457, mass=0.02, rho0=998.29, k=3.0, visc=3.5, surf_tens=0.0728,
                    g=(0.0, -9.82, 0.0), dt=0.0, box_c=(0.0, 0.0, 0.0), box_e=0.2,
                    tnx=0, tnz=0, t_org=(0.0, 0.0, 0.0), t_cell=1.0, fx_exp=60, erosion=0,
                    kc=0.05, ke=0.3, kd=0.3, band=0.0457, lanes=4, slabs=4, yz=YZ_DEFAULT)
    defaults.update(kw)
    for k, v in defaults.items():
        if isinstance(v, (tuple, list, np.ndarray)):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    return s


def derive(scene):
    c = OrcConsts()
    lib().orc_derive(C.byref(scene), C.byref(c))
    return c


def pred_threshold(h):
    return lib().orc_pred_threshold(float(np.float32(h)))


def row_width_f32(x):
    """The FP32 kernels' clipped-row half-width (before the margin)."""
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    w = np.empty_like(x)
    lib().orc_row_width_f32_n(x.ctypes.data, len(x), w.ctypes.data)
    return w


def root_f32(d2):
    """(r, 1/r) of FP32 squared distances as the FP32 kernels compute them."""
    d2 = np.ascontiguousarray(d2, dtype=np.float32).ravel()
    r = np.empty_like(d2)
    inv = np.empty_like(d2)
    lib().orc_root_f32_n(d2.ctypes.data, len(d2), r.ctypes.data, inv.ctypes.data)
    return r, inv


def cell_keys(consts, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    keys = np.zeros(len(pos), dtype=np.int32)
    lib().orc_cell_keys(C.byref(consts), _ptr(pos), len(pos), _ptr(keys))
    return keys


def neighbours_brute(pos, thr, csr=True):
    """Neighbour sets of the reference predicate, self included.
    Returns (counts, offsets, idx) with idx ascending per particle."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n = len(pos)
    counts = np.zeros(n, dtype=np.int32)
    total = lib().orc_neighbours_brute(_ptr(pos), n, float(thr), _ptr(counts), None, None)
    if not csr:
        return counts, None, None
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    idx = np.zeros(max(total, 1), dtype=np.int32)
    lib().orc_neighbours_brute(_ptr(pos), n, float(thr), _ptr(counts), _ptr(off), _ptr(idx))
    return counts, off, idx[:total]


def terrain_mesh(heights, cell, origin):
    """Grid::UpdateGrid vertex stream (6 floats per cell) of a float heightfield [nz, nx]."""
    H = np.ascontiguousarray(heights, dtype=np.float32)
    nz, nx = H.shape
    org = np.asarray(origin, dtype=np.float32)
    out = np.zeros(nx * nz * 6, dtype=np.float32)
    lib().orc_terrain_mesh(_ptr(H), nx, nz, float(np.float32(cell)), _ptr(org), _ptr(out))
    return out


def terrain_indices(nx, nz):
    n = lib().orc_terrain_indices(nx, nz, None)
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_terrain_indices(nx, nz, _ptr(out))
    return out


def refbin_grid_mesh(img, d):
    """The reference's Grid(d, d, d) UpdateGrid on the 512 x 512 uint8 image img
    (indexed img[x, z], as GetHeightfieldAt): (vertex floats, indices)."""
    R = reflib()
    img = np.ascontiguousarray(img, dtype=np.uint8)
    assert img.shape == (512, 512)
    nv = C.c_int()
    ni = R.refx_grid_mesh(_ptr(img), d, None, 0, None, 0, C.byref(nv))
    v = np.zeros(nv.value, dtype=np.float32)
    i = np.zeros(ni, dtype=np.uint32)
    R.refx_grid_mesh(_ptr(img), d, _ptr(v), nv.value, _ptr(i), ni, C.byref(nv))
    return v, i


def heights(consts, vol_fx):
    vol_fx = np.ascontiguousarray(vol_fx, dtype=np.int64)
    out = np.zeros(vol_fx.shape, dtype=np.float32)
    lib().orc_heights(C.byref(consts), vol_fx.size, _ptr(vol_fx), _ptr(out))
    return out


def orc_step(scene, pos, vel, sed, ids, vol=None, height=None, real=8, comp=1, diag=True):
    """One step of the library's algorithm on copies of the inputs.
    Returns dict(pos, vel, sed, vol, height, active, and diag arrays)."""
    n = len(pos)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3).copy()
    sed = np.ascontiguousarray(sed, dtype=np.int64).copy()
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    if len(sed) != n or len(ids) != n:
        raise ValueError("sed and ids need one entry per particle")
    if scene.tnx > 0 and scene.tnz > 0:
        # the C step reads and writes tnx * tnz volumes and heights
        nt = scene.tnx * scene.tnz
        if vol is None or height is None or np.size(vol) != nt or np.size(height) != nt:
            raise ValueError(f"a scene with a {scene.tnx} x {scene.tnz} terrain needs {nt} volumes and heights")
    if vol is not None:
        vol = np.ascontiguousarray(vol, dtype=np.int64).copy()
        height = np.ascontiguousarray(height, dtype=np.float32).copy()
    out = dict()
    d = OrcDiag()
    if diag:
        out["rho"] = np.zeros(n)
        for k in ("acc", "pforce", "vforce", "sforce", "normal"):
            out[k] = np.zeros((n, 3))
        out["nbmax"] = np.zeros(n, dtype=np.int32)
        out["nbcount"] = np.zeros(n, dtype=np.int32)
        for k in ("rho", "acc", "pforce", "vforce", "sforce", "normal", "nbmax", "nbcount"):
            setattr(d, k, out[k].ctypes.data)
    rc = lib().orc_step(C.byref(scene), real, comp, n, _ptr(pos), _ptr(vel), _ptr(sed), _ptr(ids),
                        _ptr(vol), _ptr(height), C.byref(d))
    if rc < 0:
        raise RuntimeError(f"orc_step failed: {rc}")
    out.update(pos=pos, vel=vel, sed=sed, vol=vol, height=height, active=rc)
    return out


def orc_erode(scene, pos, vel, sed, vol, height, real=4):
    n = len(pos)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
    sed = np.ascontiguousarray(sed, dtype=np.int64).copy()
    vol = np.ascontiguousarray(vol, dtype=np.int64).copy()
    height = np.ascontiguousarray(height, dtype=np.float32)
    active = lib().orc_erode(C.byref(scene), n, real, _ptr(pos), _ptr(vel), _ptr(sed),
                             _ptr(vol), _ptr(height))
    return sed, vol, active
