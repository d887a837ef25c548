"""The C++ drop-in (include/sphe_fluid_system.hpp) driven the way the
reference's front end drives FluidSystemSPH, against the reference step."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import sphe
from sphe import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_cpp_front_end_matches_reference(tmp_path):
    exe = tmp_path / "front"
    lib_dir = os.path.join(ROOT, "sph-erosion_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "front_end.cpp"), "-o", str(exe),
                        os.path.join(lib_dir, "libsphe.so"), "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    steps = 5
    out = subprocess.run([str(exe), str(steps)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    R = O.ref32_step(O.lattice(1000), O.default_params(dt=0.01), steps)
    for line in lines[:2]:
        f = line.split()
        pid = int(f[0])
        pos = np.array([float(x) for x in f[1:4]])
        vel = np.array([float(x) for x in f[4:7]])
        rho, nb = float(f[7]), int(f[8])
        assert np.abs(pos - R["Position"][pid]).max() <= 1e-5
        assert np.abs(vel - R["Velocity"][pid]).max() <= 1e-3 * max(1.0, np.abs(R["Velocity"][pid]).max())
        assert abs(rho - R["Density"][pid]) <= 1e-5 * R["Density"][pid]
        assert nb == R["NeighbId"][pid]
    assert lines[2] == "n 1125"   # AddParticles(125) appends a 5^3 lattice
    assert lines[3] == "n 1000"   # Reset re-creates Initialize's count
    assert lines[4] == "range_error 1"


def _hip_copy(dptr, nbytes):
    """Device to host copy through the HIP runtime (the test reads the
    library's device mesh buffer)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    out = np.zeros(nbytes // 4, dtype=np.float32)
    rc = hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(dptr), C.c_size_t(nbytes), C.c_int(2))
    assert rc == 0
    return out


def test_terrain_mesh_is_the_reference_grid_mesh():
    """In voxel units (cell 1, cell centres at integer x, z, integer heights)
    the device-computed mesh equals the reference's Grid::UpdateGrid /
    genIndices output bit for bit (fixture mesh_ref.npz)."""
    g = np.load(os.path.join(GOLD, "mesh_ref.npz"))
    h = g["h_xz"].astype(np.float32)
    d = h.shape[0]
    p = sphe.default_params(dt=0.0)
    ctx = sphe.Context()
    sid = ctx.add_scene(p, dict(nx=d, nz=d, origin=(-0.5, 0.0, -0.5), cell=1.0, heights=h.T.copy()))
    assert np.array_equal(ctx.terrain_heights(sid), h.T)
    verts, idx = ctx.terrain_mesh(sid)
    assert np.array_equal(verts.reshape(-1).view(np.uint32), g["verts"].view(np.uint32))
    assert np.array_equal(idx, g["indices"])


def test_terrain_mesh_after_erosion_matches_oracle():
    """After erosion steps the mesh of the eroded heights (host copy and the
    device buffer) equals the oracle's restatement bit for bit."""
    import ctypes as C
    sc = scenes.cfg1(n=4096, erosion=sphe.EROSION_ATOMIC)
    ctx = sphe.Context()
    sid = ctx.add_synthetic(sc)
    ctx.step(30)
    H = ctx.terrain_heights(sid)
    ref = O.terrain_mesh(H, sc.terrain["cell"], sc.terrain["origin"])
    verts, idx = ctx.terrain_mesh(sid)
    assert np.array_equal(verts.reshape(-1).view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(idx, O.terrain_indices(H.shape[1], H.shape[0]))
    dptr, n = C.c_void_p(), C.c_size_t()
    assert sphe.lib().sphe_terrain_mesh_device(ctx._p, sid, C.byref(dptr), C.byref(n)) == 0
    assert n.value == ref.size and dptr.value
    assert np.array_equal(_hip_copy(dptr.value, n.value * 4).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("precision", [sphe.FP32, sphe.FP64])
def test_particle_feed_gathered_on_the_device(precision):
    """Two scenes hold the same particles, one with ids 0 .. n-1 and one with
    ids 2i (the same order inside every cell, so the two step to the same
    bits).  After steps, positions, velocities, densities and pressures of the
    first are gathered in id order on the device, with no host sync; the
    second reads through the host; the two are equal bit for bit, positions
    and velocities are the FP32 roundings of the synced state, and
    sphe_positions_device holds the same positions (the second scene has
    none).  Reading changes nothing in the steps that follow."""
    import ctypes as C

    sc = scenes.cfg1(n=4096)
    ctx = sphe.Context(precision)
    a = ctx.add_synthetic(sc)
    b = ctx.add_scene(sc.params, sc.terrain)
    ctx.set_particles(b, sc.pos, sc.vel, ids=2 * sc.ids)
    ctx.step(6)
    bufs = (sphe.BUF_POS, sphe.BUF_VEL, sphe.BUF_DENSITY, sphe.BUF_PRESS)
    got = [ctx.read(a, w) for w in bufs]  # device gathers, before anything syncs the host
    dptr, n = C.c_void_p(), C.c_size_t()
    assert sphe.lib().sphe_positions_device(ctx._p, a, C.byref(dptr), C.byref(n)) == 0
    assert n.value == 3 * sc.n and dptr.value
    dev = _hip_copy(dptr.value, n.value * 4).reshape(-1, 3)
    assert sphe.lib().sphe_positions_device(ctx._p, b, C.byref(dptr), C.byref(n)) != 0
    want = [ctx.read(b, w) for w in bufs]  # the host path
    assert np.all(got[2] > 0)
    for g, w in zip(got + [dev], want + [want[0]]):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))
    gp, gv = ctx.state_f64(a)
    assert np.array_equal(got[0], gp.astype(np.float32)) and np.array_equal(got[1], gv.astype(np.float32))
    ctx.step(3)
    ea, eb = ctx.state_f64(a), ctx.state_f64(b)
    assert np.array_equal(ea[0], eb[0]) and np.array_equal(ea[1], eb[1])
    ctx.close()


def test_buffer_views_and_batches():
    """sphe_buffer's host views equal sphe_read_buffer's copies; a batch made
    by sphe_batch_create (terrains included) and stepped by sphe_batch_step
    equals a context given the same scenes one by one, bit for bit."""
    import ctypes as C

    from sphe import _lib as L

    scs = [scenes.cfg1(seed_idx=k, n=1500 + 700 * k) for k in range(3)]
    ctx = sphe.Context()
    sids = [ctx.add_synthetic(s) for s in scs]
    ctx.step(3)
    for s in sids:
        for which in (sphe.BUF_POS, sphe.BUF_VEL, sphe.BUF_DENSITY, sphe.BUF_PRESS, sphe.BUF_SED,
                      sphe.BUF_ID, sphe.BUF_MASS):
            assert np.array_equal(ctx.view(s, which), ctx.read(s, which))
    params = (sphe.Params * 3)(*[s.params for s in scs])
    tds = (sphe.TerrainDesc * 3)()
    hts = []
    for k, s in enumerate(scs):
        h = np.ascontiguousarray(s.terrain["heights"], dtype=np.float32)
        hts.append(h)
        tds[k].nx, tds[k].nz = int(s.terrain["nx"]), int(s.terrain["nz"])
        for a in range(3):
            tds[k].origin[a] = float(s.terrain["origin"][a])
        tds[k].cell = float(s.terrain["cell"])
        tds[k].heights = h.ctypes.data
    b = C.c_void_p()
    assert L.lib().sphe_batch_create(sphe.FP32, C.byref(params), C.byref(tds), 3, C.byref(b)) == 0
    bc = sphe.Context.__new__(sphe.Context)
    bc.precision, bc._p, bc._terrains = sphe.FP32, b, hts
    try:
        for k, s in enumerate(scs):
            bc.set_particles(k, s.pos, s.vel, s.ids)
        ref = sphe.Context()
        for s in scs:
            ref.add_synthetic(s)
        assert L.lib().sphe_batch_step(b, 4) == 0
        ref.step(4)
        for k in range(3):
            pa, va = bc.state_f64(k)
            pb, vb = ref.state_f64(k)
            assert np.array_equal(pa, pb) and np.array_equal(va, vb)
            assert np.array_equal(bc.terrain_volume_fx(k), ref.terrain_volume_fx(k))
    finally:
        bc.close()
