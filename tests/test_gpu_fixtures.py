"""GPU parity on the reference's fixtures and on the branches that need a
particular state: terrain contact (collision and bulk erosion), coincident
pairs, the box edge cases, the lattice initialiser, GetParticle's force
fields and heightfield ingest.

Bars:
  * the library's FP32 step equals the oracle's same-order twin bit for bit;
  * against the reference's own step (the golden fixtures written by the
    reference's FluidSystemSPH): density within 1e-5 relative, NeighbId exact,
    force fields within 1e-5 of the fixture's largest magnitude, positions
    within 1e-6; the box cases and the lattice bit for bit;
  * single steps against the FP64 oracle from the same state: positions
    within 1e-6, velocities within 1e-4 of the largest speed, density within
    1e-5 relative, sediment and terrain-volume changes within 1e-4 of the
    largest change.
"""
import os

import numpy as np
import pytest

import oracle as O
import sphe
from gpu_common import bilinear_heights, contact_block, gpu_context, oracle_inputs, orc_scene_from
from sphe import scenes

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DENSITY_TOL = 1e-5


def recs(buf):
    return np.frombuffer(buf.tobytes(), dtype=O.PARTICLE_DTYPE)


def f32_bits_equal(a, b):
    """Bit equality of float32 arrays, NaN equal to NaN (the sign and payload
    of a NaN differ between the CPU and the GPU)."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


# ------------------------------------------------------------- terrain contact

@pytest.mark.parametrize("erosion", [sphe.EROSION_ATOMIC, sphe.EROSION_TWO_PHASE])
def test_terrain_contact_bit_exact_twin(erosion):
    """A block driven into the terrain: every step some particles collide with
    the heightfield and about three quarters erode or deposit.  Three steps
    equal the twin bit for bit (state, sediment, terrain volumes and heights);
    each step is within the FP64 oracle's tolerances from the same state."""
    sc = contact_block(erosion=erosion)
    ctx, sid = gpu_context(sc)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    cell = float(np.float32(sc.terrain["cell"]))
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for step in range(3):
        ctx.step(1)
        gp, gv = ctx.state_f64(sid)
        gsed = ctx.sediment_fx(sid)
        gvol = ctx.terrain_volume_fx(sid).ravel()
        gh = ctx.terrain_heights(sid).ravel()
        tw = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
        assert np.array_equal(gp, tw["pos"]), (step, np.abs(gp - tw["pos"]).max())
        assert np.array_equal(gv, tw["vel"]), (step, np.abs(gv - tw["vel"]).max())
        assert np.array_equal(gsed, tw["sed"]) and np.array_equal(gvol, tw["vol"].ravel())
        assert np.array_equal(gh, tw["height"].ravel())
        # the branches fired: contact with the terrain and the exchange
        hs = bilinear_heights(st["height"].reshape(sc.terrain["nz"], sc.terrain["nx"]), cell, gp[:, 0], gp[:, 2])
        collided = int((np.abs(gp[:, 1] - hs) < 2e-6).sum())
        exchanged = int((gsed != st["sed"]).sum())
        assert collided > 0 and exchanged >= 0.1 * sc.n, (step, collided, exchanged)
        # FP64 oracle from the same (FP32) state
        r64 = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=8)
        assert np.abs(gp - r64["pos"]).max() <= 1e-6
        assert np.abs(gv - r64["vel"]).max() <= 1e-4 * np.abs(r64["vel"]).max()
        assert (np.abs(tw["rho"] - r64["rho"]) / r64["rho"]).max() <= DENSITY_TOL
        dsed64 = r64["sed"] - st["sed"]
        assert np.abs(gsed - r64["sed"]).max() <= 1e-4 * np.abs(dsed64).max()
        dvol64 = r64["vol"].ravel() - st["vol"].ravel()
        assert np.abs(gvol - r64["vol"].ravel()).max() <= 1e-4 * np.abs(dvol64).max()
        st = dict(pos=gp, vel=gv, sed=gsed, vol=gvol, height=gh)
    t0, s0, u0 = ctx.mass(sid)
    assert s0 > 0


def test_terrain_contact_atomic_equals_two_phase():
    a, sa = gpu_context(contact_block(erosion=sphe.EROSION_ATOMIC))
    b, sb = gpu_context(contact_block(erosion=sphe.EROSION_TWO_PHASE))
    _, _, u0 = a.mass(sa)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.terrain_volume_fx(sa), b.terrain_volume_fx(sb))
    assert np.array_equal(a.sediment_fx(sa), b.sediment_fx(sb))
    assert np.array_equal(a.state_f64(sa)[0], b.state_f64(sb)[0])
    assert a.mass(sa)[2] == u0 == b.mass(sb)[2]


# ------------------------------------------------------ reference fixtures

def test_f2_coincident_pairs_through_fluid_system():
    """Initialize(125) + AddParticles(125) puts 125 coincident pairs on the
    lattice (the dist < 1e-4 branch of gradPressure, fluid_system.h:438-440).
    One Run at dt = 0.01 through the FluidSystemSPH mirror: against the
    reference's own step (f2_coincident.npz) and bit for bit against the twin,
    GetParticle records included."""
    ref = recs(np.load(os.path.join(GOLD, "f2_coincident.npz"))["step1"])
    fs = sphe.FluidSystemSPH()
    fs.SetDeltaTime(0.01)
    fs.Initialize(125)
    fs.AddParticles(125)
    assert fs.NumParticles() == 250
    ids = fs.ctx.ids(fs.scene)
    pos0, _ = fs.ctx.state_f64(fs.scene)
    fs.Run()
    pos, vel = fs.ctx.state_f64(fs.scene)
    rho = fs.ctx.density(fs.scene)
    assert (np.abs(rho - ref["Density"]) / ref["Density"]).max() <= DENSITY_TOL
    assert np.abs(pos - ref["Position"]).max() <= 1e-6
    assert np.abs(vel - ref["Velocity"]).max() <= 1e-4 * np.abs(ref["Velocity"]).max()
    n = len(ids)
    tw = O.orc_step(orc_scene_from(sphe.default_params(dt=0.01)), pos0, np.zeros((n, 3)), np.zeros(n, np.int64),
                    ids, real=4, comp=1)
    assert np.array_equal(pos, tw["pos"]) and np.array_equal(vel, tw["vel"])
    pf_scale = np.abs(ref["PressureForce"]).max()
    for pid in range(0, 250, 7):
        rec = fs.GetParticle(pid)
        r = ref[pid]
        assert rec.Id == pid and rec.NeighbId == r["NeighbId"] == tw["nbmax"][pid]
        assert abs(rec.Density - r["Density"]) <= DENSITY_TOL * r["Density"]
        assert np.abs(np.array(rec.PressureForce[:]) - r["PressureForce"]).max() <= 1e-5 * pf_scale
        assert f32_bits_equal(rec.PressureForce[:], tw["pforce"][pid])
        assert f32_bits_equal(rec.Acceleration[:], tw["acc"][pid])
        assert f32_bits_equal([rec.Density], [tw["rho"][pid]])


def test_f4_box_cases_bit_exact():
    """Isolated particles crossing one, two or three walls, resting on a wall,
    and at rest outside the box (NaN velocity, fluid_system.h:345): one step
    of the library equals the reference's step (f4_box.npz) bit for bit."""
    g = np.load(os.path.join(GOLD, "f4_box.npz"))
    n = len(g["in_pos"])
    ctx = sphe.Context()
    sid = ctx.add_scene(sphe.default_params(dt=0.01))
    ctx.set_particles(sid, g["in_pos"], vel=g["in_vel"])
    ctx.step(1)
    pos, vel = ctx.state_f64(sid)
    assert f32_bits_equal(pos, g["Position"])
    assert f32_bits_equal(vel, g["Velocity"])
    assert np.isnan(vel[11]).all() and np.isfinite(vel[:11]).all()
    assert ctx.status() & 1
    for pid in range(n):
        rec = ctx.particle(sid, pid)
        assert f32_bits_equal(rec.Acceleration[:], g["Acceleration"][pid])
        assert f32_bits_equal([rec.Density], [g["Density"][pid]])
        assert rec.NeighbId == g["NeighbId"][pid] == -1


def test_lattice_initialize_add_reset_bit_exact():
    """Initialize / AddParticles / Reset (fluid_system.h:74-102, 232-259)
    place the reference's lattice bit for bit (lattice.npz), ids included."""
    g = np.load(os.path.join(GOLD, "lattice.npz"))
    for name, (n, add) in {"l1000": (1000, 0), "l100": (100, 0), "l125_125": (125, 125)}.items():
        fs = sphe.FluidSystemSPH()
        fs.Initialize(n)
        if add:
            fs.AddParticles(add)
        assert np.array_equal(fs.ctx.ids(fs.scene), g[f"{name}_Id"])
        assert f32_bits_equal(fs.ctx.positions(fs.scene), g[f"{name}_Position"])
        assert f32_bits_equal(fs.ctx.velocities(fs.scene), g[f"{name}_Velocity"])
    fs = sphe.FluidSystemSPH()
    fs.SetDeltaTime(0.01)
    fs.Initialize(1000)
    fs.Run(steps=3)
    fs.AddParticles(125)
    assert fs.NumParticles() == 1125
    fs.Reset()  # back to Initialize's lattice, ids from 0
    assert np.array_equal(fs.ctx.ids(fs.scene), g["l1000_Id"])
    assert f32_bits_equal(fs.ctx.positions(fs.scene), g["l1000_Position"])
    assert not np.any(fs.ctx.velocities(fs.scene))


def test_get_particle_force_fields_vs_f3():
    """The jittered 4,096-particle block of f3_jitter.npz, one step:
    GetParticle's PressureForce, SurfaceNormal, Acceleration, Density and
    NeighbId against the reference's step, and the whole record bit for bit
    against the twin's force fields."""
    g = np.load(os.path.join(GOLD, "f3_jitter.npz"))
    n = len(g["in_pos"])
    p = sphe.default_params(dt=1e-3, box_half=0.5, rest_density=1000.0, gravity=(0.0, -9.8, 0.0))
    ctx = sphe.Context()
    sid = ctx.add_scene(p)
    ctx.set_particles(sid, g["in_pos"], vel=g["in_vel"])
    ctx.step(1)
    tw = O.orc_step(orc_scene_from(p), g["in_pos"].astype(np.float64), g["in_vel"].astype(np.float64),
                    np.zeros(n, np.int64), np.arange(n, dtype=np.int32), real=4, comp=1)
    scale = {k: np.abs(g[k]).max() for k in ("PressureForce", "SurfaceNormal", "Acceleration")}
    for pid in range(0, n, 37):
        rec = ctx.particle(sid, pid)
        assert rec.Id == pid and rec.NeighbId == g["NeighbId"][pid]
        assert abs(rec.Density - g["Density"][pid]) <= DENSITY_TOL * g["Density"][pid]
        for k in ("PressureForce", "SurfaceNormal", "Acceleration"):
            got = np.array(getattr(rec, k)[:], dtype=np.float64)
            assert np.abs(got - g[k][pid]).max() <= 1e-5 * scale[k], (pid, k)
        assert np.abs(np.array(rec.Position[:]) - g["Position"][pid]).max() <= 1e-6
        for k, t in (("PressureForce", "pforce"), ("ViscosityForce", "vforce"), ("SurfaceForce", "sforce"),
                     ("SurfaceNormal", "normal"), ("Acceleration", "acc")):
            assert f32_bits_equal(getattr(rec, k)[:], tw[t][pid]), (pid, k)
        assert f32_bits_equal(rec.Position[:], tw["pos"][pid]) and f32_bits_equal(rec.Velocity[:], tw["vel"][pid])


def test_get_particle_sparse_ids_equal_dense_ids():
    """Ids far apart (no id map: the record's slot is searched) give the same
    records as dense ids with the same relative order; an unknown id fails."""
    sc = scenes.cfg1(n=2000)
    a, sa = gpu_context(sc)
    b = sphe.Context()
    sb = b.add_scene(sc.params, sc.terrain)
    b.set_particles(sb, sc.pos, sc.vel, ids=sc.ids.astype(np.int64).astype(np.int32) * 50000 + 3)
    a.step(2)
    b.step(2)
    for pid in (0, 17, 999, 1999):
        ra, rb = a.particle(sa, pid), b.particle(sb, pid * 50000 + 3)
        for f in ("Position", "Velocity", "Acceleration", "PressureForce", "SurfaceNormal", "Density"):
            va = getattr(ra, f)
            vb = getattr(rb, f)
            assert f32_bits_equal(va[:] if f != "Density" else [va], vb[:] if f != "Density" else [vb]), f
    for bad in (-1, 2000, 10 ** 9):
        with pytest.raises(sphe.SpheError):
            a.particle(sa, bad)
    with pytest.raises(sphe.SpheError):
        b.particle(sb, 4)


# ---------------------------------------------------------- heightfield ingest

def test_heightfield_ingest_is_the_reference_grid_mesh():
    """LoadHeightfield on the image of mesh_ref.npz (the 24 x 24 corner the
    reference's Grid(24, 24, 24) meshes) and GetSurfaceParts / GetIndices equal
    the reference's Grid::UpdateGrid / genIndices output bit for bit."""
    g = np.load(os.path.join(GOLD, "mesh_ref.npz"))
    img = np.ascontiguousarray(g["h_xz"])  # img[x, z] = GetHeightfieldAt(x, z)
    fs = sphe.FluidSystemSPH(params=sphe.default_params(dt=0.0))
    fs.LoadHeightfield(img, scale=1.0, cell=1.0, origin=(-0.5, 0.0, -0.5))
    verts = fs.GetSurfaceParts()
    assert np.array_equal(verts.view(np.uint32), g["verts"].view(np.uint32))
    assert np.array_equal(fs.GetIndices(), g["indices"])


def test_heightfield_ingest_non_square_image():
    """A 40-row by 24-column image: terrain cell (x, z) is pixel (row x,
    column z) as Grid::GetHeightfieldAt reads it (map[dimY * x + y] with dimY
    the row length, grid.h:104-107), so the terrain is 40 cells along x and 24
    along z; the mesh follows the oracle's restatement."""
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, size=(40, 24), dtype=np.uint8)
    flat = img.reshape(-1)  # the row-major bytes stbi_load returns
    ctx = sphe.Context()
    sid = ctx.add_scene(sphe.default_params(dt=0.0))
    ctx.set_terrain(sid, dict(nx=1, nz=1, origin=(-0.2, -0.1, -0.2), cell=0.01, heights=None))
    scale = np.float32(0.001)
    ctx.terrain_load_u8(sid, img, float(scale))
    assert ctx.terrain_dims(sid) == (40, 24)
    H = ctx.terrain_heights(sid)  # [z, x]
    want = np.empty((24, 40), dtype=np.float32)
    for x in range(40):
        for z in range(24):
            want[z, x] = np.float32(-0.1) + scale * np.float32(flat[24 * x + z])
    # heights come back from the fixed-point volumes: within a few float ulp
    assert np.abs(H - want).max() <= 1e-6
    assert H[5, 30] == pytest.approx(-0.1 + 0.001 * img[30, 5], abs=1e-6)
    verts, idx = ctx.terrain_mesh(sid)
    ref = O.terrain_mesh(H, np.float32(0.01), (-0.2, -0.1, -0.2))
    assert np.array_equal(verts.reshape(-1).view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(idx, O.terrain_indices(40, 24))


# ------------------------------------------------------- runtime robustness

def test_list_capacity_without_memory_keeps_running():
    """Lists too large for the device: set_list_capacity fails with ENOMEM,
    and the following steps run (the failed allocation leaves no error behind)
    and equal a run that never asked."""
    sc = scenes.cfg2()
    a, sa = gpu_context(sc)
    b, sb = gpu_context(sc)
    a.step(2)
    b.step(2)
    with pytest.raises(sphe.SpheError, match="ENOMEM"):
        a.set_list_capacity(1 << 20)
    a.step(3)
    b.step(3)
    assert a.status() == 0
    pa, va = a.state_f64(sa)
    pb, vb = b.state_f64(sb)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_param_pointer_survives_added_scenes():
    """sphe_param_ptr storage stays put when more scenes are added: a write
    through a pointer taken before reaches the scene's parameters."""
    import ctypes as C
    ctx = sphe.Context()
    s0 = ctx.add_scene(sphe.default_params())
    pm = sphe.lib().sphe_param_ptr(ctx._p, s0, 0)  # SPHE_P_MASS
    pg = sphe.lib().sphe_param_ptr(ctx._p, s0, 4)  # SPHE_P_GRAV
    for _ in range(200):
        ctx.add_scene(sphe.default_params())
    pm[0] = 0.031
    pg[1] = -3.0
    p = ctx.get_params(s0)
    assert p.mass == np.float32(0.031) and p.gravity[1] == -3.0
    assert C.addressof(pm.contents) == C.addressof(sphe.lib().sphe_param_ptr(ctx._p, s0, 0).contents)


def test_negative_volumes_restore():
    """A checkpoint whose terrain went below its base (negative fixed-point
    volumes) restores and continues bit for bit."""
    sc = contact_block()
    a, sa = gpu_context(sc)
    vol = a.terrain_volume_fx(sa)
    vol[:, :3] = -vol[:, :3] // 3
    a.set_terrain_volume_fx(sa, vol)
    assert np.array_equal(a.terrain_volume_fx(sa), vol)
    H = a.terrain_heights(sa)
    assert (H[:, :3] < 0).all()
    a.step(2)
    p, v = a.state_f64(sa)
    b, sb = gpu_context(sc)
    b.set_particles(sb, p, vel=v, ids=sc.ids, sed=a.sediment_fx(sa))
    b.set_terrain_volume_fx(sb, a.terrain_volume_fx(sa))
    a.step(2)
    b.step(2)
    assert np.array_equal(a.state_f64(sa)[0], b.state_f64(sb)[0])
    assert np.array_equal(a.terrain_volume_fx(sa), b.terrain_volume_fx(sb))


def test_inspection_keeps_the_steps_order_after_tuning_changes():
    """GetParticle after the slab or lane count changed still repeats the
    sums of the step that ran (its record equals the twin's for that step)."""
    sc = scenes.cfg1(n=3000)
    ctx, sid = gpu_context(sc)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    ctx.step(1)
    tw = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=4, comp=1)
    ctx.set_subcells(2)
    ctx.set_lanes(8)
    for pid in (5, 1234, 2999):
        rec = ctx.particle(sid, pid)
        assert f32_bits_equal(rec.Acceleration[:], tw["acc"][pid])
        assert f32_bits_equal(rec.PressureForce[:], tw["pforce"][pid])


def test_mass_all_equals_per_scene_mass():
    """sphe_mass_all (every scene in one launch) gives each scene's terrain
    and sediment totals and unit sums; the unit sums equal the per-scene
    reductions exactly and the FP64 totals within rounding of their order."""
    scs = [scenes.cfg1(seed_idx=k, n=1000 + 900 * k) for k in range(6)]
    ctx = sphe.Context()
    sids = [ctx.add_synthetic(s) for s in scs]
    ctx.step(20)
    t, s, u = ctx.mass_all()
    for k, sid in enumerate(sids):
        t1, s1, u1 = ctx.mass(sid)
        assert u[k] == u1
        assert abs(t[k] - t1) <= 1e-12 * t1 and abs(s[k] - s1) <= 1e-12 * max(s1, 1e-30)
    assert s.sum() > 0
    t2, s2, u2 = ctx.mass_all()
    assert np.array_equal(t, t2) and np.array_equal(s, s2) and np.array_equal(u, u2)
