"""GPU parity: libsphe (HIP, gfx950) against the oracle on the same inputs.

Bars:
  * neighbour sets and cell keys: bit-exact;
  * FP32 step: bit-exact against the oracle's same-order FP32 twin; density
    within 1e-5 relative of the FP64 oracle;
  * FP64 step: within 1e-12 relative of the FP64 oracle;
  * erosion: exact integer conservation, atomic == two-phase bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import sphe
from gpu_common import gpu_context, oracle_inputs, orc_scene_from, small_dam
from sphe import scenes

pytestmark = pytest.mark.gpu

DENSITY_TOL = 1e-5  # relative density error of the FP32 path vs FP64


def test_constants_match_oracle():
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    got = ctx.debug_consts(sid)
    c = O.derive(orc_scene_from(sc.params, sc.terrain, ctx.fixed_exponent(sid)))
    for k in ("h", "h2", "thr", "t3h2", "mc6", "pf_scale", "vf_scale", "ng", "cell", "inv_cell", "hq",
              "kcv", "up", "dn"):
        assert np.float32(got[k]) == np.float32(getattr(c, k)), k
    assert [int(got["dim_x"]), int(got["dim_y"]), int(got["dim_z"])] == list(c.dims)


def test_cell_keys_bit_exact():
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    keys = ctx.debug_cell_keys(sid)
    c = O.derive(orc_scene_from(sc.params, sc.terrain))
    ref = O.cell_keys(c, sc.pos.astype(np.float32))
    assert np.array_equal(keys, ref)


@pytest.mark.parametrize("which", ["dam4096", "ref_lattice", "coincident"])
def test_neighbour_sets_bit_exact(which):
    if which == "dam4096":
        sc = small_dam()
        ctx, sid = gpu_context(sc)
        pos = sc.pos.astype(np.float32)
        thr = O.derive(orc_scene_from(sc.params)).thr
    else:
        ctx = sphe.Context()
        sid = ctx.add_scene(sphe.default_params(dt=0.01))
        fs_pos = O.lattice(1000) if which == "ref_lattice" else np.concatenate(
            [O.lattice(125), O.lattice(125, first_id=125)])
        pos = fs_pos["Position"]
        ctx.set_particles(sid, pos, ids=fs_pos["Id"])
        thr = O.pred_threshold(0.0457)
    counts, off, idx = ctx.debug_neighbours(sid)
    bc, boff, bidx = O.neighbours_brute(pos, thr)
    assert np.array_equal(counts, bc)
    assert np.array_equal(off, boff)
    assert np.array_equal(idx, bidx)


@pytest.mark.parametrize("precision,comp", [(sphe.FP32, 1), (sphe.FP32_NAIVE, 0)])
def test_density_vs_fp64_and_twin(precision, comp):
    sc = small_dam()
    ctx, sid = gpu_context(sc, precision)
    rho = ctx.debug_density(sid)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    ref64 = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=8)["rho"]
    rel = np.abs(rho - ref64) / ref64
    assert rel.max() <= DENSITY_TOL, rel.max()
    twin = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=4, comp=comp)["rho"]
    assert np.array_equal(rho.astype(np.float32), twin.astype(np.float32))


@pytest.mark.parametrize("erosion,precision,comp", [(sphe.EROSION_OFF, sphe.FP32, 1),
                                                     (sphe.EROSION_ATOMIC, sphe.FP32, 1),
                                                     (sphe.EROSION_TWO_PHASE, sphe.FP32, 1),
                                                     (sphe.EROSION_ATOMIC, sphe.FP32_NAIVE, 0),
                                                     (sphe.EROSION_ATOMIC, sphe.FP32_F64ACC, 2)])
def test_step_fp32_bit_exact_twin(erosion, precision, comp):
    """Three steps equal the twin's bit for bit (positions, velocities,
    sediment, terrain volumes and heights): compensated FP32, plain FP32, and
    FP32 terms with FP64 force sums."""
    sc = small_dam(erosion=erosion)
    ctx, sid = gpu_context(sc, precision)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    ctx.step(3)
    gp, gv = ctx.state_f64(sid)
    gsed = ctx.sediment_fx(sid)
    gvol = ctx.terrain_volume_fx(sid)
    gh = ctx.terrain_heights(sid)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(3):
        out = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=comp)
        st = out
    assert np.array_equal(gp, st["pos"]), np.abs(gp - st["pos"]).max()
    assert np.array_equal(gv, st["vel"]), np.abs(gv - st["vel"]).max()
    assert np.array_equal(gsed, st["sed"])
    assert np.array_equal(gvol, st["vol"].reshape(gvol.shape))
    assert np.array_equal(gh, st["height"].reshape(gh.shape))
    if erosion != sphe.EROSION_OFF:
        assert gsed.sum() > 0  # the exchange did happen


def test_step_fp32_vs_fp64_oracle():
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    ctx.step(1)
    gp, gv = ctx.state_f64(sid)
    ref = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=8)
    rho = ctx.density(sid)
    assert (np.abs(rho - ref["rho"]) / ref["rho"]).max() <= DENSITY_TOL
    dv = np.abs(gv - ref["vel"]).max()
    vscale = np.abs(ref["vel"]).max()
    assert dv <= 1e-4 * vscale, (dv, vscale)


@pytest.mark.parametrize("yz", [2, 1, 3])
def test_fp64_path_vs_oracle(yz):
    """The plain FP64 path against the FP64 oracle on the same grid (every
    y/z split: the FP64 kernels have their own row and stencil code)."""
    sc = small_dam()
    ctx, sid = gpu_context(sc, sphe.FP64)
    ctx.set_yz_split(yz)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.yz = yz
    ctx.step(2)
    gp, gv = ctx.state_f64(sid)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=8)
    assert np.abs(gp - st["pos"]).max() <= 1e-12
    assert np.abs(gv - st["vel"]).max() <= 1e-9 * max(1.0, np.abs(st["vel"]).max())


def test_reference_lattice_vs_reference_step():
    """The reference app's scene (Initialize(1000), dt = 0.01, default
    parameters) against ref32, the bit-exact restatement of the reference."""
    fs = sphe.FluidSystemSPH()
    fs.SetDeltaTime(0.01)
    fs.Initialize(1000)
    fs.Run()
    p = O.default_params(dt=0.01)
    R = O.ref32_step(O.lattice(1000), p, 1)
    pos = fs.ctx.positions(fs.scene)
    rho = fs.ctx.density(fs.scene)
    assert (np.abs(rho - R["Density"]) / R["Density"]).max() <= DENSITY_TOL
    assert np.abs(pos - R["Position"]).max() <= 1e-6
    for pid in (0, 555, 999):
        rec = fs.GetParticle(pid)
        r = R[pid]
        assert rec.Id == pid
        assert rec.NeighbId == r["NeighbId"]
        assert abs(rec.Density - r["Density"]) <= 1e-5 * r["Density"]
        acc = np.array(rec.Acceleration[:])
        assert np.abs(acc - r["Acceleration"]).max() <= 1e-4 * max(1.0, np.abs(r["Acceleration"]).max())


def test_mass_conservation_exact():
    sc = small_dam(erosion=sphe.EROSION_ATOMIC)
    ctx, sid = gpu_context(sc)
    t0, s0, u0 = ctx.mass(sid)
    ctx.step(50)
    t1, s1, u1 = ctx.mass(sid)
    assert u1 == u0  # exact fixed-point units
    assert s1 > 0
    assert abs((t1 + s1) - (t0 + s0)) <= 1e-9 * (t0 + s0)


@pytest.mark.parametrize("precision", [sphe.FP32, sphe.FP64])
def test_atomic_equals_two_phase(precision):
    """Atomic and two-phase erosion give the same bits over 20 steps (integer
    exchanges: their order does not matter), in FP32 and in FP64."""
    a = small_dam(erosion=sphe.EROSION_ATOMIC)
    b = small_dam(erosion=sphe.EROSION_TWO_PHASE)
    ca, sa = gpu_context(a, precision)
    cb, sb = gpu_context(b, precision)
    ca.step(20)
    cb.step(20)
    assert np.array_equal(ca.terrain_volume_fx(sa), cb.terrain_volume_fx(sb))
    assert np.array_equal(ca.sediment_fx(sa), cb.sediment_fx(sb))
    pa, _ = ca.state_f64(sa)
    pb, _ = cb.state_f64(sb)
    assert np.array_equal(pa, pb)


def test_deterministic_and_graph_equals_direct():
    sc = small_dam()
    c1, s1 = gpu_context(sc, graph=True)
    c2, s2 = gpu_context(sc, graph=False)
    c1.step(10)
    c2.step(10)
    p1, v1 = c1.state_f64(s1)
    p2, v2 = c2.state_f64(s2)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
    assert np.array_equal(c1.terrain_volume_fx(s1), c2.terrain_volume_fx(s2))


def test_batch_equals_single_scenes():
    scs = [scenes.dam_break_scene(7 + i, n, frac, T, steps=5)
           for i, (n, frac, T) in enumerate([(3000, 0.3, 64), (1100, 0.45, 128), (5000, 0.2, 192)])]
    batch = sphe.Context()
    ids = [batch.add_synthetic(s) for s in scs]
    batch.step(5)
    for s, sid in zip(scs, ids):
        single, ss = gpu_context(s)
        single.step(5)
        pb, vb = batch.state_f64(sid)
        ps, vs = single.state_f64(ss)
        assert np.array_equal(pb, ps) and np.array_equal(vb, vs)
        assert np.array_equal(batch.terrain_volume_fx(sid), single.terrain_volume_fx(ss))


@pytest.mark.parametrize("lanes,precision", [(1, sphe.FP32), (4, sphe.FP32), (8, sphe.FP32_NAIVE),
                                             (4, sphe.FP64)])
def test_persistent_passes_change_no_bits(lanes, precision):
    """Density and force as persistent launches over the work queues give the
    same bits as the ordinary grids, on a batch with an empty scene, atomic
    and two-phase erosion, in graph sequences and direct launches."""
    spec = [(3000, 0.3, 64, sphe.EROSION_ATOMIC), None, (700, 0.45, 128, sphe.EROSION_TWO_PHASE),
            (5000, 0.2, 192, sphe.EROSION_ATOMIC), (64, 0.3, 64, sphe.EROSION_ATOMIC)]
    runs = []
    for on, graph in ((False, True), (True, True), (True, False)):
        ctx = sphe.Context(precision)
        ctx.set_lanes(lanes)
        ctx.set_persistent(on)
        ctx.use_graph(graph)
        assert ctx.get_persistent() == on
        sids = []
        for k, sp in enumerate(spec):
            if sp is None:
                ctx.add_scene(sphe.default_params())
                continue
            n, frac, T, ero = sp
            sids.append(ctx.add_synthetic(scenes.dam_break_scene(70 + k, n, frac, T, erosion=ero, steps=40)))
        ctx.step(40)
        assert ctx.status() & 3 == 0
        runs.append([ctx.state_f64(s) + (ctx.sediment_fx(s), ctx.terrain_volume_fx(s)) for s in sids])
        ctx.close()
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_save_restore_repeats_the_same_steps():
    """A restored state steps to the same bits, also after the lists grew."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    ctx.step(5)
    ctx.save_state()
    ctx.step(20)
    p1, v1 = ctx.state_f64(sid)
    s1, vol1 = ctx.sediment_fx(sid), ctx.terrain_volume_fx(sid)
    ctx.restore_state()
    ctx.step(20)
    p2, v2 = ctx.state_f64(sid)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
    assert np.array_equal(s1, ctx.sediment_fx(sid)) and np.array_equal(vol1, ctx.terrain_volume_fx(sid))


def test_autotune_restores_state_and_counts():
    """Autotune steps re-uploaded copies of the state; the cell counters are
    cleared before every recount, and the tuned result equals an untuned run
    (regression test of stale counts after a re-upload)."""
    sc = small_dam()
    c1, s1 = gpu_context(sc)
    c1.profile(True)
    c1.step(1)
    blk, rm = c1.autotune(2)
    assert blk in (64, 128, 256, 512) and c1.get_force_block() in (64, 128, 256, 512)
    # the caller's profile is kept: one step before autotune, none of its trials
    assert c1.profile_read("density")[1] == 1
    c1.profile(False)
    c1.step(5)
    c2, s2 = gpu_context(sc)
    c2.step(6)
    assert c1.status() == 0 and c2.status() == 0
    p1, v1 = c1.state_f64(s1)
    p2, v2 = c2.state_f64(s2)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)


@pytest.mark.parametrize("block,remap,fblock", [(64, 0, 0), (64, 32, 512), (128, 8, 64), (256, 3, 0),
                                                (512, 1, 128)])
def test_launch_shape_does_not_change_results(block, remap, fblock):
    sc = small_dam()
    c1, s1 = gpu_context(sc)
    c1.set_tuning(block, remap)
    c1.set_force_block(fblock)
    assert c1.get_force_block() == (fblock or block)
    c1.step(4)
    c2, s2 = gpu_context(sc)
    c2.step(4)
    p1, _ = c1.state_f64(s1)
    p2, _ = c2.state_f64(s2)
    assert c1.status() == 0
    assert np.array_equal(p1, p2)


@pytest.mark.parametrize("lanes", [1, 2, 8])
def test_lane_split_bit_exact_twin(lanes):
    """Each lane count is its own summation order; the twin with the same
    split reproduces the GPU step bit for bit."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    ctx.set_lanes(lanes)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.lanes = lanes
    ctx.step(2)
    gp, gv = ctx.state_f64(sid)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    assert np.array_equal(ctx.sediment_fx(sid), st["sed"])


@pytest.mark.parametrize("slabs,yz", [(1, 2), (2, 2), (1, 3), (2, 1), (1, 1)])
def test_slab_split_bit_exact_twin(slabs, yz):
    """The x-slab count sets the canonical order and the clipped rows; the
    twin with the same count and y/z split reproduces the GPU steps bit for
    bit."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    ctx.set_subcells(slabs)
    ctx.set_yz_split(yz)
    assert ctx.get_subcells() == slabs
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.slabs = slabs
    osc.yz = yz
    ctx.step(2)
    gp, gv = ctx.state_f64(sid)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    assert np.array_equal(ctx.sediment_fx(sid), st["sed"])


@pytest.mark.parametrize("yz,lanes", [(1, 4), (3, 4), (3, 8), (3, 1), (1, 8)])
def test_yz_split_bit_exact_twin(yz, lanes):
    """The y/z cell split sets the grid, the canonical order and the clipped
    rows; the twin with the same split reproduces the GPU steps bit for bit
    (the default split is covered by every other twin test).  Lane counts
    other than 4 walk the rows in rounds of their own width: 8 lanes step
    past a whole stencil row of split 3 (7 rows) at once."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    ctx.set_yz_split(yz)
    ctx.set_lanes(lanes)
    assert ctx.get_yz_split() == yz
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.yz = yz
    osc.lanes = lanes
    counts, off, idx = ctx.debug_neighbours(sid)
    bc, boff, bidx = O.neighbours_brute(sc.pos.astype(np.float32), O.derive(osc).thr)
    assert np.array_equal(counts, bc) and np.array_equal(idx, bidx)
    ctx.step(2)
    gp, gv = ctx.state_f64(sid)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    assert np.array_equal(ctx.sediment_fx(sid), st["sed"])


def test_autotuned_grid_steps_bit_exact_with_the_twin():
    """The grid autotuner leaves the state as it was, and its choice of
    split and slabs steps bit-exact with the twin configured the same way;
    a second call takes the cached choice."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    yz, slabs, tile = ctx.autotune_grid(2)
    assert yz in (1, 2, 3) and slabs in (1, 2, 4) and 11 <= tile <= 13
    gp0, gv0 = ctx.state_f64(sid)
    assert np.array_equal(gp0, pos) and np.array_equal(gv0, vel)
    assert ctx.autotune_grid(2) == (yz, slabs, tile)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.yz = yz
    osc.slabs = slabs
    ctx.step(2)
    gp, gv = ctx.state_f64(sid)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    assert np.array_equal(ctx.sediment_fx(sid), st["sed"])


def test_grid_tuner_skips_a_split_with_too_many_cells():
    """A box 402 cells long in x: split 3 would need 402 x 1204^2 = 5.8e8
    cells, past the 4e8 a scene may hold, so setting it fails and leaves the
    grid as it was, and the grid tuner leaves it out instead of failing.  Its
    choice among splits 1 and 2 (up to 2.6e8 cells) steps bit-exact with the
    twin configured the same way."""
    base = small_dam(n=1024)
    prm = base.params
    prm.box_half = 9.2
    sc = scenes.Scene("big_box_1024", prm, None, base.pos, base.vel, base.ids, 2)
    ctx, sid = gpu_context(sc)
    split0 = ctx.get_yz_split()
    with pytest.raises(sphe.SpheError, match="ERANGE"):
        ctx.set_yz_split(3)
    assert ctx.get_yz_split() == split0
    yz, slabs, tile = ctx.autotune_grid(1)
    assert yz in (1, 2) and slabs in (1, 2, 4) and 11 <= tile <= 13
    pos, vel, sed, _, _, _ = oracle_inputs(ctx, sid, sc)
    gp0, gv0 = ctx.state_f64(sid)
    assert np.array_equal(gp0, pos) and np.array_equal(gv0, vel)
    osc = orc_scene_from(sc.params)
    osc.yz = yz
    osc.slabs = slabs
    ctx.step(2)
    assert ctx.status() & 3 == 0
    gp, gv = ctx.state_f64(sid)
    st = dict(pos=pos, vel=vel, sed=sed)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, real=4, comp=1)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    ctx.close()


def test_scan_tile_never_changes_results():
    """Scan tiles of 2^11 and 2^13 cells, set before the first step or between
    steps (the tile flags and fill words are reset then), step a batch to
    the same bits as the default tile."""
    scs = [scenes.dam_break_scene(60 + i, n, frac, T, steps=12)
           for i, (n, frac, T) in enumerate([(3000, 0.3, 64), (1100, 0.45, 128), (5000, 0.2, 192)])]
    runs = []
    for tile, when in ((12, 0), (11, 0), (13, 0), (11, 5), (13, 5)):
        ctx = sphe.Context()
        sids = [ctx.add_synthetic(s) for s in scs]
        if when == 0:
            ctx.set_scan_tile(tile)
        ctx.step(5)
        if when:
            ctx.set_scan_tile(tile)
        assert ctx.get_scan_tile() == tile
        ctx.step(7)
        assert ctx.status() & 3 == 0
        runs.append([ctx.state_f64(s) + (ctx.terrain_volume_fx(s),) for s in sids])
        ctx.close()
    for r in runs[1:]:
        for (p, v, t), (p0, v0, t0) in zip(r, runs[0]):
            assert np.array_equal(p, p0) and np.array_equal(v, v0) and np.array_equal(t, t0)


def test_yz_split_change_keeps_the_state():
    """Changing the split after steps rebuilds the layout from the pulled
    state: the following steps equal the twin's with the new split from that
    state, bit for bit (terrain, sediment and ids included)."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    ctx.set_yz_split(1)
    _, _, _, _, _, fx = oracle_inputs(ctx, sid, sc)
    ctx.step(3)
    p3, v3 = ctx.state_f64(sid)
    s3 = ctx.sediment_fx(sid)
    vol3 = ctx.terrain_volume_fx(sid)
    h3 = ctx.terrain_heights(sid)
    ctx.set_yz_split(2)
    assert ctx.get_yz_split() == 2
    ctx.step(2)
    gp, gv = ctx.state_f64(sid)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.yz = 2
    st = dict(pos=p3, vel=v3, sed=s3, vol=vol3.ravel(), height=h3.ravel())
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    assert np.array_equal(ctx.sediment_fx(sid), st["sed"])
    assert np.array_equal(ctx.terrain_volume_fx(sid).ravel(), st["vol"].ravel())


def test_crowded_cell_taken_whole():
    """A cell of 2300 particles (more than the 1023 the slab offsets can
    count) with probes along x through it: neighbour sets equal brute force
    and two steps equal the twin bit for bit."""
    rng = np.random.default_rng(21)
    h = np.float32(0.0457)
    cell = float(h) * (1 + 2.0 ** -8)
    org = -0.2
    fr = np.concatenate([rng.uniform(0.02, 0.22, 1300), rng.uniform(0.78, 0.98, 1000)])
    blob = org + cell * np.stack([2 + fr, 2 + rng.uniform(0.3, 0.45, 2300), 2 + rng.uniform(0.3, 0.45, 2300)], 1)
    probes = org + cell * np.stack([rng.uniform(0.3, 4.7, 400), np.full(400, 2.375), np.full(400, 2.375)], 1)
    pos = np.concatenate([blob, probes, rng.uniform(-0.19, 0.19, size=(1500, 3))]).astype(np.float32)
    n = len(pos)
    p = sphe.default_params(dt=1e-5)
    ctx = sphe.Context()
    sid = ctx.add_scene(p)
    ctx.set_particles(sid, pos)
    cnt, off, idx = O.neighbours_brute(pos, O.pred_threshold(h))
    gcnt, goff, gidx = ctx.debug_neighbours(sid)
    assert np.array_equal(gcnt, cnt)
    for i in range(0, n, 11):
        assert np.array_equal(np.sort(gidx[goff[i]:goff[i + 1]]), np.sort(idx[off[i]:off[i + 1]]))
    osc = orc_scene_from(p)
    ctx.step(2)
    st = dict(pos=pos.astype(np.float64), vel=np.zeros((n, 3)), sed=np.zeros(n, np.int64))
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], np.arange(n, dtype=np.int32), real=4, comp=1)
    gp, gv = ctx.state_f64(sid)
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])


def _blob(n=3000, seed=4):
    rng = np.random.default_rng(seed)
    pos = (0.5 + rng.normal(0, 0.012, size=(n, 3))).astype(np.float32)  # dense blob
    p = sphe.default_params(h=0.03, mass=0.001, rest_density=1000.0, dt=1e-4, box_center=(0.5, 0.5, 0.5),
                            box_half=0.5, gravity=(0.0, -9.8, 0.0))
    return pos, p


def test_list_overflow_paths_match_the_twin():
    """A blob with 32-entry lists: particles whose neighbours fit the list,
    continue in an overflow chunk, find the pool exhausted, or pass twice the
    capacity (stencil walk) all give the twin's bits."""
    pos, p = _blob()
    n = len(pos)
    ctx = sphe.Context()
    sid = ctx.add_scene(p)
    ctx.set_particles(sid, pos)
    ctx.set_list_capacity(32)
    assert ctx.get_list_capacity() == 32
    counts = ctx.debug_neighbours(sid, csr=False)[0] - 1  # listed: self excluded
    assert (counts <= 32).any() and ((counts > 32) & (counts <= 64)).any() and (counts > 64).any()
    osc = orc_scene_from(p)
    ctx.step(2)
    st = dict(pos=pos.astype(np.float64), vel=np.zeros((n, 3)), sed=np.zeros(n, np.int64))
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], np.arange(n, dtype=np.int32), real=4, comp=1)
    gp, gv = ctx.state_f64(sid)
    assert ctx.status() == 4  # particles past list and pool walked their stencil
    assert np.array_equal(gp, st["pos"]) and np.array_equal(gv, st["vel"])
    ctx.step(1)  # the overflow grew the lists (and the chunks) for the following steps
    cap = ctx.get_list_capacity()
    chunk, chunks = ctx.get_overflow_capacity()
    assert cap % 128 == 0 and cap >= counts.max() and cap + chunk >= counts.max() and chunks >= 64
    ctx.step(1)
    assert ctx.status() == 0  # every neighbourhood fits now


def test_overflow_chunks_past_32_bit_entry_offsets():
    """Chunk claims that start at pool entry 2^32 + 2^20 (entry offsets past
    32 bits; the chunks are addressed in units of 16 entries) give the bits
    and the sizes of claims from entry 0, on every path of the overflow test
    above: list, chunk, full pool and stencil walk, then grown lists."""
    pos, p = _blob()
    runs = []
    for origin in (0, (1 << 32) + (1 << 20)):
        ctx = sphe.Context()
        sid = ctx.add_scene(p)
        ctx.set_particles(sid, pos)
        ctx.set_list_capacity(32)
        assert sphe.lib().sphe_debug_pool_origin(ctx._p, origin) == 0
        ctx.step(2)
        first = (ctx.status(), ctx.state_f64(sid))
        ctx.step(2)
        runs.append((first, ctx.status(), ctx.state_f64(sid), ctx.get_list_capacity(), ctx.get_overflow_capacity()))
        ctx.close()
    (f0, st0, s0, cap0, oc0), (f1, st1, s1, cap1, oc1) = runs
    assert f0[0] == f1[0] == 4  # some particles found the pool full and walked
    for a, b in ((f0[1], f1[1]), (s0, s1)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert (st0, cap0, oc0) == (st1, cap1, oc1)


def test_list_overflow_paths_fp64():
    """The FP64 kernels on the overflow test's blob with 32-entry lists (list,
    chunk, full pool and stencil walk) against the FP64 oracle, and the same
    bits as with 1,024-entry lists (a different mix of the three paths)."""
    pos, p = _blob()
    n = len(pos)
    runs = []
    for cap in (32, 1024):
        ctx = sphe.Context(sphe.FP64)
        sid = ctx.add_scene(p)
        ctx.set_particles(sid, pos)
        ctx.set_list_capacity(cap)
        ctx.step(2)
        runs.append((ctx.status(), ctx.state_f64(sid)))
        ctx.close()
    assert runs[0][0] == 4  # particles past list and pool walked their stencil
    assert np.array_equal(runs[0][1][0], runs[1][1][0]) and np.array_equal(runs[0][1][1], runs[1][1][1])
    osc = orc_scene_from(p)
    st = dict(pos=pos.astype(np.float64), vel=np.zeros((n, 3)), sed=np.zeros(n, np.int64))
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], np.arange(n, dtype=np.int32), real=8)
    gp, gv = runs[0][1]
    assert np.abs(gp - st["pos"]).max() <= 1e-12
    assert np.abs(gv - st["vel"]).max() <= 1e-9 * max(1.0, np.abs(st["vel"]).max())


def test_list_capacity_never_changes_results():
    pos, p = _blob(2000, 7)
    out = []
    for cap in (48, 96, 1024):
        ctx = sphe.Context()
        sid = ctx.add_scene(p)
        ctx.set_particles(sid, pos)
        ctx.set_list_capacity(cap)
        ctx.step(3)
        out.append(ctx.state_f64(sid))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])


def test_status_flags_nonfinite():
    # a particle at rest outside the box: the reference's response divides by
    # |v| = 0 and yields NaN velocity (fluid_system.h:345)
    ctx = sphe.Context()
    sid = ctx.add_scene(sphe.default_params(dt=0.01, gravity=(0.0, 0.0, 0.0)))
    ctx.set_particles(sid, np.array([[0.3, 0.0, 0.0], [0.0, 0.0, 0.0]], dtype=np.float32))
    ctx.step(1)
    assert ctx.status() == 1
    v = ctx.velocities(sid)
    assert np.isnan(v[0]).all() and np.isfinite(v[1]).all()


def test_debug_accel_is_the_steps_acceleration():
    """sphe_debug_accel equals the twin's acceleration bit for bit, at the
    start and after steps, and FP64 is within 1e-9 of the FP64 oracle."""
    sc = small_dam()
    ctx, sid = gpu_context(sc)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    acc = ctx.debug_accel(sid)
    twin = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=4, comp=1)
    assert np.array_equal(acc, twin["acc"])
    ctx.step(3)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(3):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    acc3 = ctx.debug_accel(sid)
    nxt = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=1)
    assert np.array_equal(acc3, nxt["acc"])
    c64, s64 = gpu_context(sc, sphe.FP64)
    a64 = c64.debug_accel(s64)
    ref = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=8)["acc"]
    scale = np.abs(ref).max()
    assert np.abs(a64 - ref).max() <= 1e-9 * scale


@pytest.mark.parametrize("yz,lanes,precision,comp", [(3, 8, sphe.FP32, 1), (1, 1, sphe.FP32, 1),
                                                      (2, 4, sphe.FP32_F64ACC, 2), (3, 2, sphe.FP32_NAIVE, 0)])
def test_inspection_kernels_bit_exact_on_every_grid(yz, lanes, precision, comp):
    """The inspection kernels repeat the step's sums in its order on every
    split, lane count and FP32 mode: sphe_debug_accel (the next step's
    acceleration) and GetParticle (the last step's acceleration and density)
    equal the twin's bit for bit."""
    sc = small_dam()
    ctx, sid = gpu_context(sc, precision)
    ctx.set_yz_split(yz)
    ctx.set_lanes(lanes)
    pos, vel, sed, vol, height, fx = oracle_inputs(ctx, sid, sc)
    osc = orc_scene_from(sc.params, sc.terrain, fx)
    osc.yz = yz
    osc.lanes = lanes
    twin = O.orc_step(osc, pos, vel, sed, sc.ids, vol, height, real=4, comp=comp)
    assert np.array_equal(ctx.debug_accel(sid), twin["acc"])
    ctx.step(2)
    st = dict(pos=pos, vel=vel, sed=sed, vol=vol, height=height)
    for _ in range(2):
        st = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=comp)
    for pid in (0, 1234, 4095):
        rec = ctx.particle(sid, pid)
        assert np.array_equal(np.array(rec.Acceleration[:], dtype=np.float32), st["acc"][pid].astype(np.float32))
        assert np.float32(rec.Density) == np.float32(st["rho"][pid])
    nxt = O.orc_step(osc, st["pos"], st["vel"], st["sed"], sc.ids, st["vol"], st["height"], real=4, comp=comp)
    assert np.array_equal(ctx.debug_accel(sid), nxt["acc"])
    ctx.close()


def test_step_integrates_the_reported_acceleration():
    """v' = v + a dt in FP32 with a from sphe_debug_accel, bit for bit, for
    every particle the step leaves strictly inside the box (no terrain)."""
    ctx = sphe.Context()
    sid = ctx.add_scene(sphe.default_params(dt=0.01))
    lat = O.lattice(1000, origin=(0.02, 0.0, 0.0))
    ctx.set_particles(sid, lat["Position"], ids=lat["Id"])
    ctx.step(2)
    p0, v0 = ctx.state_f64(sid)
    a = ctx.debug_accel(sid).astype(np.float32)
    ctx.step(1)
    p1, v1 = ctx.state_f64(sid)
    dt = np.float32(0.01)
    want = v0.astype(np.float32) + a * dt
    inside = (np.abs(p1) < 0.2).all(axis=1)
    assert inside.sum() > 500
    assert np.array_equal(v1[inside].astype(np.float32), want[inside])


def test_checkpoint_continues_bit_for_bit():
    """A scene's state (positions, velocities, sediment, ids) and terrain
    volumes moved into a new context continue the run bit for bit."""
    sc = small_dam()
    a, sa = gpu_context(sc)
    a.step(3)
    p, v = a.state_f64(sa)
    sed = a.sediment_fx(sa)
    vol = a.terrain_volume_fx(sa)
    assert sed.any()
    b, sb = gpu_context(sc)
    b.set_particles(sb, p, vel=v, ids=sc.ids, sed=sed)
    b.set_terrain_volume_fx(sb, vol)
    assert np.array_equal(b.terrain_volume_fx(sb), vol)
    a.step(2)
    b.step(2)
    pa, va = a.state_f64(sa)
    pb, vb = b.state_f64(sb)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert np.array_equal(a.sediment_fx(sa), b.sediment_fx(sb))
    assert np.array_equal(a.terrain_volume_fx(sa), b.terrain_volume_fx(sb))
    assert np.array_equal(a.terrain_heights(sa), b.terrain_heights(sb))
    with pytest.raises(Exception):
        b.set_terrain_volume_fx(sb, vol.ravel()[:-1])
    with pytest.raises(Exception):
        b.set_terrain_volume_fx(sb, np.full(vol.size, -(2 ** 62) - 1, np.int64))  # outside |v| <= 2^62


def test_full_size_column_properties():
    """The 262,144-particle column (the bench workload) after 40 steps:
    exact mass units; a replay from a saved state ends bit for bit where the
    first run did; density within 1e-5 of the FP64 path on the same state;
    symmetric neighbour sets (fl(a - b) = -fl(b - a), so j is i's neighbour
    exactly when i is j's)."""
    sc = scenes.cfg2()
    ctx, sid = gpu_context(sc)
    _, _, u0 = ctx.mass(sid)
    ctx.step(40)
    ctx.save_state()
    ctx.step(10)
    p1, v1 = ctx.state_f64(sid)
    ctx.restore_state()
    ctx.step(10)
    p2, v2 = ctx.state_f64(sid)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
    _, _, u1 = ctx.mass(sid)
    assert u1 == u0
    c64, s64 = gpu_context(sc, sphe.FP64)
    c64.set_particles(s64, p2, vel=v2, ids=ctx.ids(sid), sed=ctx.sediment_fx(sid))
    c64.set_terrain_volume_fx(s64, ctx.terrain_volume_fx(sid))
    rho = ctx.debug_density(sid)
    ref = c64.debug_density(s64)
    assert (np.abs(rho - ref) / ref).max() <= DENSITY_TOL
    counts, off, idx = ctx.debug_neighbours(sid)
    n = len(counts)
    assert counts.min() >= 1
    rows = np.repeat(np.arange(n, dtype=np.int64), counts)
    fwd = rows * n + idx.astype(np.int64)
    rev = idx.astype(np.int64) * n + rows
    assert np.array_equal(np.sort(fwd), np.sort(rev))


def test_empty_scene_and_lone_particle():
    """A scene without particles beside a populated one changes nothing; a
    lone particle (itself its only neighbour) steps bit for bit like the
    twin and accelerates at g to within one ulp."""
    sc = small_dam()
    a, sa = gpu_context(sc)
    b = sphe.Context()
    se = b.add_scene(sphe.default_params(dt=1e-3))
    sb = b.add_synthetic(sc)
    a.step(3)
    b.step(3)
    pa, va = a.state_f64(sa)
    pb, vb = b.state_f64(sb)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert b.num_particles(se) == 0
    prm = sphe.default_params(dt=1e-3)
    c = sphe.Context()
    s1 = c.add_scene(prm)
    pos = np.array([[0.0, 0.05, 0.0]])
    c.set_particles(s1, pos, vel=np.zeros((1, 3)), ids=np.array([7]))
    acc = c.debug_accel(s1)[0]
    g = np.float32(prm.gravity[1])
    assert acc[0] == 0.0 and acc[2] == 0.0
    assert abs(np.float32(acc[1]) - g) <= np.spacing(abs(g))
    c.step(1)
    p, v = c.state_f64(s1)
    twin = O.orc_step(orc_scene_from(prm), pos, np.zeros((1, 3)), np.zeros(1, np.int64),
                      np.array([7], np.int32), None, None, real=4, comp=1)
    assert np.array_equal(p, twin["pos"]) and np.array_equal(v, twin["vel"])


def test_start_table_when_scan_tiles_empty_and_refill():
    """The scan leaves the cell starts of a tile without particles unwritten
    while they still hold the tile's prefix.  A blob and a block at opposite
    ends of the box along z (the scan tiles between them hold no particles):
    after steps, after the blob jumps next to the block (the empty tiles'
    prefix changes), and after it jumps back, the grid's neighbour sets equal
    brute force on the current positions."""
    s = 0.02
    h = 2.15 * s
    prm = scenes.sph_params(h, s, 5e-4, sphe.EROSION_OFF)
    rng = np.random.default_rng(7)
    blob, _ = scenes.lattice_block(512, s, 0.40, 0.56, 0.05, 0.21, 0.30, rng)
    block, _ = scenes.lattice_block(1024, s, 0.30, 0.62, 0.75, 0.91, 0.10, rng)
    pos = np.concatenate([blob, block])
    n = len(pos)
    ctx = sphe.Context()
    sid = ctx.add_scene(prm)
    ctx.set_particles(sid, pos, ids=np.arange(n, dtype=np.int32))
    thr = O.derive(orc_scene_from(prm)).thr

    def state_by_id():
        p, v = ctx.state_f64(sid)
        order = np.argsort(ctx.ids(sid))
        return p[order], v[order]

    def check():
        p, _ = state_by_id()
        counts, off, idx = ctx.debug_neighbours(sid)
        bc, boff, bidx = O.neighbours_brute(p.astype(np.float32), thr)
        assert np.array_equal(counts, bc)
        assert np.array_equal(off, boff)
        assert np.array_equal(idx, bidx)

    ctx.step(10)
    check()
    for dz in (0.45, -0.45):
        p, v = state_by_id()
        p[:512, 2] += dz
        ctx.set_particles(sid, p, v, ids=np.arange(n, dtype=np.int32))
        ctx.step(5)
        check()
