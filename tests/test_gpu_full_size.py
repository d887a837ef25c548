"""The batched and large configurations at their full sizes on the GPU.

  256 dam-break scenes (4,641,519 particles, terrains 64..256 cells square);
  the 4,194,304-particle layer over a ridged 2048^2 terrain;
  1,024 dam-break scenes plus an 8,388,608-particle layer on a 4096^2
  terrain (13,954,298 particles).

Each is stepped into terrain contact (the batch and the layer through their
whole 300- and 1,000-step runs), then checked through properties that do not
depend on the size:
  * fixed-point water/sediment units conserved exactly per scene, with
    erosion active;
  * a replay from a saved state ends bit for bit where the first run did;
  * density within 1e-5 relative of the FP64 path on the same state;
  * neighbour sets symmetric (fl(a - b) = -fl(b - a)), and equal to brute
    force on windows of the large scenes;
  * a sample of scenes bit for bit equal to the same scenes stepped alone.
"""
import numpy as np
import pytest

import oracle as O
import sphe
from sphe import scenes

pytestmark = pytest.mark.gpu

DENSITY_TOL = 1e-5


def snapshot(ctx, sid):
    p, v = ctx.state_f64(sid)
    nx, _ = ctx.terrain_dims(sid)
    return dict(pos=p, vel=v, ids=np.asarray(ctx.ids(sid), dtype=np.int32), sed=ctx.sediment_fx(sid),
                vol=ctx.terrain_volume_fx(sid) if nx else None)


def same_state(a, b):
    return (np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
            and np.array_equal(a["sed"], b["sed"]) and np.array_equal(a["ids"], b["ids"])
            and (a["vol"] is None or np.array_equal(a["vol"], b["vol"])))


class Run:
    """A workload stepped `steps` times in one context, with the initial
    fixed-point units and the state of some sample scenes at that point."""

    def __init__(self, scs, steps, sample):
        self.scs = scs
        self.steps = steps
        self.ctx = sphe.Context(sphe.FP32)
        self.sids = [self.ctx.add_synthetic(s) for s in scs]
        self.units0 = [self.ctx.mass(s)[2] for s in self.sids]
        self.ctx.step(steps)
        self.ctx.sync()
        self.sample = sample
        self.at_steps = {i: snapshot(self.ctx, self.sids[i]) for i in sample}


def mass_units_exact(run):
    units = [run.ctx.mass(s)[2] for s in run.sids]
    assert units == run.units0
    sed = sum(run.ctx.mass(s)[1] for s in run.sids)
    assert sed > 0  # erosion did happen
    assert run.ctx.status() & 3 == 0  # bit 4 (a stencil walk while lists grew) is allowed


def replay_bit_identical(run, k, scenes_to_check):
    run.ctx.save_state()
    run.ctx.step(k)
    first = [snapshot(run.ctx, run.sids[i]) for i in scenes_to_check]
    run.ctx.restore_state()
    run.ctx.step(k)
    second = [snapshot(run.ctx, run.sids[i]) for i in scenes_to_check]
    assert all(same_state(a, b) for a, b in zip(first, second))


def density_vs_fp64(run, scenes_to_check):
    c64 = sphe.Context(sphe.FP64)
    for i in scenes_to_check:
        st = snapshot(run.ctx, run.sids[i])
        sc = run.scs[i]
        s64 = c64.add_scene(sc.params, sc.terrain)
        c64.set_particles(s64, st["pos"], vel=st["vel"], ids=st["ids"], sed=st["sed"])
        if st["vol"] is not None:
            c64.set_terrain_volume_fx(s64, st["vol"])
    worst = 0.0
    for k, i in enumerate(scenes_to_check):
        rho = run.ctx.debug_density(run.sids[i])
        ref = c64.debug_density(k)
        worst = max(worst, float((np.abs(rho - ref) / ref).max()))
    c64.close()
    assert worst <= DENSITY_TOL, worst


def neighbours_symmetric(ctx, sid, rows=None):
    """j in N(i) exactly when i in N(j), over all rows or a sample of rows."""
    counts, off, idx = ctx.debug_neighbours(sid)
    ids = np.asarray(ctx.ids(sid))
    assert np.array_equal(ids, np.arange(len(ids)))  # neighbour ids index the rows
    assert counts.min() >= 1
    n = len(counts)
    if rows is None:
        r = np.repeat(np.arange(n, dtype=np.int64), counts)
        fwd = r * n + idx.astype(np.int64)
        rev = idx.astype(np.int64) * n + r
        assert np.array_equal(np.sort(fwd), np.sort(rev))
        return counts, off, idx
    for i in rows:
        for j in idx[off[i]:off[i + 1]]:
            row = idx[off[j]:off[j + 1]]
            k = np.searchsorted(row, i)
            assert k < len(row) and row[k] == i, (i, j)
    return counts, off, idx


def brute_force_windows(pos, counts, off, idx, h, rng, windows, half):
    """`windows` cube windows around particles drawn by rng, each of half
    width half, 3/4 or 1/2 of it (the first that holds 500 .. 20,000
    particles: basins are much denser than the surface); in each, the grid's
    neighbour sets of the particles at least h from its faces equal brute
    force among the window's particles."""
    done = 0
    for _ in range(60):
        centre = pos[rng.integers(0, len(pos))]
        for hw in (half, 0.75 * half, 0.5 * half):
            sel = np.nonzero(np.all(np.abs(pos - centre) <= hw, axis=1))[0]
            if len(sel) <= 20000:
                break
        if 500 <= len(sel) <= 20000:
            window_brute_force(pos, counts, off, idx, h, centre, hw, sel)
            done += 1
            if done == windows:
                return
    raise AssertionError(f"only {done} of {windows} windows of 500 .. 20,000 particles")


def window_brute_force(pos, counts, off, idx, h, centre, half, sel):
    """The particles sel inside a cube window; for those at least h from its
    faces the grid's neighbour sets equal brute force among sel."""
    inner = np.all(np.abs(pos[sel] - centre) <= half - 1.01 * h, axis=1)
    assert inner.sum() >= 50
    bc, boff, bidx = O.neighbours_brute(pos[sel].astype(np.float32), O.pred_threshold(np.float32(h)))
    for t in np.nonzero(inner)[0]:
        i = sel[t]
        want = sel[bidx[boff[t]:boff[t + 1]]]
        assert np.array_equal(idx[off[i]:off[i + 1]], np.sort(want)), i


def alone_equals_batch(run, i):
    sc = run.scs[i]
    c = sphe.Context(sphe.FP32)
    s = c.add_synthetic(sc)
    c.step(run.steps)
    assert same_state(snapshot(c, s), run.at_steps[i])
    c.close()


# ------------------------------------------------------------------ cfg3

@pytest.fixture(scope="module")
def cfg3_run():
    return Run(scenes.cfg3(), 300, sample=[0, 77, 190, 255])  # the whole 300-step run


@pytest.mark.timeout(300)
def test_cfg3_mass_units_exact(cfg3_run):
    assert sum(s.n for s in cfg3_run.scs) == 4641519
    mass_units_exact(cfg3_run)


@pytest.mark.timeout(300)
def test_cfg3_scenes_alone_equal_the_batch(cfg3_run):
    for i in cfg3_run.sample:
        alone_equals_batch(cfg3_run, i)


@pytest.mark.timeout(300)
def test_cfg3_density_vs_fp64(cfg3_run):
    density_vs_fp64(cfg3_run, list(range(0, 256, 5)))


@pytest.mark.timeout(300)
def test_cfg3_neighbours_symmetric(cfg3_run):
    for i in (3, 100, 201):
        neighbours_symmetric(cfg3_run.ctx, cfg3_run.sids[i])


@pytest.mark.timeout(300)
def test_cfg3_replay_bit_identical(cfg3_run):
    replay_bit_identical(cfg3_run, 20, list(range(0, 256, 9)))


@pytest.mark.timeout(300)
def test_cfg3_step_sequences_equal_direct_launches():
    """The whole 300-step run as one captured 300-step sequence, as nine
    32-step sequences plus single-step graphs, and launched kernel by kernel
    ends in the same bits for all 256 scenes."""
    scs = scenes.cfg3()
    ctxs = []
    for mode in ("seq300", "seq32", "direct"):
        c = sphe.Context(sphe.FP32)
        if mode == "seq300":
            c.set_sequence_steps(300)
            assert c.get_sequence_steps() == 300
        if mode == "direct":
            c.use_graph(False)
        sids = [c.add_synthetic(s) for s in scs]
        c.step(300)
        assert c.status() & 3 == 0
        ctxs.append((c, sids))
    (a, sa), others = ctxs[0], ctxs[1:]
    for b, sb in others:
        for x, y in zip(sa, sb):
            assert same_state(snapshot(a, x), snapshot(b, y))
    for c, _ in ctxs:
        c.close()


# ------------------------------------------------------------------ cfg4

@pytest.fixture(scope="module")
def cfg4_run():
    return Run([scenes.cfg4()], 1000, sample=[])  # the whole 1,000-step run


@pytest.mark.timeout(300)
def test_cfg4_mass_units_exact(cfg4_run):
    assert cfg4_run.scs[0].n == 4194304
    mass_units_exact(cfg4_run)


@pytest.mark.timeout(300)
def test_cfg4_density_vs_fp64(cfg4_run):
    density_vs_fp64(cfg4_run, [0])


@pytest.mark.timeout(300)
def test_cfg4_neighbours_symmetric_and_brute_force_windows(cfg4_run):
    ctx, sid = cfg4_run.ctx, cfg4_run.sids[0]
    rng = np.random.default_rng(3)
    n = cfg4_run.scs[0].n
    counts, off, idx = neighbours_symmetric(ctx, sid, rows=rng.integers(0, n, 300))
    pos, _ = ctx.state_f64(sid)
    h = float(cfg4_run.scs[0].params.h)
    brute_force_windows(pos, counts, off, idx, h, rng, 3, 4 * h)


@pytest.mark.timeout(300)
def test_cfg4_replay_bit_identical(cfg4_run):
    replay_bit_identical(cfg4_run, 10, [0])


# ------------------------------------------------------------------ cfg5

@pytest.fixture(scope="module")
def cfg5_run():
    return Run(scenes.cfg5(), 120, sample=[0, 511, 1023, 1024])


@pytest.mark.timeout(300)
def test_cfg5_mass_units_exact(cfg5_run):
    assert len(cfg5_run.scs) == 1025 and sum(s.n for s in cfg5_run.scs) == 13954298
    mass_units_exact(cfg5_run)


@pytest.mark.timeout(300)
def test_cfg5_scenes_alone_equal_the_batch(cfg5_run):
    for i in cfg5_run.sample:  # three small scenes and the 8.4M-particle layer
        alone_equals_batch(cfg5_run, i)


@pytest.mark.timeout(300)
def test_cfg5_density_vs_fp64(cfg5_run):
    density_vs_fp64(cfg5_run, [0, 300, 700, 1024])


@pytest.mark.timeout(300)
def test_cfg5_neighbours(cfg5_run):
    ctx = cfg5_run.ctx
    for i in (2, 600):
        neighbours_symmetric(ctx, cfg5_run.sids[i])
    sid = cfg5_run.sids[1024]
    rng = np.random.default_rng(5)
    n = cfg5_run.scs[1024].n
    counts, off, idx = neighbours_symmetric(ctx, sid, rows=rng.integers(0, n, 200))
    pos, _ = ctx.state_f64(sid)
    h = float(cfg5_run.scs[1024].params.h)
    brute_force_windows(pos, counts, off, idx, h, rng, 1, 5 * h)


@pytest.mark.timeout(300)
def test_cfg5_replay_bit_identical(cfg5_run):
    replay_bit_identical(cfg5_run, 5, [0, 400, 1024])
