"""The oracle's grid step (the library's algorithm) against brute force, the
reference restatement and closed forms of the library-defined terrain and
erosion model."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PI_F = float(np.float32(3.141592))


def jitter_state():
    g = np.load(os.path.join(GOLD, "f3_jitter.npz"))
    return g


def f3_scene():
    return O.make_scene(dt=1e-3, rho0=1000.0, g=(0.0, -9.8, 0.0), box_e=0.5)


@pytest.mark.parametrize("yz", [1, 2, 3])
def test_grid_neighbours_equal_brute_force(yz):
    rng = np.random.default_rng(11)
    pos = rng.uniform(-0.19, 0.19, size=(3000, 3)).astype(np.float32)
    pos[:300] = pos[300:600]  # coincident pairs
    sc = O.make_scene(dt=0.0, yz=yz)
    n = len(pos)
    out = O.orc_step(sc, pos, np.zeros((n, 3)), np.zeros(n, np.int64), np.arange(n, dtype=np.int32),
                     real=4, comp=1)
    cnt, off, idx = O.neighbours_brute(pos, O.pred_threshold(np.float32(0.0457)))
    assert np.array_equal(out["nbcount"], cnt)
    nbmax = np.array([max([j for j in idx[off[i]:off[i + 1]] if j != i], default=-1) for i in range(n)])
    assert np.array_equal(out["nbmax"], nbmax)


def _clip_cases():
    """Inputs that stress the clipped stencil rows: a dense blob of more than
    1023 particles in one cell (no slab offsets), lattices whose points sit
    on cell and slab boundaries and at exactly h apart, and particles outside
    the grid (clamped into the edge cells)."""
    rng = np.random.default_rng(21)
    h = np.float32(0.0457)
    cell = float(h) * (1 + 2.0 ** -8)
    org = -0.2
    # cell (2, 2, 2) holds 2300 particles in its first and last x slab, all in
    # one y/z cell also when cells are split along y and z, and a
    # line of probes along x through it ends its rows at every slab
    fr = np.concatenate([rng.uniform(0.02, 0.22, 1300), rng.uniform(0.78, 0.98, 1000)])
    blob = org + cell * np.stack([2 + fr, 2 + rng.uniform(0.3, 0.45, 2300), 2 + rng.uniform(0.3, 0.45, 2300)], 1)
    probes = org + cell * np.stack([rng.uniform(0.3, 4.7, 400), np.full(400, 2.375), np.full(400, 2.375)], 1)
    blob = np.concatenate([blob, probes]).astype(np.float32)
    bg = rng.uniform(-0.19, 0.19, size=(1500, 3)).astype(np.float32)
    g = np.arange(-3, 4)
    lat_h = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * h  # spacing h
    lat_c = (org + np.stack(np.meshgrid(np.arange(12), np.arange(6), np.arange(6), indexing="ij"), -1)
             .reshape(-1, 3) * (cell / 4.0)).astype(np.float32)  # on slab and cell edges
    out = rng.uniform(-0.3, 0.3, size=(300, 3)).astype(np.float32)  # partly outside the box
    return {"blob": np.concatenate([blob, bg]), "lattice_h": lat_h.astype(np.float32),
            "edges": lat_c, "outside": np.concatenate([out, bg[:500]])}


@pytest.mark.parametrize("yz", [1, 2, 3])
@pytest.mark.parametrize("slabs", [1, 2, 4])
@pytest.mark.parametrize("case", ["blob", "lattice_h", "edges", "outside"])
@pytest.mark.parametrize("real", [4, 8])
def test_clipped_rows_keep_every_neighbour(yz, slabs, case, real):
    """The clipped stencil rows (x slabs, y/z cells of 1, 1/2 or 1/3 the x edge,
    skipped rows, cells taken whole above 1023 particles) still reach every
    neighbour the all-pairs FP32 predicate finds: per-particle counts and
    largest neighbour ids equal brute force."""
    pos = _clip_cases()[case]
    n = len(pos)
    sc = O.make_scene(dt=0.0, slabs=slabs, yz=yz)
    out = O.orc_step(sc, pos, np.zeros((n, 3)), np.zeros(n, np.int64), np.arange(n, dtype=np.int32),
                     real=real, comp=1)
    thr = O.pred_threshold(np.float32(0.0457)) if real == 4 else float(np.float32(0.0457)) ** 2
    if real == 4:
        cnt, off, idx = O.neighbours_brute(pos, thr)
    else:
        d = pos.astype(np.float64)
        dx, dy, dz = (d[:, None, a] - d[None, :, a] for a in range(3))
        m = ((dx * dx + dy * dy) + dz * dz) <= thr
        cnt = m.sum(1)
        idx = None
    assert np.array_equal(out["nbcount"], cnt)
    if idx is not None:
        nbmax = np.array([max([j for j in idx[off[i]:off[i + 1]] if j != i], default=-1) for i in range(n)])
        assert np.array_equal(out["nbmax"], nbmax)


@pytest.mark.parametrize("yz", [1, 2, 3])
def test_cell_keys_cover_neighbours(yz):
    """Every neighbour pair lies in adjacent cells along x and within yz
    cells along y and z (the 3 x (2 yz + 1)^2 stencil), and the split cells
    are exactly 1/yz of the x edge."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 1.0, size=(4000, 3)).astype(np.float32)
    sc = O.make_scene(h=0.05, box_c=(0.5, 0.5, 0.5), box_e=0.5, yz=yz)
    c = O.derive(sc)
    keys = O.cell_keys(c, pos)
    nx, ny = c.dims[0], c.dims[1]
    k = np.stack([keys % nx, (keys // nx) % ny, keys // (nx * ny)], 1)
    cnt, off, idx = O.neighbours_brute(pos, c.thr)
    reach = np.array([1, yz, yz])
    for i in range(0, 4000, 7):
        nb = idx[off[i]:off[i + 1]]
        assert (np.abs(k[nb] - k[i]) <= reach).all()
    assert np.float32(c.inv_yz) == np.float32(c.inv_cell) * np.float32(yz)
    c1 = O.derive(O.make_scene(h=0.05, box_c=(0.5, 0.5, 0.5), box_e=0.5, yz=1))
    assert c.dims[0] == c1.dims[0] and yz * c1.dims[1] - (yz - 1) <= c.dims[1] <= yz * c1.dims[1]


def test_fp64_grid_step_vs_reference_f3():
    g = jitter_state()
    n = len(g["in_pos"])
    out = O.orc_step(f3_scene(), g["in_pos"], g["in_vel"], np.zeros(n, np.int64), np.arange(n, dtype=np.int32),
                     real=8)
    rel = np.abs(out["rho"] - g["Density"]) / g["Density"]
    assert rel.max() <= 2e-6
    scale = np.abs(g["Acceleration"]).max()
    assert np.abs(out["acc"] - g["Acceleration"]).max() <= 2e-5 * scale
    assert np.array_equal(out["nbmax"], g["NeighbId"])
    assert np.abs(out["pos"] - g["Position"]).max() <= 1e-7
    assert np.abs(out["pforce"] - g["PressureForce"]).max() <= 2e-5 * np.abs(g["PressureForce"]).max()
    assert np.abs(out["normal"] - g["SurfaceNormal"]).max() <= 2e-5 * np.abs(g["SurfaceNormal"]).max()


@pytest.mark.parametrize("comp,tol", [(1, 5e-7), (0, 3e-6)])
def test_fp32_twin_vs_fp64(comp, tol):
    g = jitter_state()
    n = len(g["in_pos"])
    args = (g["in_pos"], g["in_vel"], np.zeros(n, np.int64), np.arange(n, dtype=np.int32))
    r64 = O.orc_step(f3_scene(), *args, real=8)
    r32 = O.orc_step(f3_scene(), *args, real=4, comp=comp)
    assert (np.abs(r32["rho"] - r64["rho"]) / r64["rho"]).max() <= tol


def test_orc_isolated_and_lattice_closed_forms():
    L = O.lattice(1000)
    n = len(L)
    out = O.orc_step(O.make_scene(dt=0.01), L["Position"], L["Velocity"], np.zeros(n, np.int64), L["Id"], real=8)
    H, M = float(np.float32(0.0457)), float(np.float32(0.02))
    s = 0.025
    c = 315.0 / (64.0 * PI_F * H ** 9)
    offs = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]
    sx = [float(np.float32(-0.2 + i * 0.025)) for i in range(10)]  # lattice x as float
    assert abs(sx[1] - sx[0] - s) < 1e-7
    rho = sum(M * c * (H * H - (i * i + j * j + k * k) * s * s) ** 3 for i, j, k in offs)
    assert abs(out["rho"][555] - rho) <= 1e-5 * rho
    P = np.zeros((1, 3))
    one = O.orc_step(O.make_scene(dt=0.01), P, P, np.zeros(1, np.int64), np.zeros(1, np.int32), real=8)
    assert abs(one["rho"][0] - M * 315.0 / (64.0 * PI_F * H ** 3)) <= 1e-12 * 328.3
    assert one["acc"][0][1] == float(np.float32(-9.82))


def flat_terrain(T=8, height=0.3, fx=60):
    hts = np.full((T, T), height, dtype=np.float32)
    vol = np.round((hts.astype(np.float64)) * (1.0 / T) ** 2 * 2.0 ** fx).astype(np.int64)
    return hts, vol


def terrain_scene(T=8, fx=60, dt=0.01, erosion=1, **kw):
    base = dict(h=0.0457, dt=dt, box_c=(0.5, 0.5, 0.5), box_e=0.5, g=(0.0, 0.0, 0.0), tnx=T, tnz=T,
                t_org=(0.0, 0.0, 0.0), t_cell=float(np.float32(1.0 / T)), fx_exp=fx, erosion=erosion)
    base.update(kw)
    return O.make_scene(**base)


def test_terrain_collision_closed_form_flat():
    """A particle falling through a flat surface at H: projected onto it, and
    v_y' = -(0.5 d / (dt |v|)) v_y = 0.5 d / dt for vertical motion."""
    T, H0 = 8, np.float32(0.3)
    sc = terrain_scene(T, erosion=0)
    c = O.derive(sc)
    hts = O.heights(c, flat_terrain(T, H0)[1])
    pos = np.array([[0.5, 0.301, 0.5]])
    vel = np.array([[0.0, -1.0, 0.0]])
    out = O.orc_step(sc, pos, vel, np.zeros(1, np.int64), np.zeros(1, np.int32), flat_terrain(T, H0)[1], hts,
                     real=8)
    hs = float(hts[0, 0])
    dt = float(np.float32(0.01))  # parameters are FP32 values, as in the reference
    d = hs - (0.301 - dt)
    assert out["pos"][0][1] == hs
    assert abs(out["vel"][0][1] - 0.5 * d / dt) <= 1e-12
    assert out["vel"][0][0] == 0.0 and out["vel"][0][2] == 0.0


def test_terrain_normal_of_a_plane():
    """Cell-centred samples of the plane y = 0.2 + 0.3 x: the bilinear patch
    is the plane, the contact normal is (-0.3, 1, 0)/sqrt(1.09)."""
    T = 16
    xc = (np.arange(T) + 0.5) / T
    hts = np.tile((0.2 + 0.3 * xc).astype(np.float32), (T, 1))
    sc = terrain_scene(T, erosion=0)
    pos = np.array([[0.41, 0.2 + 0.3 * 0.41 + 0.001, 0.5]])
    vel = np.array([[0.0, -1.0, 0.0]])
    vol = np.zeros((T, T), np.int64)
    out = O.orc_step(sc, pos, vel, np.zeros(1, np.int64), np.zeros(1, np.int32), vol, hts, real=8)
    x = out["pos"][0]
    assert abs(x[1] - (0.2 + 0.3 * x[0])) < 1e-6  # on the plane
    v = out["vel"][0]
    n = np.array([-0.3, 1.0, 0.0]) / np.sqrt(1.09)
    vin = np.array([0.0, -1.0, 0.0])
    d = (0.2 + 0.3 * 0.41) - (pos[0][1] - 0.01)
    expect = vin - n * (1 + 0.5 * d / (0.01 * 1.0)) * np.dot(vin, n)
    assert np.abs(v - expect).max() < 1e-5


def test_erosion_single_particle_closed_form():
    """In the band, at rest sediment 0: the particle takes ke kc |v| m / rho0
    from the cell under it (one fixed-point rounding)."""
    T, fx = 8, 60
    hts, vol = flat_terrain(T, 0.3, fx)
    sc = terrain_scene(T, fx)
    c = O.derive(sc)
    hts = O.heights(c, vol)
    pos = np.array([[0.3, 0.32, 0.7]])  # 0.02 above the surface, band h
    vel = np.array([[0.6, 0.0, 0.8]])   # |v| = 1, horizontal: no collision
    out = O.orc_step(sc, pos, vel, np.zeros(1, np.int64), np.zeros(1, np.int32), vol, hts, real=8)
    m, rho0, kc, ke = (float(np.float32(v)) for v in (0.02, 998.29, 0.05, 0.3))
    expect = ke * kc * 1.0 * m / rho0
    got = out["sed"][0] * 2.0 ** -fx
    assert abs(got - expect) <= 1e-9 * expect
    i, k = int(0.3 * T), int(0.7 * T)
    dv = vol - out["vol"]
    assert dv[k, i] == out["sed"][0]
    assert dv.sum() == out["sed"][0] and np.count_nonzero(dv) == 1


def test_deposition_and_band():
    T, fx = 8, 60
    _, vol = flat_terrain(T, 0.3, fx)
    sc = terrain_scene(T, fx)
    hts = O.heights(O.derive(sc), vol)
    s0 = np.int64(1 << 40)
    pos = np.array([[0.3, 0.31, 0.7], [0.6, 0.3 + 0.2, 0.6]])  # second far above the band
    vel = np.zeros((2, 3))
    out = O.orc_step(sc, pos, vel, np.array([s0, s0]), np.array([0, 1], np.int32), vol, hts, real=8)
    kd = float(np.float32(0.3))
    assert abs(out["sed"][0] - s0 * (1 - kd)) <= 1
    assert out["sed"][1] == s0
    assert (vol - out["vol"]).sum() == out["sed"][0] - s0


@pytest.mark.parametrize("real", [4, 8])
def test_erosion_below_the_terrain_base(real):
    """Erosion is not capped at a cell's volume: a fast particle over a flat
    terrain at its base (volume 0) takes volume from the cell, which goes
    negative (its height below the base) while water and sediment units stay
    exactly conserved."""
    T, fx = 8, 60
    vol = np.zeros((T, T), np.int64)
    sc = terrain_scene(T, fx)
    c = O.derive(sc)
    hts = O.heights(c, vol)
    pos = np.array([[0.3, 0.02, 0.7]])  # inside the band above height 0
    vel = np.array([[0.6, 0.0, 0.8]])
    st = dict(pos=pos, vel=vel, sed=np.zeros(1, np.int64), vol=vol, height=hts)
    for _ in range(2):
        st = O.orc_step(sc, st["pos"], st["vel"], st["sed"], np.zeros(1, np.int32), st["vol"], st["height"],
                        real=real, comp=1)
        assert int(st["vol"].sum()) + int(st["sed"].sum()) == 0
    k, i = int(0.7 * T), int(0.3 * T)
    assert st["vol"][k, i] < 0 and st["height"].reshape(T, T)[k, i] < 0.0
    assert st["sed"][0] == -st["vol"][k, i]


@pytest.mark.parametrize("real", [4, 8])
def test_erosion_exchange_saturates(real):
    """One exchange is bounded by 2^52 fixed-point units: with the unit 2^-100
    the capacity of a particle at 1 m/s (ke kc m / rho0, about 3e-7) is far
    above 2^-48, and the particle takes exactly 2^52 units (an exchange above
    the bound is clamped to it, not dropped)."""
    T, fx = 8, 100
    vol = np.zeros((T, T), np.int64)
    sc = terrain_scene(T, fx)
    hts = O.heights(O.derive(sc), vol)
    pos = np.array([[0.3, 0.02, 0.7]])
    vel = np.array([[0.6, 0.0, 0.8]])
    out = O.orc_step(sc, pos, vel, np.zeros(1, np.int64), np.zeros(1, np.int32), vol, hts, real=real, comp=1)
    assert out["sed"][0] == 2 ** 52
    assert int(out["vol"].sum()) == -(2 ** 52)


@pytest.mark.parametrize("real", [4, 8])
def test_multi_step_mass_exact(real):
    T, fx = 16, 62
    rng = np.random.default_rng(2)
    xc = (np.arange(T) + 0.5) / T
    hts = (0.1 + 0.05 * np.sin(6 * xc)[None, :] * np.cos(5 * xc)[:, None]).astype(np.float32)
    vol = np.round(hts.astype(np.float64) * (1.0 / T) ** 2 * 2.0 ** fx).astype(np.int64)
    sc = terrain_scene(T, fx, dt=1e-3, g=(0.0, -9.8, 0.0), h=0.05, mass=0.012, rho0=1000.0)
    hts = O.heights(O.derive(sc), vol)
    n = 600
    pos = np.stack([rng.uniform(0.05, 0.4, n), rng.uniform(0.16, 0.22, n), rng.uniform(0.05, 0.95, n)], 1)
    pos = pos.astype(np.float32).astype(np.float64)
    st = dict(pos=pos, vel=np.zeros_like(pos), sed=np.zeros(n, np.int64), vol=vol, height=hts)
    total0 = int(vol.sum())
    exchanged = 0
    for _ in range(40):
        st = O.orc_step(sc, st["pos"], st["vel"], st["sed"], np.arange(n, dtype=np.int32), st["vol"],
                        st["height"], real=real, comp=1)
        exchanged += st["active"]
        assert int(st["vol"].sum()) + int(st["sed"].sum()) == total0
    assert exchanged > 0 and st["sed"].sum() > 0


def test_box_center_zero_is_reference_collision():
    """The generalised cube with centre 0 and half extent len reproduces the
    reference's collisionS on the F4 cases bit for bit (isolated particles:
    the only force is gravity, so both steps are the same operations)."""
    g = np.load(os.path.join(GOLD, "f4_box.npz"))
    n = len(g["in_pos"])
    sc = O.make_scene(dt=0.01)
    out = O.orc_step(sc, g["in_pos"], g["in_vel"], np.zeros(n, np.int64), np.arange(n, dtype=np.int32), real=4,
                     comp=1)
    assert out["pos"].astype(np.float32).tobytes() == g["Position"].tobytes()
    v32 = out["vel"].astype(np.float32)
    assert np.array_equal(np.isnan(v32), np.isnan(g["Velocity"]))
    ok = ~np.isnan(g["Velocity"])
    assert np.array_equal(v32[ok], g["Velocity"][ok])
