"""The oracle's ref32 restatement against the reference step.

Pins: the golden fixtures written by the reference's own FluidSystemSPH
(tests/golden/make_golden.py), a live comparison with oracle/_ref when it was
built here, and closed forms that a dropped term, a wrong sign or a wrong
constant would break.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PI_F = float(np.float32(3.141592))  # Erosion/sphere.h:8
H = float(np.float32(0.0457))
M = float(np.float32(0.02))


def load(name):
    return np.load(os.path.join(GOLD, name))


def recs(buf):
    return np.frombuffer(buf.tobytes(), dtype=O.PARTICLE_DTYPE)


def assert_records_equal(a, b, fields=O.PARTICLE_DTYPE.names):
    for f in fields:
        assert a[f].tobytes() == b[f].tobytes(), f


def test_lattice_fixtures():
    g = load("lattice.npz")
    for name, (n, add) in {"l1000": (1000, 0), "l100": (100, 0), "l125_125": (125, 125)}.items():
        L = O.lattice(n)
        if add:
            L = np.concatenate([L, O.lattice(add, first_id=len(L))])
        for f in ("Id", "Position", "Velocity", "Acceleration"):
            assert L[f].tobytes() == g[f"{name}_{f}"].tobytes(), (name, f)
    # Initialize(100): i < cbrt(100) = 4.64 gives 5 per axis
    assert len(O.lattice(100)) == 125


@pytest.mark.parametrize("steps,key", [(1, "step1"), (10, "step10")])
def test_f1_lattice_bit_exact(steps, key):
    ref = recs(load("f1_lattice.npz")[key])
    out = O.ref32_step(O.lattice(1000), O.default_params(dt=0.01), steps)
    assert_records_equal(out, ref)


@pytest.mark.parametrize("steps,key", [(1, "step1"), (3, "step3")])
def test_f2_coincident_bit_exact(steps, key):
    ref = recs(load("f2_coincident.npz")[key])
    L = np.concatenate([O.lattice(125), O.lattice(125, first_id=125)])
    out = O.ref32_step(L, O.default_params(dt=0.01), steps)
    assert_records_equal(out, ref)


def test_f3_jitter_bit_exact():
    g = load("f3_jitter.npz")
    P = np.zeros(len(g["in_pos"]), dtype=O.PARTICLE_DTYPE)
    P["Id"] = np.arange(len(P))
    P["NeighbId"] = -1
    P["Position"] = g["in_pos"]
    P["Velocity"] = g["in_vel"]
    p = O.default_params(dt=1e-3)
    p.len, p.p0 = 0.5, 1000.0
    p.g[1] = -9.8
    out = O.ref32_step(P, p, 1)
    for f in ("Position", "Velocity", "Acceleration", "Density", "Pressure", "NeighbId", "PressureForce",
              "SurfaceNormal"):
        assert out[f].tobytes() == g[f].tobytes(), f


def test_f4_box_bit_exact():
    g = load("f4_box.npz")
    P = np.zeros(len(g["in_pos"]), dtype=O.PARTICLE_DTYPE)
    P["Id"] = np.arange(len(P))
    P["NeighbId"] = -1
    P["Position"] = g["in_pos"]
    P["Velocity"] = g["in_vel"]
    out = O.ref32_step(P, O.default_params(dt=0.01), 1)
    for f in ("Position", "Velocity", "Acceleration", "Density", "NeighbId"):
        assert out[f].tobytes() == g[f].tobytes(), f
    # a particle outside the box with |v'| = 0 gets NaN velocity: the
    # response divides by |v'| (fluid_system.h:345)
    assert np.isnan(out["Velocity"][11]).all()
    assert np.isfinite(out["Velocity"][:11]).all()


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("seed", [1, 2])
def test_live_reference_random_states(seed):
    rng = np.random.default_rng(seed)
    n = 700
    P = np.zeros(n, dtype=O.PARTICLE_DTYPE)
    P["Id"] = np.arange(n)
    P["NeighbId"] = -1
    P["Position"] = rng.uniform(-0.12, 0.12, size=(n, 3)).astype(np.float32)
    P["Velocity"] = rng.normal(0, 0.5, size=(n, 3)).astype(np.float32)
    # a few coincident pairs
    P["Position"][1:20:2] = P["Position"][0:19:2]
    p = O.default_params(dt=0.005)
    a = O.ref32_step(P, p, 5)
    b = O.refbin_step(P, p, 5)
    assert_records_equal(a, b)


def test_isolated_particle_closed_form():
    """Density of an isolated particle is its self term m 315/(64 pi h^3);
    it free-falls with a = g exactly."""
    P = np.zeros(1, dtype=O.PARTICLE_DTYPE)
    p = O.default_params(dt=0.01)
    out = O.ref32_step(P, p, 1)
    rho = M * 315.0 / (64.0 * PI_F * H ** 3)
    assert abs(out["Density"][0] - rho) <= 2e-6 * rho
    assert abs(rho - 328.2934) < 1e-3
    np.testing.assert_array_equal(out["Acceleration"][0], np.float32([0.0, -9.82, 0.0]))
    assert out["Velocity"][0][1] == np.float32(np.float32(-9.82) * np.float32(0.01))


def test_lattice_interior_density_closed_form():
    """Interior particle of the reference lattice (spacing 0.025): 26
    neighbours at distances 1, sqrt 2, sqrt 3 spacings plus itself."""
    s = 0.025
    offs = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]
    c = 315.0 / (64.0 * PI_F * H ** 9)
    rho = sum(M * c * (H * H - (i * i + j * j + k * k) * s * s) ** 3 for i, j, k in offs)
    assert abs(rho - 1263.8149) < 1e-3
    out = O.ref32_step(O.lattice(1000), O.default_params(dt=0.01), 1)
    interior = [555, 444, 333]
    for i in interior:
        assert abs(out["Density"][i] - rho) <= 1e-5 * rho
        assert abs(out["Pressure"][i] - 3.0 * (rho - 998.29)) <= 1e-3 * 3.0 * (rho - 998.29)


def test_box_response_closed_form():
    """x' = 0.25 beyond the +x wall at len 0.2 with v = (1,0,0), dt = 0.01:
    d = 0.05, v' = v - n (1 + 0.5 d / (dt |v|)) (v.n) = (-2.5, 0, 0), x' = 0.2."""
    P = np.zeros(1, dtype=O.PARTICLE_DTYPE)
    P["Position"] = [[0.24, 0.0, 0.0]]
    P["Velocity"] = [[1.0, 0.0, 0.0]]
    p = O.default_params(dt=0.01)
    p.g[1] = 0.0
    out = O.ref32_step(P, p, 1)
    assert out["Position"][0][0] == np.float32(0.2)
    assert abs(out["Velocity"][0][0] - (-2.5)) < 1e-5
    assert out["Velocity"][0][1] == 0.0 and out["Velocity"][0][2] == 0.0


def test_two_particle_pressure_antisymmetry():
    """Two particles at distance 0.02 along x: equal and opposite pressure
    forces of magnitude (p_i/rho_i^2 + p_j/rho_j^2) m 45/(pi h^6) (h-r)^2 rho."""
    P = np.zeros(2, dtype=O.PARTICLE_DTYPE)
    P["Id"] = [0, 1]
    P["Position"] = [[0.0, 0.0, 0.0], [0.02, 0.0, 0.0]]
    p = O.default_params(dt=0.0)
    out = O.ref32_step(P, p, 1)
    r = float(np.float32(0.02))
    c6 = 315.0 / (64.0 * PI_F * H ** 9)
    rho = M * c6 * (H ** 6 + (H * H - r * r) ** 3)
    pr = 3.0 * (rho - float(np.float32(998.29)))
    mag = 2.0 * pr / rho ** 2 * M * 45.0 / (PI_F * H ** 6) * (H - r) ** 2 * rho
    pf = out["PressureForce"]
    assert abs(pf[0][0] + pf[1][0]) <= 1e-5 * abs(pf[0][0])
    assert abs(abs(pf[0][0]) - abs(mag)) <= 1e-4 * abs(mag)
    # pressure below rest density pulls the pair together: particle 0 is
    # pulled toward +x
    assert pr < 0 and pf[0][0] > 0


def test_predicate_threshold_exhaustive_walk():
    """T(h) is the largest float with sqrtf(T) <= h: checked by walking the
    floats around it, and the predicate d2 <= T equals sqrtf(d2) <= h."""
    for h in (0.0457, 0.0083984375, 0.00619, 0.1, 1.0, 2.15):
        hf = np.float32(h)
        T = np.float32(O.pred_threshold(hf))
        assert np.sqrt(T) <= hf
        assert np.sqrt(np.nextafter(T, np.float32(np.inf))) > hf
        walk = T.view(np.uint32) + np.arange(-2000, 2000, dtype=np.int64)
        vals = walk.astype(np.uint32).view(np.float32)
        assert np.array_equal(np.sqrt(vals) <= hf, vals <= T)
    # h = 0.0457: T is exactly fl32(h*h) (bits 0x3b08df0c)
    assert np.float32(O.pred_threshold(np.float32(0.0457))).view(np.uint32) == 0x3B08DF0C


def test_h2_never_above_threshold():
    """fl32(h*h) <= T(h) for every h: sqrt of a correctly rounded square rounds
    back to h.  The density kernels rely on it: a candidate outside the
    radius (d2 > T) has h2 - d2 < 0, so its clamped term is 0 without a test."""
    rng = np.random.default_rng(5)
    hs = np.concatenate([np.float32(2.0) ** rng.uniform(-14, 4, 20000),
                         np.float32(2.0) ** np.arange(-14, 5)]).astype(np.float32)
    hs = np.concatenate([hs, np.nextafter(hs, np.float32(0)), np.nextafter(hs, np.float32(np.inf))])
    h2 = hs * hs  # float32 products, rounded to nearest
    T = np.array([O.pred_threshold(h) for h in hs], dtype=np.float32)
    assert np.all(h2 <= T)
    assert np.all(np.sqrt(h2) == hs)


def test_fp32_root_is_within_ulps_of_sqrt():
    """The FP32 kernels' r = |d| and 1/r (three Newton steps for 1/sqrt from
    an integer estimate) against correctly rounded sqrt and 1/sqrt: within
    2.5 and 2 ulp wherever d2 / 2 is a normal float (two Newton steps would
    be ~80 ulp off), r = 0 for d2 = 0, and below the coincident-pair radius
    for smaller d2 (there 1/r is discarded and h - r rounds to h)."""
    rng = np.random.default_rng(11)
    d2 = np.concatenate([
        np.exp2(rng.uniform(-124, 127, 400_000)),
        np.arange(1, 3000, dtype=np.float64) ** 2,
        np.exp2(np.arange(-124, 127, dtype=np.float64)),
        [np.nextafter(np.float32(1), np.float32(0)), np.finfo(np.float32).max],
    ]).astype(np.float32)
    r, inv = O.root_f32(d2)
    ex = np.sqrt(d2.astype(np.float64))

    def ulps(a, e):
        return np.abs(a.astype(np.float64) - e) / np.spacing(np.abs(e).astype(np.float32)).astype(np.float64)

    assert ulps(r, ex).max() <= 2.5
    assert ulps(inv, 1.0 / ex).max() <= 2.0
    r0, _ = O.root_f32(np.zeros(1, np.float32))
    assert r0[0] == 0.0
    sub = np.exp2(rng.uniform(-149, -124, 10_000)).astype(np.float32)
    rs, ins = O.root_f32(sub)
    assert np.isfinite(rs).all() and np.isfinite(ins).all()
    assert (rs < np.nextafter(np.float32(1e-4), np.float32(1))).all()
    h = np.float32(0.0457)
    assert (np.float32(h) - rs == h).all()


def test_fp32_row_width_within_margin():
    """The FP32 clipped-row half-width is within 2e-3 of sqrt(x) on [0, 1]
    (the 2^-6 row margin covers it; a planted error above the margin fails
    the clipped-row superset tests), exact at 0 and finite everywhere."""
    x = np.concatenate([np.linspace(0.0, 1.0, 400_001), np.exp2(-np.arange(0, 149, 0.25))]).astype(np.float32)
    w = O.row_width_f32(x)
    assert np.isfinite(w).all()
    assert np.abs(w.astype(np.float64) - np.sqrt(x.astype(np.float64))).max() <= 2e-3
    assert O.row_width_f32(np.zeros(1))[0] == 0.0


def test_neighbour_brute_force_matches_reference_predicate():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.1, 0.1, size=(400, 3)).astype(np.float32)
    thr = O.pred_threshold(H)
    cnt, off, idx = O.neighbours_brute(pos, thr)
    d = pos[:, None, :] - pos[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    acc = np.sqrt(d2.astype(np.float32)) <= np.float32(H)
    assert np.array_equal(cnt, acc.sum(1))
    for i in range(0, 400, 37):
        assert np.array_equal(idx[off[i]:off[i + 1]], np.nonzero(acc[i])[0])


def test_terrain_mesh_matches_reference_grid():
    """The oracle's mesh of a float heightfield equals the reference's
    Grid::UpdateGrid / genIndices (grid.h:118-176) bit for bit in voxel units
    (cell 1, cell centres at integer x, z), fixture mesh_ref.npz."""
    g = np.load(os.path.join(GOLD, "mesh_ref.npz"))
    h = g["h_xz"].astype(np.float32)  # h[x, z] = GetHeightfieldAt(x, z)
    d = h.shape[0]
    v = O.terrain_mesh(h.T, 1.0, (-0.5, 0.0, -0.5))
    assert v.shape == g["verts"].shape
    assert np.array_equal(v.view(np.uint32), g["verts"].view(np.uint32))
    assert np.array_equal(O.terrain_indices(d, d), g["indices"])
