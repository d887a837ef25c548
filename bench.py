#!/usr/bin/env python3
"""Benchmark of the SPH-Erosion step on one MI355X.

Workload (default): the single large scene, 262,144 particles in a column
(s = 1/256, h = 2.15 s) dropped onto a 512 x 512 value-noise heightfield
(6 octaves, amplitude 0.3), dt = 5e-4, erosion on (atomic), FP32 with
compensated sums.  A step is one pass of the whole hot path (sort, density,
forces + integrate + collisions + erosion, terrain heights) over the scene.

Prints one JSON line: particle-steps/s over the timed region (K steps as
the library runs them by default, one HIP graph per step, inputs resident in
HBM); the same K steps replayed from a saved state kernel by kernel, with
each kernel's start and end in HIP events on the library's stream, for the
per-kernel times and the HBM roofline of the dominant kernel (the replay
must end in a bit-identical state); the density error against the
library's FP64 path on the final state; the accuracy protocol (one step in
FP32 compensated, FP32 plain and FP64 from a shared state); the plain FP64
GPU path over the same steps; and the reference step timed on the host's
cores (the CPU oracle's restatement of it, the only use of the oracle here).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "sph-erosion_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    from sphe import scenes
    if name == "cfg2":
        return [scenes.cfg2()], "column 262,144 particles on 512^2 value-noise terrain (6 oct, amp 0.3), h=2.15/256, dt=5e-4, erosion atomic"
    if name == "cfg1":
        return [scenes.cfg1()], "dam break 4,096 particles on 64^2 terrain, h=0.0457, dt=1e-3, erosion atomic"
    if name == "cfg3":
        return scenes.cfg3(), "256 batched dam-break scenes, N log-uniform [2048, 65536], T in {64..256}"
    if name == "cfg4":
        return [scenes.cfg4()], "layer of 4,194,304 particles over ridged 2048^2 terrain, dt=5e-4"
    if name == "cfg5":
        return scenes.cfg5(), ("1,024 batched dam-break scenes, N log-uniform [1024, 16384], plus one "
                               "8,388,608-particle layer on a 4096^2 terrain, dt=5e-4")
    if name == "cfg5_share":
        sc = scenes.cfg5_share()
        return [sc], (f"one GPU's x-slab of cfg5's 8,388,608-particle layer: {sc.n:,} particles "
                      "(1/8 of the layer plus a one-cell ghost layer) on the 4096^2 terrain")
    raise SystemExit(f"unknown workload {name}")


def snapshot(ctx, sids, f32=False):
    """Per scene: positions, velocities (rounded to FP32 if f32), ids,
    sediment and terrain volumes, as the library reads them back."""
    out = []
    for s in sids:
        p, v = ctx.state_f64(s)
        if f32:
            p = p.astype(np.float32).astype(np.float64)
            v = v.astype(np.float32).astype(np.float64)
        nx, _ = ctx.terrain_dims(s)
        out.append(dict(pos=p, vel=v, ids=np.asarray(ctx.ids(s), dtype=np.int32), sed=ctx.sediment_fx(s),
                        vol=ctx.terrain_volume_fx(s) if nx else None))
    return out


def load_state(ctx, sids, states):
    for s, st in zip(sids, states):
        ctx.set_particles(s, st["pos"], vel=st["vel"], ids=st["ids"], sed=st["sed"])
        if st["vol"] is not None:
            ctx.set_terrain_volume_fx(s, st["vol"])


def context_like(scs, precision, blk, rm):
    import sphe
    c = sphe.Context(precision)
    sids = [c.add_synthetic(s) for s in scs]
    c.set_tuning(blk, rm)
    return c, sids


def density_error(ctx, sids, scs, blk, rm):
    """Max relative density error of the FP32 pass against the library's FP64
    path on the same state (the state the timed region ended in)."""
    import sphe
    c64, s64 = context_like(scs, sphe.FP64, blk, rm)
    load_state(c64, s64, snapshot(ctx, sids))
    rho = np.concatenate([ctx.debug_density(s) for s in sids])
    ref = np.concatenate([c64.debug_density(s) for s in s64])
    c64.close()
    return float((np.abs(rho - ref) / ref).max())


def accuracy_protocol(scs, blk, rm, warm=50):
    """50 FP64 warm-up steps; the state rounded to FP32; from it, one step in
    FP32 compensated, FP32 plain and FP64: density and acceleration of that
    step, terrain height and sediment changes, each against FP64."""
    import sphe
    c64, s64 = context_like(scs, sphe.FP64, blk, rm)
    c64.step(warm)
    c64.sync()
    st = snapshot(c64, s64, f32=True)
    c64.close()
    res = {}
    for name, prec in (("fp64", sphe.FP64), ("fp32_compensated", sphe.FP32), ("fp32_plain", sphe.FP32_NAIVE)):
        c, ss = context_like(scs, prec, blk, rm)
        load_state(c, ss, st)
        rho = np.concatenate([c.debug_density(s) for s in ss])
        acc = np.concatenate([c.debug_accel(s) for s in ss])
        h0 = [c.terrain_heights(s).astype(np.float64) for s in ss]
        d0 = [c.sediment_fx(s) for s in ss]
        c.step(1)
        dh = np.concatenate([(c.terrain_heights(s) - h).ravel() for s, h in zip(ss, h0)])
        ds = np.concatenate([c.sediment_fx(s) - d for s, d in zip(ss, d0)])
        res[name] = (rho, acc, dh, ds)
        c.close()
    rho64, acc64, dh64, ds64 = res["fp64"]
    n64 = np.linalg.norm(acc64, axis=1)
    out = dict(protocol=f"{warm} FP64 warm-up steps, state rounded to FP32, one step per precision from it; "
                        "errors against the FP64 step",
               particles=int(len(rho64)))
    for name in ("fp32_compensated", "fp32_plain"):
        rho, acc, dh, ds = res[name]
        rr = np.abs(rho - rho64) / rho64
        ok = n64 > 0
        ar = np.linalg.norm(acc - acc64, axis=1)[ok] / n64[ok]
        out[name] = dict(rho_rel_max=float(rr.max()), rho_rel_p99=float(np.percentile(rr, 99)),
                         acc_rel_p50=float(np.median(ar)), acc_rel_p99=float(np.percentile(ar, 99)),
                         acc_rel_max=float(ar.max()),
                         height_change_abs_max=float(np.abs(dh - dh64).max()) if dh.size else 0.0,
                         height_change_fp64_max=float(np.abs(dh64).max()) if dh.size else 0.0,
                         sediment_change_differs=int((ds != ds64).sum()),
                         sediment_change_rel_max=(float(np.abs(ds - ds64).max() / np.abs(ds64).max())
                                                  if ds.size and np.abs(ds64).max() > 0 else 0.0))
    return out


def cpu_baseline(sc, target_s=12.0):
    """The reference step (oracle ref32: all pairs, AoS, FP32, OpenMP over i)
    on this host's cores, over a strided sample of i (each still over all j)."""
    import oracle as O
    n = sc.n
    P = np.zeros(n, dtype=O.PARTICLE_DTYPE)
    P["Id"] = np.arange(n)
    # the reference box is centred at 0: shift the unit box to [-0.5, 0.5]
    P["Position"] = (sc.pos - 0.5).astype(np.float32)
    P["Velocity"] = sc.vel.astype(np.float32)
    prm = O.default_params(dt=sc.params.dt)
    prm.h, prm.mass, prm.p0, prm.k = sc.params.h, sc.params.mass, sc.params.rest_density, sc.params.stiffness
    prm.visc, prm.surf_tens, prm.len = sc.params.visc, sc.params.surf_tens, 0.5
    prm.g[0], prm.g[1], prm.g[2] = sc.params.gravity[0], sc.params.gravity[1], sc.params.gravity[2]
    stride = max(1, n // 256)
    t = time.perf_counter()
    m = O.ref32_step_sampled(P, prm, stride)
    dt0 = time.perf_counter() - t
    per_i = dt0 / max(m, 1)
    want = int(target_s / max(per_i, 1e-9))
    stride2 = max(1, n // max(want, 1))
    t = time.perf_counter()
    m2 = O.ref32_step_sampled(P, prm, stride2)
    dt1 = time.perf_counter() - t
    return dict(value=m2 / dt1, unit="particle-steps/s", cores=O.ref32_threads(), kind="port",
                sample=f"reference step (all pairs, 3 passes over all {n} j) for every {stride2}-th i "
                       f"= {m2} particles, {dt1:.1f} s; particle-steps/s = particles processed / time")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--trace", type=int, default=0, help="steps to trace before the warmup (diagnostics)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the timed region (for profiler runs): no statistics, accuracy, FP64 or CPU legs")
    ap.add_argument("--mode", default="step", choices=["step", "copy"],
                    help="copy: time a plain 1 GiB device copy (HBM reference and counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import sphe
    if a.mode == "copy":
        ctx = sphe.Context(sphe.FP32)
        nbytes = 1 << 30
        ms = ctx.debug_copy(nbytes, 20)
        print(json.dumps(dict(mode="copy", bytes=nbytes, avg_ms=ms, gbs=2 * nbytes / (ms * 1e-3) / 1e9)), flush=True)
        return
    if a.no_extras:
        a.no_cpu = a.no_fp64 = a.no_tune = True
    log(f"bench: {sphe.lib().sphe_version().decode()} workload {a.workload}")
    scs, desc = make_workload(a.workload)
    N = sum(s.n for s in scs)

    ctx = sphe.Context(sphe.FP32)
    sids = [ctx.add_synthetic(s) for s in scs]
    if not a.no_tune:
        blk, rm = ctx.autotune(5)
        log(f"bench: autotuned block={blk} force_block={ctx.get_force_block()} xcd_remap={rm}")
    blk, rm = ctx.get_tuning()
    fblk = ctx.get_force_block()
    if a.trace:
        # ms/step and neighbour counts along the trajectory, 50-step chunks
        for c in range(0, a.trace, 50):
            t = time.perf_counter()
            ctx.step(50)
            ctx.sync()
            ms = (time.perf_counter() - t) * 1e3 / 50
            cnt = ctx.debug_neighbours(sids[0], csr=False)[0]
            log(f"trace: steps {c + 50:5d} ms/step {ms:.3f} neighbours mean {cnt.mean():.1f} max {cnt.max()}")
    ctx.step(a.warmup)
    ctx.sync()
    ctx.save_state()
    mass0 = [ctx.mass(s) for s in sids]

    # timed region: K steps as the library runs them by default (each step
    # one HIP graph launch on the library's stream), inputs resident in HBM
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(a.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    status = ctx.status()
    mass1 = [ctx.mass(s) for s in sids]
    drift = max(abs((m1[0] + m1[1]) - (m0[0] + m0[1])) / (m0[0] + m0[1]) for m0, m1 in zip(mass0, mass1))
    units_exact = all(m0[2] == m1[2] for m0, m1 in zip(mass0, mass1))
    end_state = [ctx.state_f64(s) for s in sids]

    # the same K steps again from the saved state, kernel by kernel: each
    # kernel's start and end land in HIP events on the library's stream
    ctx.restore_state()
    ctx.profile(True)
    ctx.sync()
    tp = time.perf_counter()
    ctx.step(a.steps)
    ctx.sync()
    prof_elapsed = time.perf_counter() - tp
    kernels = {}
    for name in ctx.profile_names():
        ms, cnt = ctx.profile_read(name)
        kernels[name] = dict(avg_ms=ms, launches=cnt, bytes=ctx.kernel_bytes(name) if name != "two_phase" else None)
    ctx.profile(False)
    replay_identical = all(np.array_equal(e[0], r[0]) and np.array_equal(e[1], r[1])
                           for e, r in zip(end_state, [ctx.state_f64(s) for s in sids]))

    value = N * a.steps / elapsed
    dom = max((k for k in kernels if kernels[k]["bytes"]), key=lambda k: kernels[k]["avg_ms"])
    dk = kernels[dom]
    achieved = dk["bytes"] / (dk["avg_ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            tr = json.load(f)
        traffic = tr.get(a.workload, {}).get(dom)
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic, kernel=dom,
                    bytes_per_launch=dk["bytes"], avg_ms=dk["avg_ms"])

    out = dict(metric="particle-steps/s at <=1e-5 rel density err vs FP64; HBM GB/s fraction",
               value=value, unit="particle-steps/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=1e3 * elapsed / a.steps, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="fp32", data="synthetic",
               config=dict(workload=f"{a.workload}: {desc}", particles=N, scenes=len(scs),
                           precision="fp32 (compensated density sums, FP64-accumulated force sums)", block=blk, force_block=fblk, density_xcd_run=rm,
                           x_slabs=ctx.get_subcells(), yz_split=ctx.get_yz_split()),
               roofline=roofline)
    out["value_kernel_by_kernel"] = N * a.steps / prof_elapsed
    out["replay_bit_identical"] = replay_identical
    out["kernels_ms"] = {k: round(v["avg_ms"], 5) for k, v in kernels.items()}
    out["mass_drift_rel"] = drift
    out["mass_units_exact"] = units_exact
    out["nonfinite"] = status

    if a.no_extras:
        print(json.dumps(out), flush=True)
        return
    # neighbour statistics of the state the timed region ended in
    counts = np.concatenate([ctx.debug_neighbours(s, csr=False)[0] for s in sids])
    rho = np.concatenate([ctx.debug_density(s) for s in sids])
    rho0 = np.array([s.params.rest_density for s in scs for _ in range(s.n)])
    out["neighbours"] = dict(mean=float(counts.mean()), p50=float(np.median(counts)),
                             p99=float(np.percentile(counts, 99)), max=int(counts.max()))
    out["density_over_rest"] = dict(mean=float((rho / rho0).mean()), max=float((rho / rho0).max()))

    log("bench: density error vs the FP64 path")
    out["density_rel_err_max"] = density_error(ctx, sids, scs, blk, rm)
    log("bench: accuracy protocol")
    # after 50 steps as specified; after 400 the cfg2 column is on the
    # terrain, so the height and sediment changes are compared too
    out["accuracy"] = accuracy_protocol(scs, blk, rm, 50)
    out["accuracy_400"] = accuracy_protocol(scs, blk, rm, 400)

    if not a.no_fp64:
        log("bench: FP64 plain GPU path")
        c64 = sphe.Context(sphe.FP64)
        for s in scs:
            c64.add_synthetic(s)
        c64.set_tuning(blk, rm)
        c64.set_force_block(fblk)
        c64.step(a.warmup)
        c64.sync()
        t = time.perf_counter()
        c64.step(a.steps)
        c64.sync()
        v64 = N * a.steps / (time.perf_counter() - t)
        out["fp64_value"] = v64
        out["speedup_vs_fp64"] = value / v64
        c64.close()

    if not a.no_cpu and len(scs) == 1:
        log("bench: CPU baseline (reference step on host cores)")
        out["cpu_baseline"] = cpu_baseline(scs[0], a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
