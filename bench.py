#!/usr/bin/env python3
"""Benchmark of the SPH-Erosion step on one MI355X.

Workload (default): the long-horizon batched production configuration,
1,024 dam-break scenes (N log-uniform in [1,024, 16,384], terrains 64..256
cells square) plus one 8,388,608-particle layer on a 4096^2 value-noise
terrain, 13,954,298 particles in one context, dt = 5e-4, erosion on
(atomic), FP32 with Kahan-compensated density and force sums.  A step is one
pass of the whole hot path (sort, density, forces + integrate + collisions +
erosion, terrain heights) over every scene.

The run is a window of the workload's trajectory: `--settle` steps first
(untimed; by default 1,000 of the 2,000-step run, where the water is on the
terrain and eroding), then `--warmup` steps, then the K timed steps.

Prints one JSON line:
  * value: particle-steps/s over the K timed steps as the library runs them by
    default (captured step sequences, inputs resident in HBM);
  * the same K steps replayed from a saved state kernel by kernel, each
    kernel's start and end in HIP events on the library's stream: per-kernel
    times, the HBM roofline of the dominant kernel and of the force pass (the
    replay must end in a bit-identical state);
  * terrain contact in the window: particles that exchanged sediment with the
    terrain, terrain cells that changed;
  * water/sediment bookkeeping per scene over the window (FP64 reductions,
    exact fixed-point units);
  * density error against the library's FP64 path on the final state; the
    accuracy protocol (one step in each FP32 mode and FP64 from a shared
    state); the front end's frame loop (a step, then the largest scene's
    positions read back to the host); the plain FP64 GPU path over the same
    window; and the reference step on the host's cores (the CPU oracle's
    restatement of it, the only use of the oracle here).
Exits with status 3 when the step reported a non-finite state or a particle it
could not place.

`--mode horizon` instead runs the workload's whole trajectory from step 0 in
chunks of 250 steps and prints one JSON line with the time per chunk and the
conservation after each (exit status 3 if the units drift or bits 1 or 2 are
set); `--mode copy` and `--mode summary` serve the rocprofv3 profiles.
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
SIMDS = 256 * 4        # MI355X: 256 CUs, 4 SIMDs each
CLOCK_GHZ = 2.4        # MI355X peak engine clock
COUNTERS = os.path.join(ROOT, "profiles", "counters.json")
SETTLE = {"cfg1": 100, "cfg2": 400, "cfg3": 500, "cfg4": 600, "cfg5": 1000, "cfg5_share": 1000}


T_START = time.perf_counter()


def log(msg):
    print(f"[{time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def make_workload(name):
    from sphe import scenes
    if name == "cfg1":
        return [scenes.cfg1()], "dam break 4,096 particles on 64^2 terrain, h=0.0457, dt=1e-3, erosion atomic", 200
    if name == "cfg2":
        return ([scenes.cfg2()], "column 262,144 particles on 512^2 value-noise terrain (6 oct, amp 0.3), "
                "h=2.15/256, dt=5e-4, erosion atomic", 500)
    if name == "cfg3":
        return scenes.cfg3(), "256 batched dam-break scenes, N log-uniform [2048, 65536], T in {64..256}", 300
    if name == "cfg4":
        return [scenes.cfg4()], "layer of 4,194,304 particles over ridged 2048^2 terrain, dt=5e-4", 1000
    if name == "cfg5":
        return scenes.cfg5(), ("1,024 batched dam-break scenes, N log-uniform [1024, 16384], plus one "
                               "8,388,608-particle layer on a 4096^2 terrain, dt=5e-4, erosion atomic"), 2000
    if name == "cfg5_share":
        sc = scenes.cfg5_share()
        return [sc], (f"one GPU's x-slab of cfg5's 8,388,608-particle layer: {sc.n:,} particles "
                      "(1/8 of the layer plus a one-cell ghost layer) on the 4096^2 terrain"), 2000
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


def save_states(ctx, sids, path, settled):
    """Every scene's particles (FP32 values, id order) and terrain volumes,
    concatenated, with the per-scene counts, in one .npz."""
    st = snapshot(ctx, sids)
    np.savez(path, settled=settled, n=np.array([len(s["ids"]) for s in st]),
             pos=np.concatenate([s["pos"] for s in st]).astype(np.float32),
             vel=np.concatenate([s["vel"] for s in st]).astype(np.float32),
             ids=np.concatenate([s["ids"] for s in st]), sed=np.concatenate([s["sed"] for s in st]),
             nvol=np.array([0 if s["vol"] is None else s["vol"].size for s in st]),
             vol=np.concatenate([np.zeros(0, np.int64)] + [s["vol"].ravel() for s in st if s["vol"] is not None]))


def load_states(ctx, sids, path):
    with np.load(path) as d:  # each member read once (NpzFile reads on every access)
        ns, nvs, pos, vel, ids, sed, vol = (d[k] for k in ("n", "nvol", "pos", "vel", "ids", "sed", "vol"))
    p0 = v0 = 0
    for s, n, nv in zip(sids, ns, nvs):
        ctx.set_particles(s, pos[p0:p0 + n], vel=vel[p0:p0 + n], ids=ids[p0:p0 + n], sed=sed[p0:p0 + n])
        if nv:
            ctx.set_terrain_volume_fx(s, vol[v0:v0 + nv])
        p0 += n
        v0 += nv


def context_from(scs, states, precision, tuning):
    import sphe
    c = sphe.Context(precision)
    sids = [c.add_scene(s.params, s.terrain) for s in scs]
    for s, st in zip(sids, states):
        c.set_particles(s, st["pos"], vel=st["vel"], ids=st["ids"], sed=st["sed"])
        if st["vol"] is not None:
            c.set_terrain_volume_fx(s, st["vol"])
    c.set_yz_split(tuning[3])
    c.set_subcells(tuning[4])
    c.set_scan_tile(tuning[5])
    c.set_tuning(tuning[0], tuning[1])
    c.set_force_block(tuning[2])
    return c, sids


def density_error(ctx, sids, scs, end_state, tuning):
    """Max relative density error of the FP32 pass against the library's FP64
    path on the same state (the state the timed window ended in)."""
    import sphe
    c64, s64 = context_from(scs, end_state, sphe.FP64, tuning)
    rho = np.concatenate([ctx.debug_density(s) for s in sids])
    ref = np.concatenate([c64.debug_density(s) for s in s64])
    c64.close()
    return float((np.abs(rho - ref) / ref).max())


def front_end_loop(ctx, sids, scs, steps):
    """The drop-in front end's frame loop (main.cpp:313 Run, then drawing):
    one step, then the positions of the largest scene read back into a host
    buffer the caller keeps (12 bytes per particle over PCIe), `steps` times.
    Not the headline rate: it includes the read-back and a synchronisation
    per step."""
    import ctypes as C

    import sphe
    big = max(range(len(sids)), key=lambda i: scs[i].n)
    n_big = scs[big].n
    buf = np.zeros((n_big, 3), dtype=np.float32)
    ptr = buf.ctypes.data_as(C.c_void_p)

    def read():
        rc = sphe.lib().sphe_read_buffer(ctx._p, sids[big], sphe.BUF_POS, ptr, buf.nbytes)
        if rc != 0:
            raise RuntimeError(f"sphe_read_buffer: {rc}")

    read()  # the host buffer's pages are touched once before the clock starts
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        ctx.step(1)
        read()
    el = time.perf_counter() - t
    n = sum(s.n for s in scs)
    return dict(steps=steps, readback_particles=n_big, readback_bytes_per_step=12 * n_big,
                ms_per_step=1e3 * el / steps, particle_steps_per_s=n * steps / el)


def horizon(workload, chunk):
    """The workload's whole run from its initial state, as one context steps
    it by default (FP32 compensated, untuned launch shapes), in chunks of
    `chunk` steps: per chunk the time per step, the state of the neighbour
    lists and pool, whether the fixed-point units are still exact, and the
    FP64 relative drift of water plus sediment per scene."""
    import sphe
    scs, desc, run_steps = make_workload(workload)
    ctx = sphe.Context(sphe.FP32)
    sids = [ctx.add_synthetic(s) for s in scs]
    t0m, s0m, u0m = ctx.mass_all()
    tot0 = t0m + s0m
    n = sum(s.n for s in scs)
    chunks, done, total_s, exact, drift_max, bits = [], 0, 0.0, True, 0.0, 0
    while done < run_steps:
        k = min(chunk, run_steps - done)
        ctx.sync()
        t = time.perf_counter()
        ctx.step(k)
        ctx.sync()
        el = time.perf_counter() - t
        done += k
        total_s += el
        st = ctx.status()
        t1m, s1m, u1m = ctx.mass_all()
        drift = float((np.abs((t1m + s1m) - tot0) / np.where(tot0 > 0, tot0, 1.0)).max())
        exact = exact and bool(np.array_equal(u0m, u1m))
        drift_max = max(drift_max, drift)
        bits |= st
        chunks.append(dict(steps=done, ms_per_step=1e3 * el / k, status=st, units_exact=exact, drift=drift,
                           list_capacity=ctx.get_list_capacity(), overflow=list(ctx.get_overflow_capacity())))
        log(f"bench: horizon {workload} step {done}: {1e3 * el / k:.2f} ms/step, status {st}, "
            f"units exact {exact}, drift {drift:.2e}")
    ctx.close()
    print(json.dumps(dict(mode="horizon", workload=f"{workload}: {desc}", particles=n, steps=run_steps,
                          seconds=total_s, ms_per_step=1e3 * total_s / run_steps,
                          particle_steps_per_s=n * run_steps / total_s, units_exact=exact,
                          mass_drift_rel_max=drift_max, status_bits=bits, chunks=chunks)), flush=True)
    if bits & 3 or not exact:
        sys.exit(3)


def accuracy_protocol(scs, start_state, tuning, warm=50):
    """From the state the timed window started in: `warm` FP64 steps; that
    state rounded to FP32; from it one step in FP64 and in each FP32 mode:
    density and acceleration of the step, terrain height and sediment
    changes, each against FP64."""
    import sphe
    c64, s64 = context_from(scs, start_state, sphe.FP64, tuning)
    c64.step(warm)
    c64.sync()
    st = snapshot(c64, s64, f32=True)
    c64.close()
    res = {}
    for name, prec in (("fp64", sphe.FP64), ("fp32_compensated", sphe.FP32),
                       ("fp32_f64_force_sums", sphe.FP32_F64ACC), ("fp32_plain", sphe.FP32_NAIVE)):
        c, ss = context_from(scs, st, prec, tuning)
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
    out = dict(protocol=f"state at the window start, {warm} FP64 steps, rounded to FP32; one step per "
                        "precision from it; errors against the FP64 step",
               particles=int(len(rho64)), exchanging_particles=int((ds64 != 0).sum()))
    for name in ("fp32_compensated", "fp32_f64_force_sums", "fp32_plain"):
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def ref_records(sc):
    """The scene as the reference's particle records: its box is centred at 0,
    so the unit box [0, 1]^3 shifts to [-0.5, 0.5]^3 (len 0.5)."""
    import oracle as O
    P = np.zeros(sc.n, dtype=O.PARTICLE_DTYPE)
    P["Id"] = np.arange(sc.n)
    P["Position"] = (sc.pos - 0.5).astype(np.float32)
    P["Velocity"] = sc.vel.astype(np.float32)
    prm = O.default_params(dt=sc.params.dt)
    prm.h, prm.mass, prm.p0, prm.k = sc.params.h, sc.params.mass, sc.params.rest_density, sc.params.stiffness
    prm.visc, prm.surf_tens, prm.len = sc.params.visc, sc.params.surf_tens, 0.5
    prm.g[0], prm.g[1], prm.g[2] = sc.params.gravity[0], sc.params.gravity[1], sc.params.gravity[2]
    return P, prm


def ref_time_per_i(sc, target_s):
    """Seconds of the reference step per particle i (each i over all n j, three
    passes) from a strided sample of i sized to about target_s seconds."""
    import oracle as O
    P, prm = ref_records(sc)
    stride = max(1, sc.n // 64)
    t = time.perf_counter()
    m = O.ref32_step_sampled(P, prm, stride)
    per_i = (time.perf_counter() - t) / max(m, 1)
    stride2 = max(1, sc.n // max(1, int(target_s / max(per_i, 1e-12))))
    t = time.perf_counter()
    m2 = O.ref32_step_sampled(P, prm, stride2)
    return (time.perf_counter() - t) / max(m2, 1), m2, stride2


def cpu_baseline(scs, target_s=12.0):
    """The reference step (the oracle's ref32: all pairs, AoS, FP32, OpenMP over
    i) on this host's cores.  The cost of one step of every scene is built
    from samples and extrapolated: a scene above 100,000 particles from a
    strided sample of its i (each i still over all j); the smaller scenes from
    whole steps of up to 24 of them, scaled by n^2 (all-pairs cost)."""
    import oracle as O
    big = [s for s in scs if s.n > 100_000]
    small = [s for s in scs if s.n <= 100_000]
    total_t = 0.0
    parts = []
    budget = target_s / (2 if big and small else 1)
    for sc in big:
        per_i, m, stride = ref_time_per_i(sc, budget / len(big))
        total_t += per_i * sc.n
        parts.append(f"{sc.n:,}-particle scene: every {stride}-th i ({m:,} of them, each over all j)")
    if small:
        order = np.argsort([s.n for s in small])
        pick = [small[i] for i in order[np.linspace(0, len(small) - 1, min(24, len(small))).astype(int)]]
        spent, cost, n2 = 0.0, 0.0, 0.0
        for sc in pick:
            P, prm = ref_records(sc)
            t = time.perf_counter()
            O.ref32_step(P, prm, 1)
            dt = time.perf_counter() - t
            cost += dt
            n2 += float(sc.n) ** 2
            spent += dt
            if spent > budget:
                break
        total_t += cost / n2 * sum(float(s.n) ** 2 for s in small)
        parts.append(f"{len(small)} scenes of <= 100,000 particles: whole steps of {len(pick)} spread over "
                     "their sizes, scaled by n^2")
    N = sum(s.n for s in scs)
    extrapolated = len(scs) > 1 or bool(big)
    return dict(value=N / total_t, unit="particle-steps/s", cores=O.ref32_threads(), kind="port",
                cpu_model=cpu_model(), extrapolated=extrapolated,
                sample="reference step (all pairs, 3 passes over all j, FP32, AoS) timed on " + "; ".join(parts) +
                       f"; one step of all {N:,} particles would take {total_t:.1f} s; "
                       "particle-steps/s = particles / that time")


def roofline(name, k, counters):
    achieved = k["bytes"] / (k["avg_ms"] * 1e-3) / 1e9
    c = counters.get(name, {})
    out = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
               traffic=c.get("hbm_bytes"), kernel=name, bytes_per_launch=k["bytes"], avg_ms=k["avg_ms"])
    if c.get("valu_insts") is not None:
        # VALU issue: each wave64 VALU instruction occupies a 16-lane SIMD for
        # 4 cycles (an upper bound for part of the mix), over the kernel time
        out["valu_issue_frac"] = c["valu_insts"] * 4 / (SIMDS * CLOCK_GHZ * 1e9 * k["avg_ms"] * 1e-3)
    return out


def summarize_profiles(base, workload, steps):
    """Summary of the rocprofv3 runs of one window (run from --load-state
    with --no-extras): the kernel trace (prof_trace), one --pmc run each for
    FETCH_SIZE, WRITE_SIZE, the SQ instruction counts (prof_SQ) and the stall
    counters (prof_STALL), and FETCH_SIZE / WRITE_SIZE of the 1 GiB copy
    (prof_copy_FETCH_SIZE, prof_copy_WRITE_SIZE), each written with -o run
    under base/<name>.  The last `steps` launches of a kernel are the replay
    (direct launches), the `steps` before them the window (graph launches).
    Writes base/prof_summary.md and base/prof_counters.json (the per-kernel
    HBM bytes and VALU instructions this bench reads from
    profiles/counters.json)."""
    import collections
    import re
    import sqlite3

    def short(n):
        m = re.search(r"(k_[a-z0-9_]+)", n)
        return m.group(1) if m else n[:40]

    def rows(name, sql):
        db = os.path.join(base, name, "run_results.db")
        if not os.path.exists(db):
            return []
        con = sqlite3.connect(db)
        try:
            return list(con.execute(sql))
        finally:
            con.close()

    def counters(name):
        out = collections.defaultdict(lambda: collections.defaultdict(list))
        for kname, cname, val in rows(name, "select kernel_name, counter_name, value from counters_collection "
                                            "order by dispatch_id"):
            out[short(kname)][cname].append(val)
        return out

    def mean(v):
        return sum(v) / len(v) if v else None

    def win(v):  # the window's and the replay's launches (not the ones before)
        return v[-2 * steps:]

    durs = collections.defaultdict(list)
    for name, dur in rows("prof_trace", "select name, duration from kernels order by start"):
        durs[short(name)].append(dur / 1e3)  # us
    total = sum(sum(v) for v in durs.values())
    fetch, write = counters("prof_FETCH_SIZE"), counters("prof_WRITE_SIZE")
    sq, stall = counters("prof_SQ"), counters("prof_STALL")
    cf = mean(counters("prof_copy_FETCH_SIZE")["k_copy"]["FETCH_SIZE"])
    cw = mean(counters("prof_copy_WRITE_SIZE")["k_copy"]["WRITE_SIZE"])
    gib_kib = (1 << 30) / 1024
    fscale, wscale = gib_kib / cf, gib_kib / cw
    lines = [f"Copy calibration (1 GiB plain copy, `bench.py --mode copy`): FETCH_SIZE {cf:.0f} KiB per launch "
             f"(scale {fscale:.3f}), WRITE_SIZE {cw:.0f} KiB (scale {wscale:.3f}).", "",
             "| kernel | launches | avg us, window (graph) | avg us, replay (direct) | % of GPU time | "
             "FETCH_SIZE KiB | WRITE_SIZE KiB | HBM traffic per launch (MB) | VALU instr per wave | waves |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    out = {}
    for k, v in sorted(durs.items(), key=lambda kv: -sum(kv[1])):
        stepped = len(v) >= 2 * steps
        wnd = mean(v[-2 * steps:-steps]) if stepped else mean(v)
        rep = mean(v[-steps:]) if stepped else mean(v)
        f, w = mean(win(fetch[k]["FETCH_SIZE"])), mean(win(write[k]["WRITE_SIZE"]))
        vi, wv = mean(win(sq[k]["SQ_INSTS_VALU"])), mean(win(sq[k]["SQ_WAVES"]))
        hbm = (f * fscale + w * wscale) * 1024 if f is not None and w is not None else None
        lines.append(f"| {k} | {len(v)} | {wnd:.1f} | {rep:.1f} | {100 * sum(v) / total:.1f} | {f or 0:.0f} | "
                     f"{w or 0:.0f} | {(hbm or 0) / 1e6:.1f} | {(vi / wv) if vi and wv else 0:.0f} | {wv or 0:.0f} |")
        key = k.replace("k_", "").replace("_f32", "").replace("_count", "")
        out[key] = dict(hbm_bytes=hbm, valu_insts=vi, waves=wv, avg_us_replay=rep, avg_us_window=wnd)
        if k in stall:
            out[key]["stall"] = {c: mean(win(x)) for c, x in stall[k].items()}
    lines += ["", "Stall counters per launch (wave-cycles summed over the device; GRBM_GUI_ACTIVE summed over "
                  "the XCDs):", ""]
    for k in ("k_force_f32", "k_density", "k_reorder"):
        for src in (stall, sq):
            if k in src:
                lines.append(f"- {k}: " + ", ".join(f"{c} {mean(win(x)):.3e}" for c, x in sorted(src[k].items())))
    with open(os.path.join(base, "prof_summary.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(base, "prof_counters.json"), "w") as f:
        json.dump({workload: out}, f, indent=1)
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cfg5")
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed steps before the warm-up (default: per workload, into terrain contact)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="only the timed region (for profiler runs): no statistics, accuracy, FP64 or CPU legs")
    ap.add_argument("--mode", default="step", choices=["step", "copy", "summary", "horizon"],
                    help="copy: time a plain 1 GiB device copy (HBM reference and counter calibration); "
                         "summary: summarise the rocprofv3 runs under --prof-dir (no GPU); "
                         "horizon: the workload's whole run from step 0, timed in chunks, with its "
                         "conservation checked at every chunk")
    ap.add_argument("--prof-dir", default=".", help="where the rocprofv3 runs are (--mode summary)")
    ap.add_argument("--save-state", default=None,
                    help="write the state the timed window starts from (after settle and warm-up) to this .npz")
    ap.add_argument("--load-state", default=None,
                    help="start from a state written by --save-state (no settle; give --warmup 0 for the same "
                         "window): lets a profiler see only the window's kernels")
    ap.add_argument("--grid", default=None, help="y/z split, x slabs, log2 scan tile, e.g. 2,4,12 (with --no-tune)")
    ap.add_argument("--block", type=int, default=None, help="density block size (with --no-tune)")
    ap.add_argument("--force-block", type=int, default=None, help="force block size (with --no-tune)")
    ap.add_argument("--sequence-steps", type=int, default=None,
                    help="steps per captured graph launch (the library's default if not given)")
    a = ap.parse_args()
    if a.mode == "summary":
        summarize_profiles(a.prof_dir, a.workload, a.steps)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import sphe
    if a.mode == "copy":
        ctx = sphe.Context(sphe.FP32)
        nbytes = 1 << 30
        ms = ctx.debug_copy(nbytes, 20)
        print(json.dumps(dict(mode="copy", bytes=nbytes, avg_ms=ms, gbs=2 * nbytes / (ms * 1e-3) / 1e9)), flush=True)
        return
    if a.mode == "horizon":
        horizon(a.workload, 250)
        return
    if a.no_extras:
        a.no_cpu = a.no_fp64 = a.no_tune = True
    settle = SETTLE.get(a.workload, 0) if a.settle is None else a.settle
    log(f"bench: {sphe.lib().sphe_version().decode()} workload {a.workload}")
    scs, desc, run_steps = make_workload(a.workload)
    if a.settle is None:  # the window stays inside the workload's run
        settle = max(0, min(settle, run_steps - a.warmup - a.steps))
    N = sum(s.n for s in scs)

    ctx = sphe.Context(sphe.FP32)
    if a.sequence_steps:
        ctx.set_sequence_steps(a.sequence_steps)
    sids = [ctx.add_synthetic(s) for s in scs]
    if a.no_tune:
        if a.grid:
            yz, slabs, tile = (int(x) for x in a.grid.split(","))
            ctx.set_yz_split(yz)
            ctx.set_subcells(slabs)
            ctx.set_scan_tile(tile)
        if a.block:
            ctx.set_tuning(a.block, ctx.get_tuning()[1])
        if a.force_block:
            ctx.set_force_block(a.force_block)
    if a.load_state:
        load_states(ctx, sids, a.load_state)
        with np.load(a.load_state) as d:
            settle = int(d["settled"])
        # the lists grow to the state as the settle steps grew them in the run
        # that saved it: a step publishes the overflow, the next call grows
        # the lists before its launch; then back to the loaded state
        ctx.save_state()
        for _ in range(2):
            ctx.step(1)
            ctx.sync()
        ctx.restore_state()
        log(f"bench: loaded the state after {settle} steps from {a.load_state}")
    if not a.load_state:
        for done in range(0, settle, 250):
            ctx.step(min(250, settle - done))
            ctx.sync()
            log(f"bench: settled {min(done + 250, settle)} of {settle} steps")
    if not a.no_tune:
        # on the settled state: the grid first, then the block sizes on it
        yz, slabs, tile = ctx.autotune_grid(3)
        blk, rm = ctx.autotune(5)
        log(f"bench: autotuned yz_split={yz} x_slabs={slabs} scan_tile=2^{tile} block={blk} "
            f"force_block={ctx.get_force_block()} xcd_remap={rm}")
    blk, rm = ctx.get_tuning()
    fblk = ctx.get_force_block()
    tuning = (blk, rm, fblk, ctx.get_yz_split(), ctx.get_subcells(), ctx.get_scan_tile())
    ctx.step(a.warmup)
    ctx.sync()
    if a.save_state:
        save_states(ctx, sids, a.save_state, settle + a.warmup)
        log(f"bench: saved the window's start state to {a.save_state}")
    ctx.save_state()
    start_state = snapshot(ctx, sids) if not a.no_extras else None
    sed0 = np.concatenate([ctx.sediment_fx(s) for s in sids])
    vol0 = [ctx.terrain_volume_fx(s) for s in sids]
    t0m, s0m, u0m = ctx.mass_all()

    # timed window: K steps as the library runs them by default (captured
    # step sequences on the library's stream), inputs resident in HBM
    log(f"bench: timing {a.steps} steps")
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(a.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    status = ctx.status()
    t1m, s1m, u1m = ctx.mass_all()
    tot0 = t0m + s0m
    drift = np.abs((t1m + s1m) - tot0) / np.where(tot0 > 0, tot0, 1.0)
    units_exact = bool(np.array_equal(u0m, u1m))
    sed1 = np.concatenate([ctx.sediment_fx(s) for s in sids])
    cells_changed = sum(int((ctx.terrain_volume_fx(s) != v).sum()) for s, v in zip(sids, vol0))
    end_state = snapshot(ctx, sids)

    # the same K steps again from the saved state, kernel by kernel: each
    # kernel's start and end land in HIP events on the library's stream
    log("bench: replaying the window kernel by kernel")
    ctx.restore_state()
    ctx.mass_all()  # uploads the restored state before the clock starts
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
    replay = snapshot(ctx, sids)
    replay_identical = all(np.array_equal(e["pos"], r["pos"]) and np.array_equal(e["vel"], r["vel"])
                           and np.array_equal(e["sed"], r["sed"]) for e, r in zip(end_state, replay))

    counters = {}
    if os.path.exists(COUNTERS):
        with open(COUNTERS) as f:
            counters = json.load(f).get(a.workload, {})
    value = N * a.steps / elapsed
    dom = max((k for k in kernels if kernels[k]["bytes"]), key=lambda k: kernels[k]["avg_ms"])
    first = settle + a.warmup + 1
    out = dict(metric="particle-steps/s at <=1e-5 rel density err vs FP64; HBM GB/s fraction",
               value=value, unit="particle-steps/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=1e3 * elapsed / a.steps, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="fp32", data="synthetic",
               config=dict(workload=f"{a.workload}: {desc}", particles=N, scenes=len(scs),
                           window=f"steps {first}..{first + a.steps - 1} of the {run_steps}-step run "
                                  f"({settle} settle + {a.warmup} warm-up steps before it, untimed)",
                           precision="fp32 (Kahan-compensated density and force sums)", block=blk,
                           force_block=fblk, density_xcd_run=rm, x_slabs=ctx.get_subcells(),
                           yz_split=ctx.get_yz_split(), scan_tile_log2=ctx.get_scan_tile(),
                           sequence_steps=ctx.get_sequence_steps()),
               roofline=roofline(dom, kernels[dom], counters))
    out["roofline_force"] = roofline("force", kernels["force"], counters)
    out["value_kernel_by_kernel"] = N * a.steps / prof_elapsed
    out["replay_bit_identical"] = replay_identical
    out["kernels_ms"] = {k: round(v["avg_ms"], 5) for k, v in kernels.items()}
    out["terrain_contact"] = dict(particles_exchanging=int((sed1 != sed0).sum()),
                                  particles_exchanging_frac=float((sed1 != sed0).mean()),
                                  particles_carrying_sediment=int((sed1 > 0).sum()),
                                  terrain_cells_changed=cells_changed)
    out["mass_drift_rel_max_per_scene"] = float(drift.max())
    out["mass_units_exact"] = units_exact
    out["status"] = status
    # bit 4: some particle walked its stencil in the window (lists too short)
    out["lists"] = dict(list_capacity=ctx.get_list_capacity(), overflow_chunk=ctx.get_overflow_capacity()[0],
                        overflow_chunks=ctx.get_overflow_capacity()[1], stencil_walks_in_window=bool(status & 4))

    if not a.no_extras:
        counts = np.concatenate([ctx.debug_neighbours(s, csr=False)[0] for s in sids])
        rho = np.concatenate([ctx.debug_density(s) for s in sids])
        rho0 = np.concatenate([np.full(s.n, s.params.rest_density) for s in scs])
        out["neighbours"] = dict(mean=float(counts.mean()), p50=float(np.median(counts)),
                                 p99=float(np.percentile(counts, 99)), max=int(counts.max()))
        out["density_over_rest"] = dict(mean=float((rho / rho0).mean()), max=float((rho / rho0).max()))
        log("bench: density error vs the FP64 path")
        out["density_rel_err_max"] = density_error(ctx, sids, scs, end_state, tuning)
        log("bench: accuracy protocol")
        out["accuracy"] = accuracy_protocol(scs, start_state, tuning)
        out["front_end_loop"] = front_end_loop(ctx, sids, scs, 10)

    if not a.no_fp64:
        log("bench: FP64 plain GPU path over the same window")
        ctx.close()  # the FP64 context gets the device to itself
        c64, s64 = context_from(scs, start_state, sphe.FP64, tuning)
        # the window run once untimed sizes its neighbour lists for every
        # step of it (as the settle did for the FP32 context), then back to
        # the window's start
        c64.save_state()
        c64.step(a.warmup + a.steps)
        c64.sync()
        c64.restore_state()
        c64.step(a.warmup)
        c64.sync()
        t = time.perf_counter()
        c64.step(a.steps)
        c64.sync()
        v64 = N * a.steps / (time.perf_counter() - t)
        st64 = c64.status()
        out["fp64_value"] = v64
        out["speedup_vs_fp64"] = value / v64
        out["fp64_stencil_walks_in_window"] = bool(st64 & 4)
        c64.close()

    if not a.no_cpu:
        log("bench: CPU baseline (reference step on host cores)")
        out["cpu_baseline"] = cpu_baseline(scs, a.cpu_seconds)
    print(json.dumps(out), flush=True)
    if status & 3:
        log(f"bench: the step reported status {status} (1: non-finite state, 2: a particle it could not place)")
        sys.exit(3)


if __name__ == "__main__":
    main()
