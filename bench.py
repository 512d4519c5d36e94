#!/usr/bin/env python3
"""Throughput of the pixel-track stage (CA doublets -> connect -> fishbone ->
ntuplets -> broken-line fit) on one MI355X.

One step = one pass of the stage over a batch of PU50-like synthetic events
resident in HBM (default 100 events, seeds 1..100: 1500 helices from 50
vertices, pT power law above 0.3 GeV, |eta| < 2.5, +5% noise hits).
Prints ONE JSON line (see the keys below).  Per-kernel durations come from HIP
events recorded on the library's stream after every kernel of the timed steps;
the roofline entry is computed for the kernel with the largest total time.
"""
import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cms-julia_amd"))

METRIC = "events/s CA+fit at PU50 (1 GPU); FP32 fit+adjoint tracks/s vs FP64 fit @1e-4"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

# Flops of one broken-line fit (add/sub + mul + div + sqrt), counted by running
# the fit restatement with an op-counting scalar (tests/test_oracle_fit.py pins
# the counts); transcendental calls are excluded.  N = hits used (5 -> 4).
FIT_FLOPS = {3: 1468, 4: 2007}


# The measurement model of the path: algorithmic HBM bytes per event of each
# stage, from the per-event counters (H hits, D doublets, C neighbour entries,
# T stored tuples, Th their hits), the bytes each stage must move at least once.
STAGES = {
    # stage: (kernels that implement it, bytes per event)
    "doublets (F2)": (("k_doublet_count", "k_doublet_scan", "k_doublet_fill", "k_ohc_rebuild"),
                      lambda H, D, C, T, Th: 16 * H + 32 * D + 5168),
    "connect (F3)": (("k_connect",), lambda H, D, C, T, Th: 8 * H + 34 * D + 4 * C),
    "fishbone (F4)": (("k_fish_list", "k_fishbone"), lambda H, D, C, T, Th: 12 * H + 32 * D),
    "ntuplets, duplicates, multiplicity (F5-F8)": (
        ("k_dfs_count", "k_dfs_scan", "k_dfs_emit", "k_tuples_post", "k_multiplicity"),
        lambda H, D, C, T, Th: 28 * D + 4 * C + 6 * Th + 12 * T),
}
STAGE_OF_KERNEL = {k: st for st, (ks, _) in STAGES.items() for k in ks}
BOX_CPU_SHARE = 16  # host CPUs granted to a one-GPU job on the measurement boxes


def counts(c):
    return (c["n_hits"], c["n_cells"], c["n_connections"], np.minimum(c["n_tuples"], 32768),
            np.minimum(c["n_tuple_hits"], 163840))


def stage_bytes(stage, c):
    """Algorithmic bytes of one stage over the batch (sum over the events)."""
    return int(np.sum(STAGES[stage][1](*counts(c))))


def fit_bytes(c):
    """24 N + 98 bytes per fitted track (N = 3, or 4 for 4- and 5-hit tuples,
    the fifth hit's id read too), 24,576 fits per length at most."""
    nb = c["n_by_len"]
    f3, f4, f5 = (np.minimum(nb[:, k], 24576) for k in (3, 4, 5))
    return int(np.sum(170 * f3 + 194 * f4 + 196 * f5))


def kernel_model_bytes(name, c):
    """A finer per-kernel byte model (what each kernel's own data layout reads
    and writes once: records, lists, counters), summed over the events.
    c: dict of per-event counter arrays."""
    H, S, D = c["n_hits"], c["n_inner_slots"], c["n_cells"]
    C, T = c["n_connections"], np.minimum(c["n_tuples"], 32768)
    Th = np.minimum(c["n_tuple_hits"], 163840)
    hist = 1281 * 4 + 11 * 4
    if name == "k_doublet_count":
        # inner-hit record per slot 16 B, histogram-ordered hit records 16 B per hit,
        # histogram offsets; out: count 4 + mask 8 + candidate range 16 B per slot
        b = 16 * S + 16 * H + hist + 28 * S
    elif name == "k_doublet_fill":
        # per slot: count, mask, range (28 B) in, first cell id (4 B) out; per cell:
        # outer-hit record + xy (24 B) in, cell record 4 + pair 1 + flags 1 + count 1
        # + outer-hit coordinates 16 + outer-hit list entry 8 + cursor 4 B out
        b = 32 * S + 24 * D + 35 * D
    elif name == "k_connect":
        # per hit: list cursor 4 B, slot ranges 8 B, coordinates 16 B; per cell: list
        # entry 8 B, inner-hit coordinates 16 B + module 2 B, outer-hit coordinates of
        # the cell as a candidate 16 B, neighbour count 1 B; per neighbour entry 4 B
        b = 28 * H + 43 * D + 4 * C
    elif name == "k_fishbone":
        # per hit: cursor + coordinates 20 B; per cell: list entry 8 B, inner-hit
        # coordinates 16 B + module 2 B, flags 1 B
        b = 20 * H + 27 * D
    elif name in ("k_dfs_count", "k_dfs_emit"):
        # per neighbour entry 4 B + neighbour flags 1 B; per stored tuple: offsets 4 B,
        # hits 2 B each, cells 16 B, quality 4 B (emit)
        b = 5 * C + 4 * D // 10 + ((4 + 16 + 4) * T + 2 * Th if name == "k_dfs_emit" else 0)
    elif name == "k_fit_tuples":
        # per fit (mostly triplets at PU50): hits and module frames gathered, outputs
        # written, about 24 N + 98 B with N = 3
        b = 170 * c["n_fitted"]
    else:
        b = None
    return None if b is None else int(np.sum(b))


def profiled_traffic(kernel, path=ROOT / "profiles" / "hbm_traffic_pu50_batch100.md"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by a plain copy), or None."""
    try:
        for line in path.read_text().splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) < 4 or "`" not in cells[0]:
                continue
            name = cells[0].split("`")[1].split("<")[0].split("::")[-1]  # template arguments dropped
            if name == kernel:
                return int(float(cells[-1].replace(",", "")) * 1e6)
    except (OSError, ValueError):
        pass
    return None


def cpu_info():
    """The host the CPU baselines run on."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    return {"model": model, "host_logical_cpus": os.cpu_count(), "cpus_available_to_this_process": affinity,
            "threads_used": max(1, min(affinity, BOX_CPU_SHARE))}


def oracle():
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the checker, used here only as the CPU baseline
    return O


def run_cpu_baseline(events, frames, budget_s=12.0, label="PU50"):
    """Oracle (serial restatement of the reference stage, incl. classification)
    on host threads, one event per thread at a time like the reference's
    --numberOfStreams runner (all cores the job is given: the measurement
    boxes grant BOX_CPU_SHARE CPUs per GPU)."""
    O = oracle()
    info = cpu_info()
    threads = info["threads_used"]
    # warm-up (excluded, main.jl:148-151)
    O.make_tuples(events[0], frames, stop_after=0, debug=False)
    # size the sample to about budget_s of CPU work
    t0 = time.perf_counter()
    for h in events[:4]:
        O.make_tuples(h, frames, stop_after=0, debug=False)
    per_event = (time.perf_counter() - t0) / 4
    n_target = int(max(len(events), budget_s / max(per_event, 1e-4)))
    work = [events[i % len(events)] for i in range(n_target)]
    lock = threading.Lock()
    idx = [0]

    def worker():
        while True:
            with lock:
                i = idx[0]
                idx[0] += 1
            if i >= len(work):
                return
            O.make_tuples(work[i], frames, stop_after=0, debug=False)

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    return dict(value=round(len(work) / wall, 2), unit="events/s", cores=threads, kind="port",
                sample=f"{len(work)} {label} events (the {len(events)} benchmark events cycled), oracle = serial C++ "
                       f"restatement of the Julia stage incl. classification, {threads} threads, "
                       f"{per_event * 1e3:.1f} ms/event single-thread",
                single_thread_events_per_s=round(1.0 / per_event, 2), host=info)


def timed_cpu(fn, min_s=2.0):
    """Seconds per call of fn, repeated until min_s has passed (after one warm-up call)."""
    fn()
    n, t0 = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= min_s:
            return dt / n


def host_tasks(events, frames, n_tasks, seconds=3.0):
    """The reference's runner model on the GPU: n_tasks host threads, each with
    its own context (the per-task scratch of ca_hit_ntuplet_generator.jl:9),
    each calling the host-buffer entry point one event at a time."""
    from pixeltrack import ca
    gens = [ca.CAHitNtupletGeneratorOnGPU(frames) for _ in range(n_tasks)]
    for g in gens:
        g.make_tuples(events[0])
    done = [0] * n_tasks
    stop = [False]

    def task(i):
        k = i
        while not stop[0]:
            gens[i].make_tuples(events[k % len(events)])
            done[i] += 1
            k += n_tasks

    ths = [threading.Thread(target=task, args=(i,)) for i in range(n_tasks)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    time.sleep(seconds)
    stop[0] = True
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    for g in gens:
        g.ctx.close()
    return {"tasks": n_tasks, "events": int(sum(done)), "events_per_s": round(sum(done) / wall, 1),
            "events_per_s_per_task": round(sum(done) / wall / n_tasks, 1)}


def host_tasks_native(n_threads, max_batch, seconds=3.0):
    """The same runner model from C++ threads (cms-julia_amd/lib/pt_host_tasks,
    a child process): a context per thread (max_batch 0) or one batcher shared
    by all threads (pt_batcher_make_tuples), PU50 events of seeds 1..100, each
    thread refilling its own TrackSOA, which the library clears."""
    import subprocess
    exe = ROOT / "cms-julia_amd" / "lib" / "pt_host_tasks"
    if not exe.exists():
        return {"skipped": f"{exe} is not built"}
    r = subprocess.run([str(exe), str(n_threads), str(seconds), "100", "1", str(max_batch)], capture_output=True,
                       text=True, timeout=120)
    if r.returncode != 0:
        raise RuntimeError(f"pt_host_tasks failed ({r.returncode}): {r.stderr.strip()}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def fit_only(ctx, n_tracks=1 << 20, steps=20, cpu=False):
    """Fit-only workload: 1,048,576 pre-gathered tracks (seed 7, lengths
    3/4/5 in a 30/55/15% mix in random order, pT log-uniform 0.5-10 GeV,
    per-layer sigma_rphi 8-15 um and sigma_z 15-40 um), resident in HBM.
    tracks/s of the FP64 and mixed fits under each schedule (persistent
    after the autotuner has chosen its launch shape)."""
    from pixeltrack import _abi as A, ca, generator as G
    tr = G.fit_tracks(n_tracks, seed=7)
    fb = ca.FitBatch(ctx, tr["hits"], tr["ge"], tr["nhit"])
    out = {"workload": f"{n_tracks} tracks, lengths 3/4/5 = 30/55/15% (seed 7), pT 0.5-10 GeV",
           "tracks": n_tracks}
    try:
        for prec, pname in ((A.FIT_FP64, "fp64"), (A.FIT_MIXED, "mixed")):
            tuned, tuned_ms = fb.autotune(prec)
            out[f"{pname}_tuning"] = dict(tuned, ms=round(tuned_ms, 4))
            for sched, sname in ((A.SCHED_PLAIN, "plain"), (A.SCHED_PER_LENGTH, "per_length"),
                                 (A.SCHED_PERSISTENT, "persistent")):
                fb.run_async(prec, sched)
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(steps):
                    fb.run_async(prec, sched)
                ctx.sync()
                dt = (time.perf_counter() - t0) / steps
                out[f"{pname}_{sname}_tracks_per_s"] = round(n_tracks / dt, 1)
                out[f"{pname}_{sname}_ms"] = round(dt * 1e3, 4)
        out["mixed_vs_fp64_persistent"] = round(out["mixed_persistent_tracks_per_s"] /
                                                out["fp64_persistent_tracks_per_s"], 3)
        # vector flop rate of the persistent fits (the mixed fit's flops are partly float, partly
        # double: its fraction is quoted against the FP32 peak, an upper bound of its ceiling)
        flops = float(np.mean([FIT_FLOPS[3] if n == 3 else FIT_FLOPS[4] for n in np.asarray(tr["nhit"])]))
        out["flops_per_track"] = round(flops, 1)
        for pname, peak in (("fp64", FP64_PEAK_TFLOPS), ("mixed", FP32_PEAK_TFLOPS)):
            tf = out[f"{pname}_persistent_tracks_per_s"] * flops / 1e12
            out[f"{pname}_persistent_tflops"] = round(tf, 2)
            out[f"{pname}_persistent_frac_of_peak"] = round(tf / peak, 4)
        out["persistent_vs_per_length_mixed"] = round(out["mixed_persistent_tracks_per_s"] /
                                                      out["mixed_per_length_tracks_per_s"], 3)
        if cpu:
            O, info = oracle(), cpu_info()
            th = info["threads_used"]
            dt = timed_cpu(lambda: O.fit_bl(tr["hits"], tr["ge"], tr["nhit"], n_threads=th))
            out["cpu_baseline"] = {"value": round(n_tracks / dt, 1), "unit": "tracks/s", "cores": th, "kind": "port",
                                   "sample": f"the same {n_tracks} tracks, FP64 oracle fit (restatement of "
                                             f"launchBrokenLineKernelsOnCPU's fit), {th} threads"}
            out["fp64_persistent_vs_cpu"] = round(out["fp64_persistent_tracks_per_s"] / (n_tracks / dt), 1)
        out["riemann"] = riemann_fit_only(ctx, fb, tr, n_tracks, steps, cpu)
    finally:
        fb.free()
    return out


def riemann_fit_only(ctx, fb, tr, n_tracks, steps, cpu, cpu_sample=262144):
    """The Riemann helix fit (pt_fit_riemann_*) on the fit-only tracks: tracks/s
    under each schedule (persistent after the autotuner), against the FP64
    broken-line fit; CPU baseline: the oracle's Riemann restatement on the
    first cpu_sample tracks, all host threads."""
    from pixeltrack import _abi as A
    tuned, tuned_ms = fb.autotune(A.FIT_RIEMANN)
    out = {"tuning": dict(tuned, ms=round(tuned_ms, 4))}
    for sched, sname in ((A.SCHED_PLAIN, "plain"), (A.SCHED_PER_LENGTH, "per_length"),
                         (A.SCHED_PERSISTENT, "persistent")):
        fb.run_riemann_async(sched)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            fb.run_riemann_async(sched)
        ctx.sync()
        dt = (time.perf_counter() - t0) / steps
        out[f"{sname}_tracks_per_s"] = round(n_tracks / dt, 1)
        out[f"{sname}_ms"] = round(dt * 1e3, 4)
    fb.run_async(A.FIT_FP64, A.SCHED_PERSISTENT)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fb.run_async(A.FIT_FP64, A.SCHED_PERSISTENT)
    ctx.sync()
    bl = n_tracks / ((time.perf_counter() - t0) / steps)
    out["persistent_vs_fp64_broken_line"] = round(out["persistent_tracks_per_s"] / bl, 3)
    if cpu:
        O, info = oracle(), cpu_info()
        th = info["threads_used"]
        m = min(cpu_sample, n_tracks)
        h, g, nh = tr["hits"][:, :, :m], tr["ge"][:, :, :m], tr["nhit"][:m]
        dt = timed_cpu(lambda: O.fit_riemann(h, g, nh, n_threads=th))
        out["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "tracks/s", "cores": th, "kind": "port",
                               "sample": f"the first {m} of the same tracks, the oracle's Riemann fit "
                                         f"(restatement of riemann_fit.jl's helix_fit), {th} threads"}
        out["persistent_vs_cpu"] = round(out["persistent_tracks_per_s"] / (m / dt), 1)
    return out


def adjoint_only(ctx, frames, n_tracks=262144, steps=20, cpu=False):
    """Adjoint workload: 262,144 quadruplets (seed 11, pT log-uniform 0.5-50
    GeV) resident in HBM.  tracks/s of the plain FP64 fit against the mixed
    fit + adjoint of L = chi2_line + chi2_circle + |params - truth|^2 with
    respect to (x, y, z, sigma_rphi, sigma_z) of every hit: fused (one kernel
    recomputes the forward and runs the reverse sweep) and as an rrule
    (forward kernel, then the pullback kernel, which recomputes the forward)."""
    from pixeltrack import _abi as A, ca, generator as G
    tr = G.fit_tracks(n_tracks, seed=11, len_mix=(0.0, 1.0, 0.0), pt_min=0.5, pt_max=50.0)
    fb = ca.FitBatch(ctx, tr["hits"], tr["ge"], tr["nhit"])
    ab = ca.AdjointBatch(ctx, tr["hits"], tr["sig"], tr["det"], tr["nhit"], tr["truth"])
    out = {"workload": f"{n_tracks} quadruplets (seed 11), pT 0.5-50 GeV", "tracks": n_tracks}

    def timed(fn):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) / steps

    try:
        for prec in (A.FIT_FP64, A.FIT_MIXED):
            fb.autotune(prec)
        t_fp64 = timed(lambda: fb.run_async(A.FIT_FP64))
        t_mixed = timed(lambda: fb.run_async(A.FIT_MIXED))
        t_fused = timed(ab.loss_grad_async)
        N = np.full(n_tracks, 4)
        ab.vjp_async(np.zeros((5, n_tracks), np.float32), (2 * N - 5).astype(np.float32))
        o = ab.out

        def rrule():
            fb.run_async(A.FIT_MIXED)
            A.check(ctx.lib.pt_fit_bl_vjp_async(ctx.h, n_tracks, ab.d_hits, ab.d_sig, ab.d_det, ab.d_nhit,
                                                A.B_FIELD_INV_GEV, o["dstate"], o["dchi2"], o["grad"]),
                    "pt_fit_bl_vjp_async")
        t_rrule = timed(rrule)
        for k, t in (("fp64_fit", t_fp64), ("mixed_fit", t_mixed), ("mixed_fit_adjoint_fused", t_fused),
                     ("mixed_fit_then_pullback", t_rrule)):
            out[f"{k}_tracks_per_s"] = round(n_tracks / t, 1)
            out[f"{k}_ms"] = round(t * 1e3, 4)
        out["fused_adjoint_vs_fp64_fit"] = round(t_fp64 / t_fused, 3)
        out["rrule_vs_fp64_fit"] = round(t_fp64 / t_rrule, 3)
        if cpu:
            # FP64 oracle fit, its forward-mode (dual-number) loss gradient, and
            # the central finite-difference gradient: 2 x 5 N + 1 fits per track
            O, info = oracle(), cpu_info()
            th = info["threads_used"]
            n_s = 65536
            sub = {k: np.ascontiguousarray(v[..., :n_s]) for k, v in tr.items()}
            t_fit = timed_cpu(lambda: O.fit_bl(sub["hits"], sub["ge"], sub["nhit"], n_threads=th))
            t_grad = timed_cpu(lambda: O.fit_grad_many(sub, frames, n_threads=th))
            out["cpu_baseline"] = {
                "unit": "tracks/s", "cores": th, "kind": "port",
                "sample": f"the first {n_s} of the same tracks, {th} threads",
                "fp64_fit_tracks_per_s": round(n_s / t_fit, 1),
                "forward_mode_gradient_tracks_per_s": round(n_s / t_grad, 1),
                "central_fd_gradient_tracks_per_s": round(n_s / t_fit / (2 * 5 * 4 + 1), 1)}
            out["fused_adjoint_vs_cpu_forward_mode"] = round(out["mixed_fit_adjoint_fused_tracks_per_s"] /
                                                             (n_s / t_grad), 1)
    finally:
        fb.free()
        ab.free()
    return out


def pu200_stream(frames, n_events=200, group_size=8, n_groups=8, cpu=False, cpu_seconds=12.0, passes=5):
    """PU200-like stream: 200 events (seeds 1001..1200, 7000 helices from 200
    vertices, +5% noise), resident in HBM, run through the graph pipeline with
    `group_size` events per captured hipGraph and `n_groups` groups in flight
    on separate streams, the stage's fit in mixed precision.  events/s over
    `passes` passes of the stream (one pass, 25 replays, lasts about 15 ms),
    and per-event latency = GPU time of the replay that carried the event
    (p50 / p99)."""
    from pixeltrack import _abi as A, ca, generator as G
    events = [G.generate_event("pu200", seed=1001 + i) for i in range(n_events)]
    hits = np.array([e.n_hits for e in events])
    ctx = ca.Context(max_events=group_size * n_groups, frames=frames)
    batch = ca.DeviceBatch(ctx, events)
    try:
        ctx.set_fit_precision(A.FIT_MIXED)
        # CA tile sizes tuned on one group of the stream before the graphs are captured
        tb = ca.DeviceBatch(ctx, events[:group_size])
        tuning, _ = tb.autotune_ca()
        tb.free()
        pl = ca.Pipeline(ctx, group_size, n_groups)
        n_sub = n_events // group_size

        def run(n_pass):
            lat, pending = [], [False] * n_groups
            for j in range(n_pass * n_sub):
                g, i = j % n_groups, j % n_sub
                if pending[g]:
                    lat.append(pl.wait(g))
                pl.submit(g, batch, i * group_size)
                pending[g] = True
            for g in range(n_groups):
                if pending[g]:
                    lat.append(pl.wait(g))
            return lat

        run(10)  # warm-up: ten passes (about 150 ms), so that the GPU runs at its steady state when timed
        t0 = time.perf_counter()
        lat = run(passes)
        wall = time.perf_counter() - t0
        per_event = np.repeat(np.array(lat), group_size)
        cnt = [pl.counters(g, k) for g in range(n_groups) for k in range(group_size)]
        cpu_res = run_cpu_baseline(events, frames, budget_s=cpu_seconds, label="PU200") if cpu else None
        return {"cpu_baseline": cpu_res,
                "gpu_vs_cpu": round(passes * n_sub * group_size / wall / cpu_res["value"], 1) if cpu else None,
                "workload": f"{n_events} pu200 events (seeds 1001..{1000 + n_events}), mean {hits.mean():.0f} hits "
                            f"(max {hits.max()}), mixed-precision fit, the stream timed {passes} times over",
                "events_in_flight": group_size * n_groups, "events_per_graph": group_size, "graphs_in_flight": n_groups,
                "ca_tuning": tuning,
                "events_per_s": round(passes * n_sub * group_size / wall, 1),
                "latency_ms_p50": round(float(np.percentile(per_event, 50)), 4),
                "latency_ms_p99": round(float(np.percentile(per_event, 99)), 4),
                "max_hits_over_capacity": bool(hits.max() > A.MAX_HITS),
                # the counters of the last replay of each group (the last 32 events): PU200 events pass the
                # reference's doublet cap (458,752), which truncates exactly as max_num_of_doublets does
                "last_events_with_overflow": int(sum(1 for c in cnt if c["overflow"])),
                "last_events_overflow_bits": hex(int(np.bitwise_or.reduce([c["overflow"] for c in cnt]))),
                "last_events_doublet_cap_hit": int(sum(1 for c in cnt if c["overflow"] & 0x1))}
    finally:
        batch.free()
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=100)
    ap.add_argument("--preset", default="pu50")
    ap.add_argument("--streams", type=int, default=1, help="split each batch over this many streams")
    ap.add_argument("--inflight", type=int, default=10,
                    help="batches in flight: each batch of --events events has its own context and stream, and "
                         "step k runs on context k %% inflight, so a batch's latency-bound kernels overlap the "
                         "next batch's (109 MiB of device scratch per event: 10 contexts of 100 events hold "
                         "106 GiB)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-fit", action="store_true", help="skip the fit-only, adjoint and PU200 workloads")
    ap.add_argument("--only", default="all", choices=("all", "pu50", "fit", "adjoint", "pu200", "copy"),
                    help="run one workload alone (for profiling); 'copy' is a plain 256 MiB device copy used "
                         "to scale the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the pixel-track stage shards by event with no exchange; "
              f"--gpus {args.gpus} is not supported (run one process per GPU, each on every N-th event)",
              file=sys.stderr)
        sys.exit(2)

    from pixeltrack import ca, generator as G

    frames = G.frames()
    if args.only in ("fit", "adjoint", "pu200", "copy"):
        if args.only == "copy":
            import torch
            a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
            b = torch.empty_like(a)
            for _ in range(5):
                b.copy_(a)
            torch.cuda.synchronize()
            res = {"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "copies": 5}
        elif args.only == "pu200":
            res = pu200_stream(frames)
        else:
            c = ca.Context(max_events=1, frames=frames)
            res = fit_only(c) if args.only == "fit" else adjoint_only(c, frames)
            c.close()
        print(json.dumps({"workload": args.only, "result": res}))
        return
    if args.only == "pu50":
        args.no_fit = True
    seed0 = 1 if args.preset == "pu50" else 1001
    events = [G.generate_event(args.preset, seed=seed0 + i) for i in range(args.events)]
    print(f"[bench] generated {len(events)} {args.preset} events, mean hits "
          f"{np.mean([e.n_hits for e in events]):.0f}", file=sys.stderr, flush=True)

    # the host-buffer entry point, one event per call (PCIe included), measured
    # first: contexts created and destroyed earlier in a process were seen to slow
    # later calls of it about threefold
    host_path = gen = None
    if args.only == "all":
        gen = ca.CAHitNtupletGeneratorOnGPU(frames)
        gen.make_tuples(events[0])
        n_host = min(len(events), 50)
        t0 = time.perf_counter()
        for h in events[:n_host]:
            gen.make_tuples(h)
        w_p = time.perf_counter() - t0
        host_path = {"events_per_s": round(n_host / w_p, 1), "events": n_host,
                     "note": "pt_make_tuples: host hits in, host TrackSOA out, one event per call, synchronous"}
        host_path["tasks"] = host_tasks(events, frames, BOX_CPU_SHARE)
        host_path["native_tasks"] = host_tasks_native(BOX_CPU_SHARE, 0)
        host_path["native_tasks_shared_batcher"] = host_tasks_native(BOX_CPU_SHARE, 16)
        print(f"[bench] host buffers: {host_path}", file=sys.stderr, flush=True)

    n_fl = max(1, args.inflight)
    ctxs = [ca.Context(max_events=len(events), frames=frames) for _ in range(n_fl)]
    batches = [ca.DeviceBatch(c, events) for c in ctxs]
    ctx, batch = ctxs[0], batches[0]
    # CA tile sizes: tuned on the batch (pt_ca_autotune), used by every context
    ca_tuning, _ = batch.autotune_ca()
    for c in ctxs:
        c.set_ca_tuning(**ca_tuning)
    print(f"[bench] CA tuning {ca_tuning}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        for b in batches:
            b.run_async(out_zeroed=False)
    for c in ctxs:
        c.sync()
    counters = [ctx.counters(i) for i in range(len(events))]
    cnt = {k: np.array([c[k] for c in counters], dtype=np.int64) for k in counters[0]}

    # timed region: production mode (no timing marks); step k is one pass of the
    # stage over the batch on context k % inflight.  The output TrackSOAs are
    # reused without being cleared in between, so the stage clears state and
    # covariance itself (out_zeroed=False), as the reference's TrackSOA() does
    # inside make_tuples
    for c in ctxs:
        c.set_streams(args.streams)
        c.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batches[k % n_fl].run_async(out_zeroed=False)
    for c in ctxs:
        c.sync()
    wall = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {wall:.4f} s ({n_fl} batches in flight, {args.streams} streams "
          f"per batch)", file=sys.stderr, flush=True)
    for b, c in zip(batches[1:], ctxs[1:]):
        b.free()
        c.close()

    # per-kernel durations: HIP events recorded after every kernel on the
    # stream the kernels run on (a timed batch runs on one stream)
    ctx.enable_timing(True)
    per_kernel = {}
    n_inst = max(3, args.steps // 2)
    for _ in range(n_inst):
        batch.run_async(out_zeroed=False)
        for k, v in ctx.stage_times().items():
            per_kernel[k] = per_kernel.get(k, 0.0) + v
    ctx.enable_timing(False)

    ms_per_step = wall * 1e3 / args.steps
    value = len(events) * args.steps / wall
    avg = {k: v / n_inst for k, v in per_kernel.items()}
    dom = max(avg, key=avg.get)
    dom_ms = avg[dom]
    # achieved = the stage's algorithmic bytes (the measurement model above) over
    # the kernel's average duration; traffic = measured HBM bytes per launch of
    # the same kernel and workload (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled),
    # read from the committed summary of that profile, not measured in this run
    std_work = args.events == 100 and args.preset == "pu50"
    roof = dict(kernel=dom, bound="hbm", achieved=None, peak=HBM_PEAK_GBS, unit="GB/s", frac=None,
                traffic=profiled_traffic(dom) if std_work else None,
                traffic_source="profiles/hbm_traffic_pu50_batch100.md (rocprofv3 --pmc of this workload)"
                if std_work else None, avg_launch_ms=round(dom_ms, 5))
    if dom in STAGE_OF_KERNEL and len(STAGES[STAGE_OF_KERNEL[dom]][0]) == 1 and dom_ms > 0:
        b = stage_bytes(STAGE_OF_KERNEL[dom], cnt)
        ach = b / (dom_ms * 1e-3) / 1e9
        roof.update(achieved=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4), algorithmic_bytes=b,
                    algorithmic_model=STAGE_OF_KERNEL[dom])
    bk = kernel_model_bytes(dom, cnt)
    if bk is not None and dom_ms > 0:
        roof.update(kernel_model_bytes=bk, kernel_model_frac=round(bk / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    kernels_ms = {k: round(v, 4) for k, v in sorted(avg.items(), key=lambda kv: -kv[1])}
    # every stage of the measurement model against the HBM roofline
    stages = {}
    for st, (ks, _) in STAGES.items():
        ms = sum(avg.get(k, 0.0) for k in ks)
        if ms > 0:
            b = stage_bytes(st, cnt)
            stages[st] = {"kernels": list(ks), "ms": round(ms, 4), "bytes": b,
                          "gb_per_s": round(b / (ms * 1e-3) / 1e9, 1),
                          "frac_of_hbm_peak": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if "k_fit_tuples" in avg:
        b = fit_bytes(cnt)
        stages["fit (F9-F14)"] = {"kernels": ["k_fit_tuples"], "ms": round(avg["k_fit_tuples"], 4), "bytes": b,
                                  "gb_per_s": round(b / (avg["k_fit_tuples"] * 1e-3) / 1e9, 1)}
    tot_b = sum(v["bytes"] for v in stages.values())
    tot_ms = sum(avg.values())
    stages["whole stage, one batch on one stream"] = {
        "ms": round(tot_ms, 4), "bytes": tot_b, "gb_per_s": round(tot_b / (tot_ms * 1e-3) / 1e9, 1),
        "frac_of_hbm_peak": round(tot_b / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    stages["whole stage, timed steps"] = {
        "ms": round(wall * 1e3 / args.steps, 4), "bytes": tot_b,
        "gb_per_s": round(tot_b / (wall / args.steps) / 1e9, 1),
        "frac_of_hbm_peak": round(tot_b / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4)}

    # the same stage building the phi histogram on the device first (no
    # histogram handed in), and the host-buffer entry point (PCIe included)
    dev_hist = mixed_stage = None
    if args.only == "all":
        hb = ca.DeviceBatch(ctx, events, device_hist=True)
        hb.run_async(out_zeroed=False)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            hb.run_async(out_zeroed=False)
        ctx.sync()
        w_h = time.perf_counter() - t0
        hb.free()
        dev_hist = {"events_per_s": round(len(events) * args.steps / w_h, 1),
                    "ms_per_step": round(w_h * 1e3 / args.steps, 4), "batches_in_flight": 1}
        print(f"[bench] device histogram: {dev_hist}", file=sys.stderr, flush=True)
        # the same stage with the mixed-precision fit (one batch at a time)
        from pixeltrack import _abi as A
        ctx.set_fit_precision(A.FIT_MIXED)
        batch.run_async(out_zeroed=False)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.run_async(out_zeroed=False)
        ctx.sync()
        w_m = time.perf_counter() - t0
        ctx.set_fit_precision(A.FIT_FP64)
        mixed_stage = {"events_per_s": round(len(events) * args.steps / w_m, 1),
                       "ms_per_step": round(w_m * 1e3 / args.steps, 4), "batches_in_flight": 1}
        print(f"[bench] stage with the mixed fit: {mixed_stage}", file=sys.stderr, flush=True)

    fit = adj = stream = None
    if not args.no_fit:
        fit = fit_only(ctx, cpu=not args.no_cpu)
        print(f"[bench] fit-only: {fit}", file=sys.stderr, flush=True)
        adj = adjoint_only(ctx, frames, cpu=not args.no_cpu)
        print(f"[bench] adjoint: {adj}", file=sys.stderr, flush=True)
        stream = pu200_stream(frames, cpu=not args.no_cpu, cpu_seconds=args.cpu_seconds)
        print(f"[bench] pu200 stream: {stream}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(events, frames, budget_s=args.cpu_seconds)
        cpu["gpu_vs_cpu"] = round(value / cpu["value"], 1)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+int16 CA, f64 fit",
        "data": "synthetic",
        "config": {
            "workload": f"{args.preset}: {len(events)} events/step (seeds {seed0}..{seed0 + len(events) - 1}), "
                        f"mean {np.mean(cnt['n_hits']):.0f} hits, {np.mean(cnt['n_cells']):.0f} doublets, "
                        f"{np.mean(cnt['n_tuples']):.0f} tuples, {np.mean(cnt['n_fitted']):.0f} fits per event",
            "events_per_step": len(events),
            "batches_in_flight": n_fl,
            "ca_tuning": ca_tuning,
            "parallelism": "events batched in one launch sequence; an event's blocks on one XCD; batches in flight "
                           "on separate streams",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages": stages,
        "kernels_ms_per_step": kernels_ms,
        "kernels_sum_ms_single_stream": round(sum(avg.values()), 4),
        "streams": args.streams,
        "fit_only": fit,
        "adjoint": adj,
        "pu200_stream": stream,
        "stage_with_device_histogram": dev_hist,
        "stage_with_mixed_fit": mixed_stage,
        "host_buffers_pcie_inclusive": host_path,
    }
    print(json.dumps(line))
    batch.free()
    ctx.close()
    if gen is not None:
        gen.ctx.close()


if __name__ == "__main__":
    main()
