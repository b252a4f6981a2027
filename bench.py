#!/usr/bin/env python3
"""Benchmark of the EVX3 inter-analysis path on one MI355X.

Workload (default, `--workload plain`): a seeded 60-frame 1920x1080 (aligned
1920x1088, 8160 MBs) 8-bit 4:2:0 synthetic sequence; every frame after the
first is a P frame with up to 4 past references (4 from the fifth frame on),
quality 8, temporal MV prediction from the previous P frame, one frame per
launch.  A step is one frame: integer step search over all references,
half/quarter-pel refinement, selection, residual prediction, 8x8 DCT,
adaptive QP and quantisation of every MB.  Pictures are resident in HBM before
the timed region starts.

Prints one JSON line (the driver's contract) with the plain-path throughput,
the roofline of the analysis kernel (algorithmic integer ops per launch over
its HIP-event duration), per-frame latency percentiles, a bit-exactness check
of sampled frames against the CPU path, and the reference CPU path timed on
this host's cores over a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# The serving dispatcher keeps up to --serve-inflight launches on streams of
# their own; HIP maps streams onto this many hardware queues per process (its
# default is 4, which the dispatcher's streams would share with the others).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
sys.path.insert(0, ROOT)

METRIC = "1080p ME+DCT/quant frames/s (bit-exact), batched vs plain; p99 frame latency"

# Algorithmic integer ops per unit of work (one 32-bit-lane add, sub, abs,
# max, mul or shift each; loads not counted): an integer candidate's SAD,
# MAD, SATD and cost; a half-pel interpolation per axial / diagonal direction;
# the quarter-pel average; a sub-pel candidate's SAD + MAD; the DCT + QP +
# quantisation of a non-copy MB; the bidirectional average + SAD.
OPS_INT = 3729
OPS_HALF_AXIAL = 3456
OPS_HALF_DIAG = 8352
OPS_QAVG = 1152
OPS_SUBPEL = 1920
OPS_DCTQ = 22405
OPS_BIDIR = 2304
VALU_PEAK_TOPS = 78.6      # 256 CUs x 64 lanes x 2 (packed 16-bit) x 2.4 GHz
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def record_ops(rec) -> int:
    """Algorithmic ops of one frame's records."""
    dirs = rec["n_axial"].astype(np.int64) + rec["n_diag"]
    ops = (OPS_INT * rec["n_int"].astype(np.int64) + OPS_HALF_AXIAL * rec["n_axial"].astype(np.int64)
           + OPS_HALF_DIAG * rec["n_diag"].astype(np.int64) + OPS_QAVG * dirs + 2 * OPS_SUBPEL * dirs
           + OPS_DCTQ * (rec["flags"] & 1).astype(np.int64)
           + OPS_BIDIR * ((rec["flags"] >> 1) & 1).astype(np.int64))
    return int(ops.sum())


def frame_refs(t, max_refs=4):
    return [t - 1 - k for k in range(min(max_refs, t))]


def frame_bytes(w, h, n_refs):
    """Compulsory HBM bytes of one frame: source + references read once,
    coefficients + levels + records written once (int16 samples)."""
    n = (w // 16) * (h // 16)
    pic = w * h * 3 // 2 * 2
    return pic * (1 + n_refs) + n * (384 * 2 * 2 + 64)


def shape_table(table):
    """The autotuner's table (every candidate launch shape with its step
    time), compacted for the JSON line."""
    if not table:
        return None
    return [{"tuning": {k: v for k, v in r["tuning"].items()}, "step_ms": r["step_ms"],
             "bit_identical": r.get("bit_identical", True)} for r in table]


def cpu_model():
    """The host CPU as /proc/cpuinfo and lscpu describe it (model name,
    family/model/stepping, sockets x cores x threads), plus the cores this
    process may run on."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping", "cpu MHz",
                         "cache size") and k not in info:
                    info[k] = v.strip()
    except OSError:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket",
                             "Thread(s) per core", "CPU(s)", "L3 cache"):
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    name = info.get("lscpu Model name") or info.get("model name") or "unknown"
    desc = (f"{name} (family {info.get('cpu family', '?')} model {info.get('model', '?')} "
            f"stepping {info.get('stepping', '?')}; {info.get('lscpu Socket(s)', '?')} socket(s) x "
            f"{info.get('lscpu Core(s) per socket', '?')} cores x "
            f"{info.get('lscpu Thread(s) per core', '?')} threads, "
            f"{info.get('lscpu CPU(s)', os.cpu_count())} CPUs, L3 {info.get('lscpu L3 cache', '?')}; "
            f"{len(os.sched_getaffinity(0))} in this process's affinity)")
    return desc


def host_threads():
    """Threads for the multi-threaded CPU baseline: every core this process
    may run on, bounded by the host's own thread allotment per GPU where it
    sets one (OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        return min(n, int(cap)), f"min(affinity {n}, OMP_NUM_THREADS {cap})"
    return n, f"affinity {n}"


def cpu_baseline(frames, gpu_out, sample, quality, width, height, threads=1, mt_sample=()):
    """Time the CPU path on `sample` frames on one thread and, with the
    reference's own build, on `mt_sample` frames on `threads` threads in the
    reference's wavefront row order; compare every output with the GPU's.
    Returns the multi-threaded figure (the single-threaded one nested) when
    it was run."""
    from oracle.oracle import Oracle, Picture, Reference, reference_available
    if reference_available():
        impl, kind = Reference(), "reference"
    else:
        impl, kind = Oracle(), "port"
    pics = {}

    def pic(t):
        if t not in pics:
            pics[t] = Picture(*frames[t])
        return pics[t]

    fields = [n for n in gpu_out[sample[0]][0].dtype.names
              if n not in ("mad", "cost_bits", "mad_b", "cost_b_bits", "n_int", "n_axial",
                           "n_diag") and not n.startswith("reserved")]

    def run(frames_, n_threads):
        prev = gpu_out[frames_[0] - 1][0]
        exact, times = True, []
        for t in frames_:
            refs = [pic(r) for r in frame_refs(t)]
            t0 = time.perf_counter()
            if n_threads > 1:
                rec, co, lv = impl.analyze_frame_wpp(1, quality, pic(t), refs, prev=prev,
                                                     threads=n_threads)
            else:
                rec, co, lv = impl.analyze_frame(1, quality, pic(t), refs, prev=prev)
            times.append(time.perf_counter() - t0)
            g = gpu_out[t]
            same = all(np.array_equal(rec[f], g[0][f]) for f in fields)
            same = same and np.array_equal(co, g[1]) and np.array_equal(lv, g[2])
            exact = exact and same
            log(f"cpu {kind} frame {t} on {n_threads} thread(s): {sum(times):.1f}s so far, "
                f"bit-exact={same}")
            prev = rec
        elapsed = sum(times)
        ms = np.array(times) * 1e3
        return {
            "value": round(len(frames_) / elapsed, 4),
            "unit": "frames/s",
            "cores": n_threads,
            "kind": kind,
            "sample": (f"frames {frames_[0]}-{frames_[-1]} of the {width}x{height} sequence "
                       f"({len(frame_refs(frames_[0]))} refs each, quality {quality}), "
                       f"{n_threads} thread(s)"
                       + (" in wavefront row order" if n_threads > 1 else "")
                       + f", {elapsed:.1f} s"),
            "latency_ms": {"p50": round(float(np.percentile(ms, 50)), 1),
                           "p99": round(float(np.percentile(ms, 99)), 1)},
        }, exact

    single, exact = run(sample, 1)
    if threads <= 1 or kind != "reference" or not mt_sample:
        single["cpu_model"] = cpu_model()
        return single, exact
    multi, exact_mt = run(list(mt_sample), threads)
    multi["cpu_model"] = cpu_model()
    multi["single_thread"] = single
    return multi, exact and exact_mt


# 64 concurrent streams of uneven sizes: (width, height, count).
BATCH_MIX = [(640, 360, 16), (1280, 720, 24), (1920, 1080, 16), (2560, 1440, 8)]
BATCH_FRAMES = 5          # frames 0..3 are references, frame 4 is analysed


def _gen_stream(spec):
    from chicago_amd import synth
    w, h, seed = spec
    return synth.sequence_c1(seed, BATCH_FRAMES, w, h)


def run_batched(args, ctx):
    """64 streams, one P frame each, fused into one persistent launch per step
    and replayed from a hipGraph.  Stream i uses 1 + (i mod 4) references."""
    from concurrent.futures import ProcessPoolExecutor

    from chicago_amd import evxa, synth

    specs = []
    for w, h, n in BATCH_MIX:
        for _ in range(n):
            specs.append((w, h, synth.seed_of(3, len(specs), 0)))
    log(f"batched: generating {len(specs)} streams x {BATCH_FRAMES} frames")
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        seqs = list(ex.map(_gen_stream, specs))
    streams, descs = [], []
    for i, ((w, h, _), frames) in enumerate(zip(specs, seqs)):
        s = ctx.open_stream(synth.align16(w), synth.align16(h), BATCH_FRAMES)
        for t, f in enumerate(frames):
            s.upload(t, *f)
        nref = 1 + i % 4
        streams.append(s)
        descs.append(evxa.frame_desc(evxa.FRAME_P, args.quality, BATCH_FRAMES - 1,
                                     [BATCH_FRAMES - 2 - k for k in range(nref)],
                                     temporal_mvp=False, update_temporal=False))
    mbs = sum(s.n_mbs for s in streams)
    log(f"batched: uploaded, {mbs} MBs per step")

    tuned, modes, table = None, None, None
    if not args.no_tune:
        from chicago_amd import autotune
        tuned, table = autotune.tune(ctx, lambda: ctx.analyze_batch(streams, descs),
                                     lambda: [s.results()[0] for s in streams], autotune.BATCH,
                                     reps=2, log=log)
        modes = {k: {"ms_per_step": v["step_ms"], "tuning": v["tuning"],
                     "frames_1080p_eq_per_s": round(mbs / 8160 / (v["step_ms"] / 1e3), 2)}
                 for k, v in autotune.by_mode(table).items()}

    # Untimed eager step: outputs and work accounting.
    ctx.analyze_batch(streams, descs)
    ctx.synchronize()
    outs = [s.results() for s in streams]
    ops_step = sum(record_ops(o[0]) for o in outs)
    ctx.launch_times()
    # Eager steps, timed by HIP events around each fused launch.
    for _ in range(args.warmup):
        ctx.analyze_batch(streams, descs)
    ctx.synchronize()
    ctx.launch_times()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        ctx.analyze_batch(streams, descs)
    ctx.synchronize()
    wall_eager = time.perf_counter() - t0
    kms = ctx.launch_times()
    # Graph replays of the same step.
    gid = ctx.graph_create(streams, descs)
    for _ in range(args.warmup):
        ctx.graph_launch(gid)
    ctx.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        t1 = time.perf_counter()
        ctx.graph_launch(gid)
        ctx.synchronize()
        lat.append(time.perf_counter() - t1)
    wall_graph = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for o, s in zip(outs, streams)
               for a, b in zip(o, s.results()))
    ctx.graph_destroy(gid)
    log(f"batched: eager {wall_eager / args.batch_steps * 1e3:.2f} ms/step, "
        f"graph {wall_graph / args.batch_steps * 1e3:.2f} ms/step, replay bit-identical={same}")
    k_avg_s = float(np.mean(kms)) / 1e3
    frames_step = len(streams)
    step_s = wall_graph / args.batch_steps
    res = {
        "workload": ("batched: 64 streams (16x 640x360, 24x 1280x720, 16x 1920x1080, "
                     "8x 2560x1440), 1-4 refs, quality %d, one fused launch per step, "
                     "hipGraph replays" % args.quality),
        "frames_per_s": round(frames_step / step_s, 2),
        "frames_1080p_eq_per_s": round(mbs / 8160 / step_s, 2),
        "ms_per_step": round(step_s * 1e3, 3),
        "ms_per_step_eager": round(wall_eager / args.batch_steps * 1e3, 3),
        "mbs_per_step": mbs,
        "kernel_ms_avg": round(k_avg_s * 1e3, 3),
        "achieved_tops": round(ops_step / k_avg_s / 1e12, 3),
        "frac_of_valu_peak": round(ops_step / k_avg_s / 1e12 / VALU_PEAK_TOPS, 5),
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                            "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "replays_bit_identical": same,
        "launch_shape": tuned,
        "launch_shapes": shape_table(table),
        "modes": modes,
    }
    return res, streams, descs, seqs, outs


def batched_cpu_check(seqs, descs, outs, quality, picks):
    """Reference CPU path on a few of the batched streams; bit-exact check."""
    from oracle.oracle import Oracle, Picture, Reference, reference_available
    impl = Reference() if reference_available() else Oracle()
    ok = True
    for i in picks:
        frames = seqs[i]
        d = descs[i]
        refs = [Picture(*frames[d.ref_slots[k]]) for k in range(d.n_refs)]
        rec, co, lv = impl.analyze_frame(1, quality, Picture(*frames[d.src_slot]), refs)
        g = outs[i]
        fields = [n for n in rec.dtype.names if n not in (
            "mad", "cost_bits", "mad_b", "cost_b_bits", "n_int", "n_axial", "n_diag")
            and not n.startswith("reserved")]
        same = all(np.array_equal(rec[f], g[0][f]) for f in fields)
        same = same and np.array_equal(co, g[1]) and np.array_equal(lv, g[2])
        log(f"batched stream {i}: CPU bit-exact={same}")
        ok = ok and same
    return ok


def run_plain(args):
    from chicago_amd import evxa, synth

    w, h = 1920, 1080
    aw, ah = synth.align16(w), synth.align16(h)
    log(f"generating {args.frames} frames {w}x{h}")
    frames = synth.sequence_c1(synth.seed_of(1, 0, 0), args.frames, w, h)
    ctx = evxa.Context(0)
    s = ctx.open_stream(aw, ah, args.frames)
    for t, f in enumerate(frames):
        s.upload(t, *f)
    log("uploaded")
    q = args.quality
    seq = list(range(1, args.frames))

    def step(t):
        if t == 1:
            s.reset_temporal()
        s.analyze_p(q, t, frame_refs(t))

    # Launch shape: the autotuner times frames 4-6 (4 references; temporal
    # table reset first, so every shape sees the same inputs, then chained as
    # in the timed run) under each candidate.
    tuned, modes, table = None, None, None
    if not args.no_tune:
        from chicago_amd import autotune

        def tune_step():
            s.reset_temporal()
            for t in (4, 5, 6):
                s.analyze_p(q, t, frame_refs(t))
        tuned, table = autotune.tune(ctx, tune_step, lambda: s.results(), autotune.SINGLE,
                                     log=log)
        modes = {k: {"ms_per_frame": round(v["step_ms"] / 3, 4), "tuning": v["tuning"]}
                 for k, v in autotune.by_mode(table).items()}

    # Untimed pass: work accounting, outputs for the CPU comparison and
    # per-frame latency (enqueue -> device idle).
    keep = set(range(args.cpu_start - 1,
                     args.cpu_start + max(args.cpu_frames, args.cpu_mt_frames)))
    ops, outs, lat = {}, {}, []
    for t in seq:
        t0 = time.perf_counter()
        step(t)
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
        rec, co, lv = s.results(coeffs=t in keep, levels=t in keep)
        ops[t] = record_ops(rec)
        if t in keep:
            outs[t] = (rec, co, lv)
    ctx.launch_times()
    log(f"accounting pass done: {np.mean(list(ops.values())) / 1e9:.2f} Gops/frame")

    order = [seq[k % len(seq)] for k in range(args.warmup + args.steps)]
    for t in order[:args.warmup]:
        step(t)
    ctx.synchronize()
    ctx.launch_times()
    timed = order[args.warmup:]
    t0 = time.perf_counter()
    for t in timed:
        step(t)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kms = ctx.launch_times()
    assert len(kms) == len(timed), (len(kms), len(timed))
    log(f"timed {len(timed)} frames in {wall:.3f}s")
    sweep_mode = ctx.tuning()[0]["mode"] == evxa.MODE_SWEEP
    sweeps = None
    if sweep_mode:
        # The sweeps of one frame of the run (frame 6 after 5, temporal table).
        step(5)
        step(6)
        listed, searched, done_us = ctx.sweep_stats(times=True)
        ctx.launch_times()
        sweeps = {"frame": 6, "mbs_listed": listed, "mbs_searched": searched,
                  "done_us": done_us, "searches_per_mb": round(
                      sum(searched) / ((aw // 16) * (ah // 16)), 3)}

    ops_launch = float(np.mean([ops[t] for t in timed]))
    k_avg_s = float(np.mean(kms)) / 1e3
    achieved = ops_launch / k_avg_s / 1e12
    nrefs = np.mean([len(frame_refs(t)) for t in timed])
    bytes_launch = float(np.mean([frame_bytes(aw, ah, len(frame_refs(t))) for t in timed]))

    result = {
        "metric": METRIC,
        "value": round(len(timed) / wall, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": len(timed),
        "warmup": args.warmup,
        "ms_per_step": round(wall / len(timed) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic",
        "config": {
            "workload": (f"plain: {w}x{h} 4:2:0 P frames, up to 4 refs ({nrefs:.2f} mean), "
                         f"quality {q}, quarter-pel, one frame per launch"),
            "frames": args.frames,
            "mbs_per_frame": (aw // 16) * (ah // 16),
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 4),
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s",
            "frac": round(achieved / VALU_PEAK_TOPS, 5),
            "traffic": measured_traffic(sweep_mode),
            "kernel": "sweep_kernel + final_kernel" if sweep_mode else "analyze_kernel",
            "kernel_ms_avg": round(k_avg_s * 1e3, 4),
            "algorithmic_ops_per_launch": int(ops_launch),
            "algorithmic_bytes_per_launch": int(bytes_launch),
            "hbm_gbs_algorithmic": round(bytes_launch / k_avg_s / 1e9, 2),
        },
        "latency_ms": {
            "p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "what": "enqueue of one frame to device idle, sequential frames",
        },
        "launch_shape": tuned,
        "launch_shapes": shape_table(table),
        "modes": modes,
        "sweeps": sweeps,
    }
    if not args.no_cpu:
        sample = list(range(args.cpu_start, args.cpu_start + args.cpu_frames))
        mt_sample = [t for t in range(args.cpu_start, args.cpu_start + args.cpu_mt_frames)
                     if t in outs]
        cb, exact = cpu_baseline(frames, outs, sample, q, w, h, threads=args.cpu_threads,
                                 mt_sample=mt_sample)
        cb["threads_rule"] = args.cpu_threads_rule
        result["cpu_baseline"] = cb
        result["bit_exact_vs_cpu"] = exact
        result["gpu_vs_cpu"] = round(result["value"] / cb["value"], 1)
    args.plain_tuning = ctx.tuning()[0]
    s.close()
    ctx.close()
    return result


def run_pair(args):
    """The CPU-runnable case: one seeded 1280x720 P-frame pair (texture
    reference; each block moved by a random integer MV in [-16, 16]^2 plus a
    random quarter-pel state, Gaussian noise sigma 2), 1 reference, quality 8,
    no temporal table.  GPU time per frame in both modes (best launch shape
    of each), the reference build's time on 1 and on N host threads, and
    whether every output is identical."""
    from chicago_amd import autotune, evxa, synth

    w, h = 1280, 720
    ref_pic, cur_pic = synth.pair_c0(synth.seed_of(0, 0, 0), w, h)
    q = args.quality
    ctx = evxa.Context(0)
    s = ctx.open_stream(w, h, 2)
    s.upload(0, *ref_pic)
    s.upload(1, *cur_pic)
    d = evxa.frame_desc(evxa.FRAME_P, q, 1, [0], temporal_mvp=False, update_temporal=False)

    def step():
        s.analyze(d)

    _, table = autotune.tune(ctx, step, lambda: s.results(), autotune.SINGLE, reps=5, log=log)
    modes = {}
    n = 200
    for name, best in autotune.by_mode(table).items():
        ctx.set_tuning(**best["tuning"])
        for _ in range(5):
            step()
        ctx.synchronize()
        ctx.launch_times()
        t0 = time.perf_counter()
        for _ in range(n):
            step()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        kms = ctx.launch_times()
        lat = []
        for _ in range(20):
            t1 = time.perf_counter()
            step()
            ctx.synchronize()
            lat.append(time.perf_counter() - t1)
        modes[name] = {"ms_per_frame": round(wall / n * 1e3, 4),
                       "kernel_ms_avg": round(float(np.mean(kms)), 4),
                       "latency_ms_p50": round(float(np.median(lat)) * 1e3, 4),
                       "tuning": best["tuning"]}
    best_mode = min(modes, key=lambda k: modes[k]["ms_per_frame"])
    ctx.set_tuning(**modes[best_mode]["tuning"])
    step()
    got = s.results()
    res = {
        "workload": ("pair: one 1280x720 4:2:0 P frame, 1 reference, quality 8, quarter-pel, "
                     "no temporal table, one frame per launch"),
        "frames_per_s": round(1e3 / modes[best_mode]["ms_per_frame"], 2),
        "ms_per_frame": modes[best_mode]["ms_per_frame"],
        "mode": best_mode,
        "modes": modes,
    }
    s.close()
    ctx.close()
    if args.no_cpu:
        return res
    from oracle.oracle import Oracle, Picture, Reference, reference_available
    impl, kind = (Reference(), "reference") if reference_available() else (Oracle(), "port")
    R, S = Picture(*ref_pic), Picture(*cur_pic)
    fields = [f for f in got[0].dtype.names if f not in (
        "mad", "cost_bits", "mad_b", "cost_b_bits", "n_int", "n_axial", "n_diag")
        and not f.startswith("reserved")]
    exact = True
    cpu = {}
    for threads in ([1, args.cpu_threads] if kind == "reference" and args.cpu_threads > 1
                    else [1]):
        times = []
        while sum(times) < 3.0 and len(times) < 20:
            t0 = time.perf_counter()
            if threads > 1:
                want = impl.analyze_frame_wpp(1, q, S, [R], threads=threads)
            else:
                want = impl.analyze_frame(1, q, S, [R])
            times.append(time.perf_counter() - t0)
            same = (all(np.array_equal(want[0][f], got[0][f]) for f in fields)
                    and np.array_equal(want[1], got[1]) and np.array_equal(want[2], got[2]))
            exact = exact and same
        log(f"pair: cpu {kind} on {threads} thread(s): {np.median(times) * 1e3:.1f} ms/frame, "
            f"bit-exact={exact}")
        cpu[str(threads)] = {"ms_per_frame": round(float(np.median(times)) * 1e3, 2),
                             "frames_per_s": round(1.0 / float(np.median(times)), 3),
                             "runs": len(times)}
    res["cpu"] = {"kind": kind, "threads": cpu, "threads_rule": args.cpu_threads_rule}
    res["gpu_vs_cpu_1_thread"] = round(cpu["1"]["ms_per_frame"] / res["ms_per_frame"], 1)
    res["bit_exact_vs_cpu"] = exact
    return res


def run_gop(args):
    """Hierarchical-B workload: a seeded 33-frame 3840x2160 sequence (zoom
    0.5%/frame about the centre plus +-32 px local motion on 30% of blocks),
    GOP 8 with I/P anchors and a 3-level B pyramid, quality 12 (reference Bs
    +1, non-reference Bs +2).  One frame per launch in decode order; B frames
    run forward, backward and bidirectional prediction."""
    from chicago_amd import evxa, gop, synth

    w, h = 3840, 2160
    n = args.gop_frames
    log(f"gop: generating {n} frames {w}x{h}")
    frames = synth.sequence_c2(synth.seed_of(2, 0, 0), n, w, h)
    roles = gop.gop8_schedule(n, 12)
    ctx = evxa.Context(0)
    s = ctx.open_stream(synth.align16(w), synth.align16(h), n)
    for t, f in enumerate(frames):
        s.upload(t, *f)
    log("gop: uploaded")
    work = [r for r in roles if r.kind != "I"]

    def step(r):
        if r.kind == "P":
            if not r.temporal_mvp:
                s.reset_temporal()
            s.analyze_p(r.quality, r.index, r.refs)
        else:
            s.analyze_b(r.quality, r.index, r.refs[0], r.refs[1], target_ids=[0, 1])

    tuned, modes, table = None, None, None
    if not args.no_tune:
        from chicago_amd import autotune
        rb = next(r for r in work if r.kind == "B")
        tuned, table = autotune.tune(ctx, lambda: step(rb), lambda: s.results(), autotune.SINGLE,
                                     reps=2, log=log)
        modes = {k: {"ms_per_b_frame": v["step_ms"], "tuning": v["tuning"]}
                 for k, v in autotune.by_mode(table).items()}

    # Untimed pass: per-frame latency, work accounting, outputs of a sample.
    sample = {work[0].index, next(r.index for r in work if r.kind == "B")}
    ops, lat, outs, ptables = 0, [], {}, {}
    for r in work:
        t0 = time.perf_counter()
        step(r)
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
        rec, co, lv = s.results(coeffs=r.index in sample, levels=r.index in sample)
        ops += record_ops(rec)
        if r.kind == "P":
            ptables[r.index] = rec
        if r.index in sample:
            outs[r.index] = (rec, co, lv)
    ctx.launch_times()
    for _ in range(args.warmup and 1):
        for r in work:
            step(r)
    ctx.synchronize()
    ctx.launch_times()
    t0 = time.perf_counter()
    for r in work:
        step(r)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kms = ctx.launch_times()
    res = {
        "workload": (f"gop: {w}x{h} GOP 8 (I/P anchors, 3-level hierarchical B), "
                     f"{len(work)} analysed frames in decode order, quality 12, one frame per launch"),
        "frames_per_s": round(len(work) / wall, 3),
        "frames_1080p_eq_per_s": round(len(work) * (s.n_mbs / 8160) / wall, 3),
        "ms_per_frame": round(wall / len(work) * 1e3, 3),
        "kernel_ms_avg": round(float(np.mean(kms)), 3),
        "achieved_tops": round(ops / (float(np.sum(kms)) / 1e3) / 1e12, 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "launch_shape": tuned,
        "launch_shapes": shape_table(table),
        "modes": modes,
    }
    if not args.no_cpu:
        from oracle.oracle import Oracle, Picture, Reference, reference_available
        impl = Reference() if reference_available() else Oracle()
        exact = True
        for r in work:
            if r.index not in outs:
                continue
            cur = Picture(*frames[r.index])
            if r.kind == "P":
                prev = ptables.get(r.refs[0]) if r.temporal_mvp else None
                want = impl.analyze_frame(1, r.quality, cur, [Picture(*frames[r.refs[0]])], prev=prev)
            else:
                want = impl.analyze_frame(2, r.quality, cur, [Picture(*frames[k]) for k in r.refs],
                                          target_ids=[0, 1])
            got = outs[r.index]
            fields = [f for f in want[0].dtype.names if f not in (
                "mad", "cost_bits", "mad_b", "cost_b_bits", "n_int", "n_axial", "n_diag")
                and not f.startswith("reserved")]
            same = all(np.array_equal(want[0][f], got[0][f]) for f in fields)
            same = same and np.array_equal(want[1], got[1]) and np.array_equal(want[2], got[2])
            log(f"gop frame {r.index} ({r.kind}): CPU bit-exact={same}")
            exact = exact and same
        res["bit_exact_vs_cpu"] = exact
    s.close()
    ctx.close()
    return res


def run_serve(args):
    """Serving: a real-time request loop over live streams at 60 fps (2 ms
    exponential arrival jitter; 50% 1280x720, 35% 1920x1080, 15% 2560x1440;
    exponential lifetimes of mean 20 s), P frames with one reference and the
    previous frame's MV table, continuous batching through the native
    dispatcher.  Streams are admitted by HBM budget and by a compute capacity
    measured first: a burst of frames of the mix through the same dispatcher
    gives the sustainable MB/s, and admission keeps the offered MB/s within
    `--serve-util` of it.  Run in both modes (MB wavefront, sweeps); the one
    with the lower p99 latency at the main load is reported first."""
    from chicago_amd import autotune, evxa, serve

    runs = {}
    for name, mode in (("sweep", autotune.SWEEP), ("wavefront", autotune.WAVEFRONT)):
        ctx = evxa.Context(0)
        shapes = [c for c in autotune.SERVE if c["mode"] == mode]
        table = None
        if args.no_tune:
            ctx.set_tuning(**shapes[0])
        cap = full = None
        if args.serve_util > 0:
            if args.serve_capacity > 0:
                full = args.serve_capacity
            else:
                full, table = calibrate_serving(ctx, args, None if args.no_tune else shapes)
            cap = full * args.serve_util
        elif not args.no_tune:
            _, table = calibrate_serving(ctx, args, shapes)
            log(f"serve ({name}): admission capacity {cap / 1e6:.2f} M MB/s")
        res = serve.run(ctx, seed=args.serve_seed, n_live=args.serve_streams,
                        duration=args.serve_seconds, quality=args.quality,
                        max_inflight=args.serve_inflight, max_frames=args.serve_max_frames,
                        capacity_mbps=cap, log=log)
        # The same schedule at lower admitted loads: latency against throughput.
        if cap is not None and args.serve_curve:
            res["curve"] = []
            for u in (float(x) for x in args.serve_curve.split(",")):
                r = serve.run(ctx, seed=args.serve_seed, n_live=args.serve_streams,
                              duration=args.serve_seconds, quality=args.quality,
                              max_inflight=args.serve_inflight, max_frames=args.serve_max_frames,
                              capacity_mbps=full * u, log=log)
                res["curve"].append({"util": u, "frames_per_s": r["frames_per_s"],
                                     "frames_1080p_eq_per_s": r["frames_1080p_eq_per_s"],
                                     "admitted_lives": r["admitted_lives"],
                                     "latency_ms": r["latency_ms"],
                                     "latency_ms_by_height": r["latency_ms_by_height"]})
                log(f"serve ({name}) curve util {u}: {r['frames_per_s']} frames/s, p99 "
                    f"{r['latency_ms']['p99'] if r['latency_ms'] else None} ms")
        res["workload"] = (f"serve: {args.serve_streams} live streams offered at 60 fps "
                           f"(50% 720p / 35% 1080p / 15% 1440p, mean life 20 s, 2 ms jitter), "
                           f"{args.serve_seconds:g} s schedule, 1 ref, quality {args.quality}, "
                           f"{args.serve_inflight} launches in flight")
        res["admission_capacity_mb_per_s"] = cap
        res["launch_shape"] = ctx.tuning()[0]
        res["launch_shapes"] = table
        res["mode"] = name
        runs[name] = res
        ctx.close()

    def p99(r):
        return r["latency_ms"]["p99"] if r.get("latency_ms") else float("inf")
    best = min(runs.values(), key=p99)
    other = [r for r in runs.values() if r is not best][0]
    best["other_mode"] = {k: other.get(k) for k in (
        "mode", "frames_per_s", "frames_1080p_eq_per_s", "admitted_lives", "latency_ms",
        "admission_capacity_mb_per_s", "curve", "launch_shape", "launch_shapes")}
    return best


def run_recon(args):
    """Reconstruction: a 1920x1080 P frame with 4 references is analysed,
    then rebuilt into a spare picture slot (prediction plus inverse-quantised,
    inverse-transformed levels) repeatedly; the kernel's device time and its
    algorithmic HBM bytes give the rate and the roofline.  The result is
    checked bit for bit against the oracle's decode-side driver."""
    from chicago_amd import evxa, synth

    w, h = 1920, 1080
    aw, ah = synth.align16(w), synth.align16(h)
    frames = synth.sequence_c1(synth.seed_of(1, 0, 0), 6, w, h)
    ctx = evxa.Context(0)
    s = ctx.open_stream(aw, ah, 7)
    for t, f in enumerate(frames):
        s.upload(t, *f)
    s.analyze_p(args.quality, 5, frame_refs(5))
    rec, _, lv = s.results()
    ctx.launch_times()
    reps = 50
    for _ in range(5):
        s.reconstruct(6)
    ctx.synchronize()
    ctx.launch_times()
    for _ in range(reps):
        s.reconstruct(6)
    ctx.synchronize()
    ms = np.array(ctx.launch_times())
    k_s = float(np.mean(ms)) / 1e3
    n = len(rec)
    coded = int(((rec["block_type"] & 4) == 0).sum())
    bidir = int((rec["prediction_mode"] == 2).sum())
    nbytes = n * (64 + 768 + 768) + (coded + bidir) * 768
    res = {
        "workload": "reconstruct: one 1920x1080 P frame (4 refs, quality 8) into a spare slot",
        "kernel": "recon_kernel",
        "kernel_us_avg": round(k_s * 1e6, 2),
        "mbs_per_s": round(n / k_s, 0),
        "frames_per_s": round(1.0 / k_s, 1),
        "roofline": {"bound": "hbm", "achieved": round(nbytes / k_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(nbytes / k_s / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": nbytes},
        "transformed_mbs": coded,
    }
    got = s.download(6)
    # Deblocking of that reconstruction with the frame's own records
    # (evxa_deblock: a device copy of the picture, then deblock_kernel from
    # the copy back into the slot, then the apron).
    s.deblock(6)
    got_db = s.download(6)
    for _ in range(5):
        s.deblock(6)
    ctx.synchronize()
    ctx.launch_times()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.deblock(6)
    ctx.synchronize()
    call_s = (time.perf_counter() - t0) / reps
    d_s = float(np.mean(ctx.launch_times())) / 1e3
    # Reconstruction then deblocking, as two calls and as the fused call.
    t0 = time.perf_counter()
    for _ in range(reps):
        s.reconstruct(6)
        s.deblock(6)
    ctx.synchronize()
    pair_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        s.reconstruct_deblocked(6)
    ctx.synchronize()
    fused_s = (time.perf_counter() - t0) / reps
    ctx.launch_times()
    got_fused = s.download(6)
    samples = aw * ah * 3 // 2
    dbytes = samples * 4 + n * 64
    res["deblock"] = {
        "workload": "deblock: the reconstructed 1920x1080 frame, its own records, in place",
        "kernel": "deblock_kernel",
        "kernel_us_avg": round(d_s * 1e6, 2),
        "call_us": round(call_s * 1e6, 2),
        "reconstruct_then_deblock_us": round(pair_s * 1e6, 2),
        "reconstruct_deblocked_us": round(fused_s * 1e6, 2),
        "frames_per_s": round(1.0 / d_s, 1),
        "roofline": {"bound": "hbm", "achieved": round(dbytes / d_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dbytes / d_s / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": dbytes},
    }
    # SAO of the deblocked reconstruction against the frame's source (slot 5).
    s.reconstruct_deblocked(6)
    table = s.sao(6, 5, args.quality)
    got_sao = s.download(6)
    ctx.synchronize()
    ctx.launch_times()
    sreps = 20
    t0 = time.perf_counter()
    for _ in range(sreps):
        s.reconstruct_deblocked(6)
        s.sao(6, 5, args.quality, table=False)
    ctx.synchronize()
    chain_s = (time.perf_counter() - t0) / sreps
    ms = np.array(ctx.launch_times()).reshape(sreps, 4)     # recon, deblock, decide, apply
    dec_s, app_s = float(np.mean(ms[:, 2])) / 1e3, float(np.mean(ms[:, 3])) / 1e3
    res["sao"] = {
        "workload": "SAO: the deblocked 1920x1080 reconstruction against its source frame",
        "kernels": "decisions: sao_own_kernel, sao_cost_kernel, sao_merge_kernel; filter: sao_apply_kernel",
        "decide_us_avg": round(dec_s * 1e6, 2),
        "apply_us_avg": round(app_s * 1e6, 2),
        "reconstruct_deblock_sao_call_us": round(chain_s * 1e6, 2),
        "luma_types_off_edge_band": np.bincount(table["luma_type"], minlength=3).tolist(),
        "merges_left_up": [int(table["merge_left"].sum()), int(table["merge_above"].sum())],
        "roofline": {"bound": "hbm", "kernel": "sao_apply_kernel",
                     "achieved": round(samples * 4 / app_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(samples * 4 / app_s / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": samples * 4},
    }
    if not args.no_cpu:
        from oracle.oracle import Oracle, Picture
        o = Oracle()
        pics = [Picture(*f) for f in frames]
        want = o.reconstruct_frame(1, args.quality, pics[5], [pics[r] for r in frame_refs(5)],
                                   rec, lv)
        res["bit_exact_vs_cpu"] = all(np.array_equal(a, b) for a, b in zip(got, want))
        want_db = o.deblock_frame(rec, *want)
        res["deblock"]["bit_exact_vs_cpu"] = all(
            np.array_equal(a, b) and np.array_equal(c, b)
            for a, b, c in zip(got_db, want_db, got_fused))
        want_t, *want_sao = o.sao_encode_frame(frames[5], want_db, o.lam(args.quality))
        res["sao"]["bit_exact_vs_cpu"] = bool(
            all(np.array_equal(table[f], want_t[f]) for f in want_t.dtype.names
                if "reserved" not in f) and
            all(np.array_equal(a, b) for a, b in zip(got_sao, want_sao)))
    s.close()
    ctx.close()
    return res


def run_loop(args):
    """Closed-loop P coding on the device: each 1920x1080 frame is analysed
    against the reconstructions of up to 4 previous frames (temporal MV
    predictor from the previous frame), then reconstructed from its own
    levels, deblocked and SAO'd against its source; that picture is the
    next frames' reference, so the decoded picture buffer stays in HBM.
    Sources are resident before the timed region; each frame ends with the
    SAO call's synchronisation.  Frames 1, 2 and the last timed frame are
    checked against a CPU chain run after the timed region."""
    from chicago_amd import evxa, synth

    w, h = 1920, 1080
    aw, ah = synth.align16(w), synth.align16(h)
    k = args.loop_frames
    frames = synth.sequence_c1(synth.seed_of(1, 0, 0), k + 1, w, h)
    ctx = evxa.Context(0)
    if getattr(args, "plain_tuning", None):
        ctx.set_tuning(**args.plain_tuning)   # the plain workload's tuned shape
    ring = 5                                  # 4 references + the frame being coded
    s = ctx.open_stream(aw, ah, ring + k + 1)
    src_slot = lambda t: ring + t
    rec_slot = lambda t: t % ring
    for t, f in enumerate(frames):
        s.upload(src_slot(t), *f)
    s.upload(rec_slot(0), *frames[0])        # frame 0 stands in for a coded intra frame

    def code(t):
        refs = [rec_slot(t - 1 - j) for j in range(min(4, t))]
        s.analyze_p(args.quality, src_slot(t), refs, temporal_mvp=t > 1)
        s.reconstruct_deblocked(rec_slot(t))
        s.sao(rec_slot(t), src_slot(t), args.quality, table=False)

    res = {"workload": f"closed loop: {k} 1920x1080 P frames, up to 4 reconstructed references, "
                       f"quality {args.quality}; analysis, reconstruction, deblocking, SAO per frame"}
    # Frames 1 and 2 run before the timed region and the last frame ends it;
    # their records, levels and decoded pictures are compared with a CPU
    # chain afterwards.
    device_out = {}
    for t in (1, 2):
        code(t)
        rec, _, lv = s.results()
        device_out[t] = (rec, lv, s.download(rec_slot(t)))
    start = 3
    for t in range(start, min(start + 2, k + 1)):
        code(t)                               # warm-up frames
    ctx.synchronize()
    ctx.launch_times()
    t0 = time.perf_counter()
    n = 0
    for t in range(min(start + 2, k + 1), k + 1):
        code(t)
        n += 1
    ctx.synchronize()
    dt = time.perf_counter() - t0
    if n:
        rec, _, lv = s.results()
        device_out[k] = (rec, lv, s.download(rec_slot(k)))
    res["frames_timed"] = n
    res["ms_per_frame"] = round(dt / max(n, 1) * 1e3, 3)
    res["frames_per_s"] = round(n / dt, 2) if n else None
    # Device time per frame from the events around each kernel group:
    # analysis, reconstruction, deblocking, SAO decisions, SAO filter.
    ms = np.array(ctx.launch_times())
    if n and ms.size == 5 * n:
        per = ms.reshape(n, 5).mean(axis=0)
        res["device_ms_per_frame"] = {name: round(float(v), 3) for name, v in
                                      zip(("analysis", "reconstruct", "deblock", "sao_decide",
                                           "sao_apply"), per)}
    s.close()
    ctx.close()
    if not args.no_cpu:
        res["bit_exact_vs_cpu"], res["checked_frames"] = loop_cpu_chain(args, frames, device_out)
    return res


def loop_cpu_chain(args, frames, device_out):
    """The closed loop on the CPU: each frame analysed against the CPU's own
    decoded references, reconstructed, deblocked and SAO'd -- by the
    reference build (its analysis on the host's threads) where it is built,
    else by the oracle.  Compares the frames in `device_out` (records,
    levels and decoded picture)."""
    from oracle.oracle import Oracle, Picture, Reference, reference_available
    o = Oracle()
    ana = Reference() if reference_available() else None
    q = args.quality
    recons = [Picture(*frames[0])]
    prev = None
    checked, t0 = [], time.perf_counter()
    fields = [f for f in device_out[1][0].dtype.names if not f.startswith("reserved")
              and (ana is None or f not in ("mad", "cost_bits", "mad_b", "cost_b_bits", "n_int",
                                             "n_axial", "n_diag"))]
    for t in range(1, max(device_out) + 1):
        refs = [recons[t - 1 - j] for j in range(min(4, t))]
        cur = Picture(*frames[t])
        p = prev if t > 1 else None
        if ana is not None:
            orec, _, olv = ana.analyze_frame_wpp(1, q, cur, refs, prev=p, threads=args.cpu_threads)
        else:
            orec, _, olv = o.analyze_frame(1, q, cur, refs, prev=p)
        impl = ana or o
        filt = impl.deblock_frame(orec, *impl.reconstruct_frame(1, q, cur, refs, orec, olv))
        _, *want = impl.sao_encode_frame(frames[t], filt, impl.lam(q))
        if t in device_out:
            rec, lv, pic = device_out[t]
            same = (all(np.array_equal(rec[f], orec[f]) for f in fields)
                    and np.array_equal(lv, olv)
                    and all(np.array_equal(a, b) for a, b in zip(pic, want)))
            log(f"loop frame {t}: CPU chain bit-exact={same} ({time.perf_counter() - t0:.1f} s)")
            checked.append((t, bool(same)))
        recons.append(Picture(*want))
        prev = orec
    return all(c for _, c in checked), [t for t, _ in checked]


ENCODE_FIELDS = ("block_type", "prediction_mode", "prediction_target", "prediction_target_b",
                 "motion_x", "motion_y", "motion_x_b", "motion_y_b", "sp_pred", "sp_amount",
                 "sp_index", "sp_pred_b", "sp_amount_b", "sp_index_b", "q_index", "variance",
                 "intra_mode")


def run_encode(args):
    """The reference encoder's frame loop on the device at 1920x1080: its own
    GOP (an I frame, then per GOP a P anchor, a reference B and two
    non-reference Bs, in decode order, with its picture-buffer slots and B
    qualities), each frame coded by the in-loop encode (mode decision with
    the intra-frame block search, RDO for P, trellis, reconstruction into
    its slot over the slot's previous picture), then deblocked and SAO'd in
    place, as engine_encode_frame does (encode.cpp:945-983).  The sequence
    runs once untimed (its first frames and its last are checked against
    the reference build's own chain on the host) and once timed."""
    from chicago_amd import evxa, gop, synth

    w, h = 1920, 1080
    aw, ah = synth.align16(w), synth.align16(h)
    n = args.encode_frames
    q = args.quality
    frames = synth.sequence_c1(synth.seed_of(1, 0, 0), n, w, h)
    roles = gop.reference_schedule(n, q)
    ctx = evxa.Context(0)
    n_dpb = 4
    s = ctx.open_stream(aw, ah, n_dpb + n)
    src_slot = lambda d: n_dpb + d
    for d, f in enumerate(frames):
        s.upload(src_slot(d), *f)
    kind = {"I": evxa.FRAME_I, "P": evxa.FRAME_P, "B": evxa.FRAME_B}

    def code(r):
        refs = [] if r.kind == "I" else r.ref_slots
        s.encode_frame(kind[r.kind], r.quality, src_slot(r.index), r.slot, refs,
                       target_ids=refs, temporal_mvp=r.kind == "P" and r.temporal_from >= 0)
        s.deblock(r.slot)
        s.sao(r.slot, src_slot(r.index), r.quality, table=False)

    res = {"workload": f"encode: the reference encoder's GOP (I, then P + reference B + 2 Bs per "
                       f"GOP) over {n} 1920x1080 frames, quality {q}; per frame the in-loop encode "
                       f"(encode_slice), deblocking and SAO on the device",
           "frames": n, "decode_order": "".join(r.kind for r in roles)}
    check = sorted({0, 1, 2, 3, 4, len(roles) - 1})
    device_out = {}
    for k, r in enumerate(roles):
        code(r)
        if k in check:
            rec, _, lv = s.results()
            device_out[k] = (rec, lv, s.download(r.slot))
    ctx.synchronize()
    ctx.launch_times()
    t0 = time.perf_counter()
    for r in roles:
        code(r)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    res["ms_per_frame"] = round(dt / len(roles) * 1e3, 3)
    res["frames_per_s"] = round(len(roles) / dt, 2)
    ms = np.array(ctx.launch_times())
    if ms.size == 4 * len(roles):
        per = ms.reshape(len(roles), 4)
        res["device_ms_per_frame"] = {name: round(float(v), 3) for name, v in
                                      zip(("encode", "deblock", "sao_decide", "sao_apply"),
                                          per.mean(axis=0))}
        by_kind = {}
        for kk in "IPB":
            rows = [i for i, r in enumerate(roles) if r.kind == kk]
            if rows:
                by_kind[kk] = round(float(per[rows, 0].mean()), 3)
        res["encode_ms_by_type"] = by_kind
        # The MB wavefront's length: MB (bx, by) runs after (bx + 3, by - 1).
        wb, hb = aw // 16, ah // 16
        steps = wb + 4 * (hb - 1)
        res["wavefront_steps"] = steps
        res["us_per_wavefront_step"] = round(float(per[:, 0].mean()) * 1e3 / steps, 2)
    s.close()
    ctx.close()
    if args.encode_streams > 1:
        res["batched"] = run_encode_batched(args, frames, roles, device_out)
    if not args.no_cpu:
        res.update(encode_cpu_chain(args, frames, roles, device_out))
    return res


def run_encode_batched(args, frames0, roles, device_out):
    """The same frame loop for several 1920x1080 streams at once: each step
    codes one frame of every stream in one encode launch (evxa_encode_batch),
    then deblocks and SAOs each.  Stream 0 codes the single-stream sequence,
    and its checked frames must equal the single-stream run's; the others
    code sequences of their own seeds."""
    from chicago_amd import evxa, synth
    w, h = 1920, 1080
    aw, ah = synth.align16(w), synth.align16(h)
    n_streams, n = args.encode_streams, args.encode_frames
    seqs = [frames0] + [synth.sequence_c1(synth.seed_of(1, k, 0), n, w, h)
                        for k in range(1, n_streams)]
    ctx = evxa.Context(0)
    n_dpb = 4
    streams = []
    for k in range(n_streams):
        s = ctx.open_stream(aw, ah, n_dpb + n)
        for d, f in enumerate(seqs[k]):
            s.upload(n_dpb + d, *f)
        streams.append(s)
    kind = {"I": evxa.FRAME_I, "P": evxa.FRAME_P, "B": evxa.FRAME_B}

    def step(r):
        refs = [] if r.kind == "I" else r.ref_slots
        d = evxa.encode_desc(kind[r.kind], r.quality, n_dpb + r.index, r.slot, refs,
                             target_ids=refs, temporal_mvp=r.kind == "P" and r.temporal_from >= 0)
        ctx.encode_batch(streams, [d] * n_streams)
        for s in streams:
            s.deblock(r.slot)
            s.sao(r.slot, n_dpb + r.index, r.quality, table=False)

    same = True
    for k, r in enumerate(roles):
        step(r)
        if k in device_out:
            rec, _, lv = streams[0].results()
            pic = streams[0].download(r.slot)
            want = device_out[k]
            same = same and (np.array_equal(rec, want[0]) and np.array_equal(lv, want[1])
                             and all(np.array_equal(a, b) for a, b in zip(pic, want[2])))
    ctx.synchronize()
    ctx.launch_times()
    t0 = time.perf_counter()
    for r in roles:
        step(r)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms = np.array(ctx.launch_times())
    per_step = 1 + 3 * n_streams          # encode, then deblock and SAO (2) per stream
    out = {"workload": f"{n_streams} 1920x1080 streams, the same GOP, one encode launch per step "
                       f"(one frame of every stream), deblocking and SAO per stream",
           "streams": n_streams, "ms_per_step": round(dt / len(roles) * 1e3, 3),
           "frames_per_s": round(n_streams * len(roles) / dt, 2),
           "stream0_equals_single_stream_run": bool(same)}
    if ms.size == per_step * len(roles):
        out["encode_ms_per_step"] = round(float(ms.reshape(len(roles), per_step)[:, 0].mean()), 3)
    for s in streams:
        s.close()
    ctx.close()
    return out


def encode_cpu_chain(args, frames, roles, device_out):
    """The same frame loop on the host -- the reference build's encode_slice
    (its row workers on the host's threads), deblocking and SAO, where the
    reference is built, else the oracle's -- from zeroed picture-buffer
    slots; compares the frames in `device_out` (descriptors, post-trellis
    levels, decoded picture) and times encode_slice per frame on one thread
    and on the host's threads over a bounded sample."""
    from oracle.oracle import Oracle, Picture, Reference, reference_available
    impl = Reference() if reference_available() else Oracle()
    kind = {"I": 0, "P": 1, "B": 2}
    h, w = frames[0][0].shape
    zero = (np.zeros((h, w), np.int16), np.zeros((h // 2, w // 2), np.int16),
            np.zeros((h // 2, w // 2), np.int16))
    dpb = [zero] * 4
    tables = {}
    checked = []
    if isinstance(impl, Reference):
        impl.encode_threads = args.cpu_threads
    t0 = time.perf_counter()
    for k, r in enumerate(roles):
        refs = [Picture(*dpb[sl]) for sl in ([] if r.kind == "I" else r.ref_slots)]
        prev = tables.get(r.temporal_from) if r.kind == "P" else None
        cur = Picture(*frames[r.index])
        recs, lv, pic = impl.encode_frame(kind[r.kind], r.quality, cur, refs, dpb[r.slot],
                                          prev=prev, target_ids=r.ref_slots or None)
        filt = impl.deblock_frame(recs, *pic)
        _, *dec = impl.sao_encode_frame(frames[r.index], filt, impl.lam(r.quality))
        dpb[r.slot] = tuple(dec)
        if r.kind != "B":
            tables[r.index] = recs
        if k in device_out:
            grec, glv, gpic = device_out[k]
            same = (all(np.array_equal(grec[f], recs[f]) for f in ENCODE_FIELDS)
                    and np.array_equal(glv, lv)
                    and all(np.array_equal(a, b) for a, b in zip(gpic, dec)))
            log(f"encode frame {k} ({r.kind}{r.index}): CPU chain bit-exact={same} "
                f"({time.perf_counter() - t0:.1f} s)")
            checked.append((k, bool(same)))
        if k >= max(device_out):
            break
    out = {"bit_exact_vs_cpu": all(c for _, c in checked),
           "checked_frames_decode_order": [k for k, _ in checked]}
    # CPU baseline: encode_slice alone on the first P anchor and the first
    # reference B of the sequence, one thread and the host's threads.
    if isinstance(impl, Reference):
        sample = [k for k, r in enumerate(roles) if r.kind in "PB"][:2]
        base = {}
        for threads in (1, args.cpu_threads):
            impl.encode_threads = threads
            dpb = [zero] * 4
            tables = {}
            spent, nf = 0.0, 0
            for k, r in enumerate(roles[:max(sample) + 1]):
                refs = [Picture(*dpb[sl]) for sl in ([] if r.kind == "I" else r.ref_slots)]
                prev = tables.get(r.temporal_from) if r.kind == "P" else None
                t1 = time.perf_counter()
                recs, lv, pic = impl.encode_frame(kind[r.kind], r.quality,
                                                  Picture(*frames[r.index]), refs, dpb[r.slot],
                                                  prev=prev, target_ids=r.ref_slots or None)
                if k in sample:
                    spent += time.perf_counter() - t1
                    nf += 1
                filt = impl.deblock_frame(recs, *pic)
                _, *dec = impl.sao_encode_frame(frames[r.index], filt, impl.lam(r.quality))
                dpb[r.slot] = tuple(dec)
                if r.kind != "B":
                    tables[r.index] = recs
            base[threads] = nf / spent
        out["cpu_baseline"] = {
            "encode_slice_frames_per_s_1_thread": round(base[1], 3),
            f"encode_slice_frames_per_s_{args.cpu_threads}_threads": round(base[args.cpu_threads], 3),
            "sample": f"the first P anchor and the first reference B (1920x1080), reference build, "
                      f"row workers on 1 and {args.cpu_threads} thread(s)",
            "kind": "reference"}
    return out


def calibrate_serving(ctx, args, shapes=None):
    """Sustained MB/s of the dispatcher on a backlog of mix frames (every
    stream of a 64-stream mix queued for several frames).  With `shapes`,
    the burst runs under each launch shape (outputs compared bit for bit
    with the first shape's) and the fastest is left applied; returns the
    capacity and the table."""
    from chicago_amd import evxa, serve, synth
    frames = serve.pool_content(args.serve_seed, [m[:2] for m in serve.C4_MIX], 4)
    descs = serve.descs_for(4, args.quality)
    rng = np.random.default_rng(args.serve_seed)
    shares = np.array([m[2] for m in serve.C4_MIX])
    streams = []
    for _ in range(64):
        w, h, _ = serve.C4_MIX[rng.choice(3, p=shares)]
        s = ctx.open_stream(synth.align16(w), synth.align16(h), 4)
        for k, f in enumerate(frames[(w, h)]):
            s.upload(k, *f)
        streams.append(s)
    n_per = 6
    mbs = sum(s.n_mbs for s in streams) * n_per

    def burst():
        # Two launches in flight whatever the serving setting, so the
        # capacity (and the admitted load) does not depend on it.
        with evxa.Server(ctx, 2, args.serve_max_frames) as srv:
            t0 = time.perf_counter()
            for k in range(1, n_per + 1):
                for i, s in enumerate(streams):
                    srv.submit(s, descs[k % 4], i * 100 + k)
            n = 0
            while n < 64 * n_per:
                n += len(srv.poll())
                time.sleep(0.0005)
            return time.perf_counter() - t0

    table, best, base = [], None, None
    for shape in (shapes or [None]):
        if shape is not None:
            ctx.set_tuning(**shape)
        burst()                                   # warm-up (and work lists)
        wall = burst()
        out = [s.results(coeffs=False, levels=False)[0] for s in streams[:8]]
        same = base is None or all(np.array_equal(a, b) for a, b in zip(out, base))
        base = base if base is not None else out
        cap = mbs / wall
        table.append({"tuning": shape, "mb_per_s": round(cap), "bit_identical": same})
        log(f"serve: calibration {shape} {cap / 1e6:.2f} M MB/s, bit-identical={same}")
        if not same:
            raise RuntimeError(f"launch shape {shape} changed the outputs")
        if best is None or cap > best[0]:
            best = (cap, shape)
    if best[1] is not None:
        ctx.set_tuning(**best[1])
    for s in streams:
        s.close()
    return best[0], table


TRAFFIC_COPY_BYTES = 256 << 20
TRAFFIC_PROFILE = {False: os.path.join(ROOT, "profiles", "analyze_kernel_traffic_1080p.md"),
                   True: os.path.join(ROOT, "profiles", "sweep_kernel_traffic_1080p.md")}


def run_traffic(args):
    """Launches for the HBM-traffic counter passes: a plain device-to-device
    copy of TRAFFIC_COPY_BYTES (the counters' calibration: it reads and
    writes exactly that many bytes), then analysis launches of frame 5 of the
    plain sequence (4 references)."""
    import torch

    from chicago_amd import evxa, synth

    w, h = 1920, 1080
    frames = synth.sequence_c1(synth.seed_of(1, 0, 0), 6, w, h)
    a = torch.empty(TRAFFIC_COPY_BYTES, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    ctx = evxa.Context(0)
    s = ctx.open_stream(synth.align16(w), synth.align16(h), 6)
    for t, f in enumerate(frames):
        s.upload(t, *f)
    for t in range(1, 6):
        if t == 1:
            s.reset_temporal()
        s.analyze_p(args.quality, t, frame_refs(t))
    ctx.synchronize()
    s.close()
    ctx.close()
    log(f"traffic: {TRAFFIC_COPY_BYTES} B copy x3, analysis of frames 1-5")


def measured_traffic(sweep_mode=False):
    """HBM bytes per analysis launch (frame 5, 4 refs) in the given mode from
    the committed counter summary: the line `traffic_bytes_per_launch: N`."""
    try:
        with open(TRAFFIC_PROFILE[sweep_mode]) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=118)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--quality", type=int, default=8)
    ap.add_argument("--workload", choices=["plain", "batched", "pair", "gop", "serve", "recon",
                                           "loop", "encode", "all", "traffic"],
                    default="all")
    ap.add_argument("--encode-frames", type=int, default=13)
    ap.add_argument("--encode-streams", type=int, default=8,
                    help="streams of the batched encode leg (1: none)")
    ap.add_argument("--loop-frames", type=int, default=24)
    ap.add_argument("--gop-frames", type=int, default=33)
    ap.add_argument("--serve-streams", type=int, default=200)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-seed", type=int, default=1)
    ap.add_argument("--serve-inflight", type=int, default=2,
                    help="launches in flight (2: lower p99 than 4 at every load measured)")
    ap.add_argument("--serve-max-frames", type=int, default=64)
    ap.add_argument("--serve-util", type=float, default=0.6)
    ap.add_argument("--serve-capacity", type=float, default=0.0,
                    help="admission capacity in MB/s (0: measure it with a burst)")
    ap.add_argument("--serve-curve", default="0.15,0.3",
                    help="further admitted-load fractions to run (latency against throughput)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--cpu-start", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the multi-threaded CPU baseline (the host's share of cores)")
    ap.add_argument("--cpu-mt-frames", type=int, default=24)
    ap.add_argument("--cpu-frames", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "stream with no cross-GPU exchange, so per-GPU numbers are the unit.",
              file=sys.stderr)
        sys.exit(2)
    if args.cpu_threads <= 0:
        args.cpu_threads, args.cpu_threads_rule = host_threads()
    else:
        args.cpu_threads_rule = "--cpu-threads"
    if args.workload == "traffic":
        run_traffic(args)
        return
    res = run_plain(args) if args.workload in ("all", "plain") else {"value": None}
    if args.workload in ("all", "batched"):
        from chicago_amd import evxa
        ctx = evxa.Context(0)
        b, streams, descs, seqs, outs = run_batched(args, ctx)
        if not args.no_cpu:
            b["bit_exact_vs_cpu"] = batched_cpu_check(seqs, descs, outs, args.quality, [0, 17, 41])
        if res.get("value"):
            b["vs_plain"] = round(b["frames_1080p_eq_per_s"] / res["value"], 2)
        res["batched"] = b
        for s in streams:
            s.close()
        ctx.close()
    if args.workload in ("all", "pair"):
        res["pair_720p"] = run_pair(args)
    if args.workload in ("all", "gop"):
        res["gop_b"] = run_gop(args)
    if args.workload in ("all", "recon"):
        res["reconstruct"] = run_recon(args)
    if args.workload in ("all", "loop"):
        res["closed_loop"] = run_loop(args)
    if args.workload in ("all", "encode"):
        res["encode"] = run_encode(args)
    if args.workload in ("all", "serve"):
        res["serve"] = run_serve(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
