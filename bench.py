"""Benchmark of the particle step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one frame (collide + integrate + stable re-bin) over the whole
world, inputs resident in HBM.  Default workload: 1,000,000 particles on a
jittered lattice filling the lower half of a 2048x1024 world, cell size 1,
gravity (0, -0.1).  Prints one JSON line.

roofline: the dominant kernel's algorithmic bytes per launch (per particle
and per cell figures below) over its average duration, measured with HIP
events that hipExtLaunchKernel stamps with each kernel's start and end on
the library's own stream (wrach_step_profiled).
cpu_baseline: the oracle's multi-threaded frame on the host, on a bounded
number of frames of the same workload.
upload: wall time of one wrach_add_particles of the workload into a fresh
context, binned on the device and packed on the host (not part of the step).
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
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
METRIC = "particle-steps/s & %HBM roofline @1M"


def algorithmic_bytes(n: int, cells: int, split: dict) -> dict:
    """Bytes each kernel must move per launch.  Per particle: the collide
    pass reads position+velocity (16) and writes them (16) plus the cell key
    (4) -- in collide_kernel for the particles of cells it handles inside a
    wavefront, in cells_kernel for the queued cells; the scatter reads key
    (4) + position+velocity (16) and writes them (16).  Per cell (C + 2
    offsets): the frame reads one offset (4), counted with collide_kernel,
    which learns the cells from a 4-byte key per particle instead; the scan
    reads a count (4), writes an offset (4) and clears the count (4)."""
    c2 = cells + 2
    return {"collide": 36 * split["collide"] + 4 * c2, "cells": 36 * (n - split["collide"]),
            "scan": 12 * c2, "scatter": 36 * n}


IN_WAVE_MAX = 3  # collide_kernel handles cells of up to 3 particles inside a 64-particle wavefront


def particles_by_kernel(idx: np.ndarray) -> dict:
    """Particles of the cells collide_kernel handles (at most IN_WAVE_MAX
    particles, inside one 64-particle wavefront window of the packed order),
    and of the cells it queues for cells_kernel: 2..64 particles (lane
    groups), 65..704 (the wide path), more (a block each)."""
    cnt = np.diff(idx[1:].astype(np.int64))
    start = idx[1:-1].astype(np.int64)
    occ = cnt > 0
    c, s = cnt[occ], start[occ]
    inside = (s // 64) == ((s + c - 1) // 64)
    col = inside & (c <= IN_WAVE_MAX)
    q = c[~col]
    return {"collide": int(c[col].sum()), "group": int(q[q <= 64].sum()),
            "wide": int(q[(q > 64) & (q <= 704)].sum()), "block": int(q[q > 704].sum())}


def host_cores() -> dict:
    """The host's cores as the process sees them: nproc (os.cpu_count), the
    CPUs this process may run on (sched_getaffinity) and the CPU share the
    environment grants (OMP_NUM_THREADS; 16 per GPU on the MI355X boxes)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    share = os.environ.get("OMP_NUM_THREADS", "")
    share = int(share) if share.isdigit() and int(share) > 0 else None
    threads = min(affinity, share) if share else affinity
    return {"nproc": nproc, "affinity": affinity, "share": share, "threads": threads}


def cpu_baseline(wl, budget_s: float = 12.0) -> dict:
    from oracle import oracle as o
    hc = host_cores()
    threads = hc["threads"]
    w, h = wl.config.dimensions
    idx, p, v = o.pack(wl.pos, wl.vel, (0, 0, w, h), wl.config.cell_size)
    P = o.params(w, h, wl.config.cell_size, len(p), gravity=wl.config.gravity, dt=wl.config.dt)
    t0 = time.perf_counter()
    idx, p, v = o.step(P, idx, p, v, nsteps=1, nthreads=threads)
    one = time.perf_counter() - t0
    frames = int(max(3, min(5000, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(P, idx, p, v, nsteps=frames, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": len(p) * frames / dt, "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "host": hc,
            "sample": f"{frames} frames of {wl.name} ({len(p)} particles) after 1 warm-up frame, "
                      f"oracle wo_step_mt on {threads} host threads (nproc {hc['nproc']}, "
                      f"{hc['affinity']} CPUs in the affinity mask, CPU share "
                      f"{hc['share'] if hc['share'] else 'not set'}), {dt:.2f} s"}


def copy_probe(nbytes: int) -> int:
    """A plain device-to-device copy of nbytes, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters (run alone, under the profiler).  Uses
    torch in a process that never loads libwrach_hip.so."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)
    return 0


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10) -> int:
    """Bandwidth of a plain device-to-device copy (bytes read + written per
    second, HIP events around `reps` torch copies of nbytes): the practical
    HBM ceiling beside the published peak.  Run as a child process of the
    bench (torch does not initialise HIP beside the library in one process)."""
    import torch
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        b.copy_(a)
    stop.record()
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    print(json.dumps({"copy_GBs": 2 * nbytes / (ms * 1e-3) / 1e9, "bytes": nbytes, "reps": reps}), flush=True)
    return 0


def measured_copy_gbs() -> float | None:
    """copy_bandwidth() in a child process; None if it fails."""
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-bandwidth"],
                           capture_output=True, text=True, timeout=180)
        return float(json.loads(r.stdout.strip().splitlines()[-1])["copy_GBs"])
    except Exception as e:  # noqa: BLE001 -- a missing figure must not end the bench
        print(f"[bench] device copy bandwidth not measured: {e}", file=sys.stderr, flush=True)
        return None


def upload_times(wl) -> dict:
    """Wall time of wrach_add_particles of the whole workload into a fresh
    context: the store binned on the device (the default; the particles cross
    PCIe once) and packed on the host (WRACH_FLAG_HOST_PACK; the packed arrays
    cross PCIe).  Both include the host-to-device copies."""
    from dataclasses import replace

    from wrach_amd import World

    out = {"particles": wl.n}
    for name, host in (("device_ms", False), ("host_ms", True)):
        with World(replace(wl.config, host_pack=host)) as w:
            t0 = time.perf_counter()
            w.add_particles(wl.pos, wl.vel)
            w.sync()
            out[name] = round((time.perf_counter() - t0) * 1e3, 3)
    return out


MANY_WORLDS = "many-worlds-512"


def bench_many_worlds(args) -> int:
    """512 independent worlds (sides 256/512/1024, 1,000..65,536 particles
    each) as one batch of worlds of different sizes (wrach_create_worlds):
    one launch per frame kernel for all of them over a flattened (world,
    cell) key space.  Per-frame wall time (host clock around step(1) + sync,
    one-frame graph) gives p50/p99; value is all particles x frames over the
    summed frame times."""
    from wrach_amd import World, workloads

    ws = workloads.many_worlds()
    t_up = time.perf_counter()
    b = World([w.config for w in ws])
    for k, w in enumerate(ws):
        b.add_particles(w.pos, w.vel, world=k)
    b.set_graph(not args.no_graph)
    if args.autotune > 0:
        b.step(args.warmup)
        b.autotune(args.autotune)
    b.sync()
    upload_s = time.perf_counter() - t_up
    n = sum(w.n for w in ws)
    cells = b.n_indices - 2
    print(f"[bench] {MANY_WORLDS}: {len(ws)} worlds in one batch, N={n} cells={cells} (with the padding "
          f"that puts every grid on a 1,024-cell boundary), upload {upload_s:.2f} s", file=sys.stderr, flush=True)
    b.step(args.warmup)
    b.prepare(1)  # the one-frame graph the timed loop replays
    b.sync()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        b.step(1)
        b.sync()
        times.append(time.perf_counter() - t0)
    times = np.array(times)
    elapsed = float(times.sum())
    value = n * args.steps / elapsed
    grid_cells = sum(w.cells for w in ws)
    step_bytes = 72 * n + 16 * (grid_cells + 2)
    prof = min(args.profile_steps, 50)
    ms = b.step_profiled(prof)
    kernel_avg_us = {k: 1e3 * v / prof for k, v in ms.items()}
    shapes = {k: v for k, v in b.tuning().items() if k != "frame_us"}
    b.close()
    print(f"[bench] p50 {np.percentile(times, 50) * 1e3:.3f} ms p99 {np.percentile(times, 99) * 1e3:.3f} ms "
          f"kernels {kernel_avg_us}", file=sys.stderr, flush=True)
    sides = {}
    for w in ws:
        d = sides.setdefault(str(w.config.dimensions[0]), {"worlds": 0, "particles": 0})
        d["worlds"] += 1
        d["particles"] += w.n
    line = {
        "metric": "particle-steps/s, batched worlds (p50/p99 frame time)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": MANY_WORLDS, "worlds": len(ws), "particles": n, "grid_cells": grid_cells,
                   "key_space_cells": cells, "batch": "one (wrach_create_worlds)", "sides": sides,
                   "launch_shapes": shapes, "graph": not args.no_graph,
                   "p50_frame_ms": float(np.percentile(times, 50) * 1e3),
                   "p99_frame_ms": float(np.percentile(times, 99) * 1e3),
                   "frame_clock": "host, step(1) + sync"},
        "roofline": {"bound": "hbm", "achieved": step_bytes / (elapsed / args.steps) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "scope": "whole frame, host clock",
                     "kernel_avg_us": kernel_avg_us, "kernel_frames": "the profiled frames after the timed ones"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def traffic_summary(args) -> int:
    """profiles/traffic_<workload>.json from two rocprofv3 --pmc runs of
    this bench (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) and the same
    two counters over a plain device copy of known size (--copy-probe): per
    frame kernel, the median over its last --profile-steps dispatches (the
    frames bench.py times with events, after --warmup + --steps), scaled by
    what the counters read for the copy (bytes read and written)."""
    import csv
    fetch_csv, write_csv, copy_fetch_csv, copy_write_csv = args.traffic_summary

    def rows(path):
        with open(path, newline="") as f:
            return list(csv.DictReader(f))

    def per_kernel(path, counter):
        out = {}
        for r in rows(path):
            if r.get("Counter_Name") != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
        return {k: [v for _, v in sorted(vs)] for k, vs in out.items()}

    def copy_value(path, counter):
        vals = [v for k, vs in per_kernel(path, counter).items() if "copyBuffer" in k for v in vs]
        return max(vals)

    probe_kib = args.copy_probe_bytes / 1024.0
    fetch_scale = probe_kib / copy_value(copy_fetch_csv, "FETCH_SIZE")   # copy reads probe bytes
    write_scale = probe_kib / copy_value(copy_write_csv, "WRITE_SIZE")   # and writes them
    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    names = {"collide": "collide_kernel", "cells": "cells_kernel", "scan": "scan_kernel",
             "scatter": "scatter_kernel"}
    kernels = {}
    def find(table, needle):
        # "::scatter_kernel<" or "::scatter_kernel(", not upload_scatter_kernel
        return next(k for k in table if f"::{needle}<" in k or f"::{needle}(" in k)

    for short, needle in names.items():
        fk, wk = find(fetch, needle), find(write, needle)
        f = float(np.median(fetch[fk][-args.profile_steps:])) * 1024 * fetch_scale
        w = float(np.median(write[wk][-args.profile_steps:])) * 1024 * write_scale
        kernels[short] = {"hbm_bytes_per_launch": int(f + w), "fetch_bytes": int(f), "write_bytes": int(w),
                          "dispatches": len(fetch[fk])}
    out = {"_about": "HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch), "
                     f"scaled by a {args.copy_probe_bytes} B device copy (FETCH x{fetch_scale:.3f}, WRITE "
                     f"x{write_scale:.3f}); median over each frame kernel's last {args.profile_steps} "
                     "dispatches, the frames bench.py times with events",
           "frames": [args.warmup + args.steps, args.warmup + args.steps + args.profile_steps],
           "fetch_scale": fetch_scale, "write_scale": write_scale, "kernels": kernels}
    path = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)
    return 0


LEARNED_METRIC = "learned/plain speedup at equal mean overlap"
# bf16 MFMA peak of MI355X (dense), and the network's MFMA work per particle:
# 2 x (64 x 128 + 128 x 128) multiply-adds; the 128 -> 2 layer (512 flop)
# runs on the vector ALUs.
MFMA_BF16_PEAK_TFLOPS = 2500.0
NET_MFMA_FLOP = 2 * (64 * 128 + 128 * 128)


def _frame_stats(world):
    """Mean overlap after the passes of the last frame (exact sums)."""
    _, _, _, (sb, nb, sa, na) = world.solver_stats()
    return sa / 2**24 / na if na else 0.0


def _timed(world, frames: int) -> float:
    world.prepare(frames)
    world.sync()
    t0 = time.perf_counter()
    world.step(frames)
    world.sync()
    return time.perf_counter() - t0


def learned_at(wl, K: int, args) -> dict:
    """One plain step size K (passes): a fresh world of the workload, its
    mode S rollout (state_passes passes per frame) sampled at the sample
    frames with targets after K passes, Adam on the device, then the plain
    step (K passes) against the learned step (k passes) from the checkpoint
    of the same rollout at eval_frame; k is the fewest passes whose mean
    overlap after the first frame is at most the plain step's."""
    from wrach_amd import World
    from wrach_amd import learn as wlearn

    n = wl.n
    world = World(wl.config)
    world.add_particles(wl.pos, wl.vel)
    t0 = time.perf_counter()
    frames = [int(f) for f in args.sample_frames.split(",")]
    rows = wlearn.collect(world, frames, args.state_passes, K)
    world.sync()
    collect_s = time.perf_counter() - t0
    untrained = wlearn.init_params(args.seed)
    world.learn_set_params(untrained)
    t0 = time.perf_counter()
    losses = wlearn.train(world, rows, args.train_steps, batch=8192, lr=1e-3, seed=args.seed)
    world.sync()
    train_s = time.perf_counter() - t0
    trained = world.learn_params()
    print(f"[learned] K={K}: {rows} pairs, {len(losses)} Adam steps in {train_s:.2f} s, loss "
          f"{losses[0]:.5f} -> {losses[-1]:.5f}", file=sys.stderr, flush=True)
    world.set_solver("S", args.state_passes, quiet=True)
    world.step(args.eval_frame - frames[-1])
    ck = world.read()

    def first_frame_mean(mode, passes, params=None):
        world.load_packed(*ck)
        if params is not None:
            world.learn_set_params(params)
        world.set_solver(mode, passes)
        world.step(1)
        return _frame_stats(world)

    plain_mean = first_frame_mean("S", K)
    ks = {}
    for name, params in (("trained", trained), ("untrained", untrained)):
        ks[name] = None
        for k in range(1, K + 1):
            if first_frame_mean("L", k, params) <= plain_mean:
                ks[name] = k
                break
    k = ks["trained"] or K
    T = args.steps
    world.load_packed(*ck)
    world.set_solver("S", K, quiet=True)
    t_plain = _timed(world, T)
    world.load_packed(*ck)
    world.learn_set_params(trained)
    world.set_solver("L", k, quiet=True)
    t_learned = _timed(world, T)
    per = {}
    for name, mode, passes in (("plain", "S", K), ("learned", "L", k)):
        world.load_packed(*ck)
        world.set_solver(mode, passes)
        means, falls = [], 0
        for _ in range(T):
            world.step(1)
            means.append(_frame_stats(world))
            if mode == "L":
                falls += int(world.learned_stats()["fell_back"])
        per[name] = {"mean_overlap_avg": float(np.mean(means)), "mean_overlap_max": float(np.max(means)),
                     "fallbacks": falls}
    world.load_packed(*ck)
    reps = 20
    prof = world.learn_profile(reps)
    world.close()
    net_ms, feat_ms = prof["network"] / reps, prof["features"] / reps
    out = {"plain_passes": K, "learned_passes": k, "passes_needed": ks, "training_pairs": rows,
           "train_steps": len(losses), "loss_first_last": [float(losses[0]), float(losses[-1])],
           "train_s": train_s, "collect_s": collect_s, "plain_first_frame_mean_overlap": plain_mean,
           "plain_ms_per_frame": t_plain * 1e3 / T, "learned_ms_per_frame": t_learned * 1e3 / T,
           "speedup": t_plain / t_learned, "over_timed_frames": per,
           "network_us": net_ms * 1e3, "features_us": feat_ms * 1e3,
           "network_tflops": n * NET_MFMA_FLOP / (net_ms * 1e-3) / 1e12}
    print(f"[learned] K={K}: plain {out['plain_ms_per_frame']:.3f} ms/frame, learned "
          f"{out['learned_ms_per_frame']:.3f} ms/frame ({k} passes), speedup {out['speedup']:.3f}; "
          f"network {out['network_us']:.1f} us, features {out['features_us']:.1f} us", file=sys.stderr,
          flush=True)
    return out


def bench_learned(args) -> int:
    """The learned warm start on the 1M lattice (config 1's setup), at each
    plain step size in --plain-passes (see learned_at); value is the
    speedup at the first of them."""
    from wrach_amd import workloads

    wl = workloads.lattice_1m()
    Ks = [int(k) for k in str(args.plain_passes).split(",")]
    print(f"[learned] {wl.name}: N={wl.n}, plain K in {Ks}", file=sys.stderr, flush=True)
    res = [learned_at(wl, K, args) for K in Ks]
    head = res[0]
    line = {
        "metric": LEARNED_METRIC,
        "value": head["speedup"],
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": head["learned_ms_per_frame"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": wl.name, "particles": wl.n, "state_passes": args.state_passes,
                   "sample_frames": args.sample_frames, "batch": 8192, "lr": 1e-3,
                   "eval_frame": args.eval_frame,
                   "equal_overlap": "mean overlap (1 - d over the overlapping pairs of the 3x3 cells, exact "
                                    "sums) after the passes of the first frame from the checkpoint, "
                                    "learned <= plain; targets after K passes",
                   "by_plain_passes": res},
        "roofline": {"bound": "mfma", "kernel": "learn_infer", "achieved": head["network_tflops"],
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": head["network_tflops"] / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "flop_per_launch": wl.n * NET_MFMA_FLOP, "kernel_avg_us": head["network_us"],
                     "features_avg_us": head["features_us"]},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="lattice-1m-2048x1024")
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--autotune", type=int, default=0, metavar="FRAMES",
                    help="run wrach_autotune over FRAMES frames per candidate after the warmup "
                         "(untimed) and report the chosen launch shapes")
    ap.add_argument("--passes", type=int, default=0, metavar="K",
                    help="mode S: K separation passes over the 3x3 cells around each particle "
                         "before the reference step (an extension; default 0, the reference's step)")
    ap.add_argument("--quiet", action="store_true",
                    help="with --passes: no overlap measure after the passes (WRACH_SOLVER_QUIET)")
    ap.add_argument("--learned", action="store_true",
                    help="the learned warm start against the plain step at equal mean overlap "
                         "(trains the network on the device first)")
    ap.add_argument("--plain-passes", default="64,32,16",
                    help="--learned: passes of the plain step, comma-separated (value: the first)")
    ap.add_argument("--state-passes", type=int, default=4, help="--learned: passes per rollout frame")
    ap.add_argument("--sample-frames", default="6,12", help="--learned: frames sampled for training")
    ap.add_argument("--train-steps", type=int, default=20000, help="--learned: Adam steps")
    ap.add_argument("--eval-frame", type=int, default=20, help="--learned: the checkpoint frame")
    ap.add_argument("--seed", type=int, default=0, help="--learned: initialisation and batch order")
    ap.add_argument("--traffic-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV",
                                                          "COPY_WRITE_CSV"),
                    help="write profiles/traffic_<workload>.json from rocprofv3 counter CSVs of this bench "
                         "run with the same --steps/--warmup/--profile-steps, and of --copy-probe")
    ap.add_argument("--copy-probe-bytes", type=int, default=268435456,
                    help="--traffic-summary: the bytes of the --copy-probe run")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="only measure a device-to-device copy's bandwidth (the bench's child process)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.traffic_summary:
        return traffic_summary(args)
    if args.copy_bandwidth:
        return copy_bandwidth()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the world would shard by row strips with a migrant exchange; not built).",
              file=sys.stderr)
        return 2

    if args.workload == MANY_WORLDS:
        return bench_many_worlds(args)
    if args.learned:
        return bench_learned(args)

    from wrach_amd import World, workloads

    wl = workloads.BY_NAME[args.workload]()
    n, cells = wl.n, wl.cells
    print(f"[bench] {wl.name}: N={n} cells={cells}", file=sys.stderr, flush=True)
    world = World(wl.config)
    world.add_particles(wl.pos, wl.vel)
    world.set_graph(not args.no_graph)
    if args.passes > 0:
        world.set_solver("S", args.passes, quiet=args.quiet)
    world.step(args.warmup)
    world.sync()
    tuning = world.autotune(args.autotune) if args.autotune > 0 else world.tuning()
    print(f"[bench] launch shapes: {json.dumps(tuning)}", file=sys.stderr, flush=True)
    # Capture the frame graphs the timed call launches (after any autotune,
    # which drops them) without running a frame: the timed region replays.
    world.prepare(args.steps)
    world.sync()

    t0 = time.perf_counter()
    world.step(args.steps)
    world.sync()
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed
    print(f"[bench] timed {args.steps} frames in {elapsed * 1e3:.2f} ms", file=sys.stderr, flush=True)
    solver = {"mode": "R"}
    if args.passes > 0:
        before, after, ran, (sb, nb, sa, na) = world.solver_stats()
        solver = {"mode": "S", "passes": args.passes, "overlap_tol": 0.0, "quiet": args.quiet,
                  "last_frame_max_overlap": [before, after], "last_frame_passes": ran,
                  "last_frame_mean_overlap": [sb / 2**24 / nb if nb else 0.0, sa / 2**24 / na if na else 0.0],
                  "last_frame_overlapping_pairs": [nb // 2, na // 2]}
        print(f"[bench] mode S: {json.dumps(solver)}", file=sys.stderr, flush=True)

    # Per-kernel durations: kernel-stamped HIP events, plain launches, over
    # the frames right after the timed ones.
    kernel_frames = [args.warmup + args.steps, args.warmup + args.steps + args.profile_steps]
    split_before = particles_by_kernel(world.read()[0])
    ms = world.step_profiled(args.profile_steps)
    split_after = particles_by_kernel(world.read()[0])
    avg_us = {k: 1e3 * v / args.profile_steps for k, v in ms.items()}
    split = {k: (split_before[k] + split_after[k]) // 2 for k in split_before}
    bytes_ = algorithmic_bytes(n, cells, split)
    dom = max(avg_us, key=avg_us.get)
    achieved = bytes_[dom] / (avg_us[dom] * 1e-6) / 1e9
    step_bytes = sum(bytes_.values())
    # Counter-measured HBM bytes of the dominant kernel (rocprofv3 FETCH_SIZE
    # and WRITE_SIZE runs, summarised by --traffic-summary), reported as
    # `traffic` only when they were taken over the same frames as the
    # kernel times above; otherwise beside them, with their own frames.
    traffic, traffic_other = None, None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{wl.name}.json")
    if os.path.exists(tpath) and args.passes == 0:
        tj = json.load(open(tpath))
        t = tj.get("kernels", {}).get(dom)
        if isinstance(t, dict):
            if tj.get("frames") == kernel_frames:
                traffic = t.get("hbm_bytes_per_launch")
            else:
                traffic_other = {"hbm_bytes_per_launch": t.get("hbm_bytes_per_launch"),
                                 "frames": tj.get("frames")}

    world.close()
    print(f"[bench] kernel avg us: {json.dumps({k: round(v, 3) for k, v in avg_us.items()})}",
          file=sys.stderr, flush=True)
    upload = upload_times(wl)
    print(f"[bench] upload ms: {json.dumps(upload)}", file=sys.stderr, flush=True)
    copy_gbs = measured_copy_gbs()
    if copy_gbs:
        print(f"[bench] device copy {copy_gbs:.0f} GB/s (read + write)", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline and args.passes == 0:
        cpu = cpu_baseline(wl)
        print(f"[bench] cpu baseline {cpu['value']:.3e} p-steps/s", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "particles": n,
            "world": list(wl.config.dimensions),
            "cell_size": wl.config.cell_size,
            "cells": cells,
            "gravity": list(wl.config.gravity),
            "dt": wl.config.dt,
            "graph": not args.no_graph,
            "launch_shapes": {k: v for k, v in tuning.items() if k != "frame_us"},
            "autotuned": args.autotune > 0,
            "graphs_captured_before_timing": not args.no_graph,
            "timed_frames": [args.warmup, args.warmup + args.steps],
            "solver": solver,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_other_frames": traffic_other,
            "algorithmic_bytes_per_launch": bytes_[dom],
            "kernel_avg_us": avg_us,
            "kernel_frames": kernel_frames,
            "particles_by_kernel": [split_before, split_after],
            "step_algorithmic_bytes": step_bytes,
            "step_achieved_GBs": step_bytes / (elapsed / args.steps) / 1e9,
            "step_frac": step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
            "measured_copy_GBs": copy_gbs,
            "frac_of_measured_copy": achieved / copy_gbs if copy_gbs else None,
        },
        "cpu_baseline": cpu,
        "upload": upload,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
