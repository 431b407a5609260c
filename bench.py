#!/usr/bin/env python3
"""Benchmark: PyroWave encode+decode on one MI355X.

Workload (the headline configuration): a seeded 60-frame 3840x2160 4:2:0
sequence (3 px/frame pan, 5 % of rectangles redrawn per frame), 1,600,000
bytes per frame.  A step encodes and decodes two batches of 32 frames that are
already resident in HBM, each on its own stream so the two overlap on the
device: pyrowave_hip_encode_batch, then pyrowave_hip_decode_encoded_batch
straight from the encoders' device results.  The transform rooflines and the
per-stage times come from a separate pass with one batch on one stream, so
kernel durations are not stretched by the overlap.

Reported: µs per frame of the FP16-storage batched path (value), the plain
FP32 single-frame path beside it, ΔPSNR between the two, the same batched
path with the LL of levels 0-1 also stored as fp16 (the reference's precision
1 storage: less traffic, lower PSNR), FP32 batched like the headline (batching
apart from storage precision), packed-FP16 lifting batched (the reference's
precision 0), the byte budget check, per-batch p50/p99
latency, the transform rooflines from HIP events recorded on the launch
stream, 1080p frames batched in FP16 and one at a time
in FP32, one batch of 32 uneven frames (720p to 4K, 4:2:0 and 4:4:4) eager
and as a replayed hipGraph, the serving loop (64 mixed sessions: latency
p50/p99 at cadence and saturated frames/s), and the oracle (the C
restatement, on up to 16 host threads) timed on a bounded sample as the CPU
baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

W, H = 3840, 2160
BUDGET = 1_600_000
FRAMES = 60
HBM_PEAK_GBS = 8000.0
LEVELS = 5
ROOF_STAGES = ("DWT", "iDWT")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--streams", type=int, default=2, help="batches in flight, one stream each")
    ap.add_argument("--cpu-frames", type=int, default=64, help="frames of the oracle CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: the host cores this process may use, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16-ll", action="store_true",
                    help="skip the FP16_LL leg (profiled runs: its transform launches have the headline's grids)")
    ap.add_argument("--no-extra-precisions", action="store_true",
                    help="skip the FP32-batched and packed-FP16 (FP16_MATH) legs")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the 1080p, uneven-batch and serving legs (profiled runs: some of their transform "
                         "launches have the grids of the headline batch's levels)")
    ap.add_argument("--lone-only", action="store_true",
                    help="only the lone 1080p FP16 frame (eager and graph): profiled runs of the latency path")
    ap.add_argument("--serving-seconds", type=float, default=2.0,
                    help="length of each serving-loop run (64 sessions); 0 skips it")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 256 MiB read + 256 MiB write elementwise kernel "
                         "(counter calibration for HBM traffic)")
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="summarize a rocprofv3 kernel trace CSV per (kernel, workgroups) and exit")
    ap.add_argument("--counter-table", nargs="+", metavar="CSV",
                    help="average of every counter per (kernel, workgroups) in rocprofv3 --pmc CSVs, and exit")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="per-launch HBM traffic from rocprofv3 FETCH_SIZE / WRITE_SIZE counter CSVs of a "
                         "--calibrate-copy run, scaled by the copy kernel's reading, and exit")
    return ap.parse_args()


def traffic_profile(batch):
    return os.path.join(ROOT, "profiles", f"traffic_4k420_fp16_b{batch}.txt")


def measured_traffic(direction, shapes, batch, fused):
    """HBM bytes per launch of one transform direction ("forward" or
    "inverse") with the given tile shape per level, at this batch size: the
    mean over the stage's launches of the committed rocprofv3 PMC figures
    (lines `traffic_level <direction> <level> <shape> <bytes>`; with a
    fused schedule, `traffic_fused <direction> <bytes>` for the launch of
    each component's first two levels and `traffic_fused_level <direction>
    2 <shape> <bytes>` for the level-2 launch, which then holds luma alone),
    or None when the profile of that batch size, or a launch's figure, is
    missing."""
    table = {}
    try:
        with open(traffic_profile(batch)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 5 and parts[0] in ("traffic_level", "traffic_fused_level") and parts[1] == direction:
                    table[(parts[0], int(parts[2]), int(parts[3]))] = int(parts[4])
                elif len(parts) == 3 and parts[0] == "traffic_fused" and parts[1] == direction:
                    table[("traffic_fused",)] = int(parts[2])
    except OSError:
        return None
    if fused:
        launches = [table.get(("traffic_fused",)), table.get(("traffic_fused_level", 2, shapes[2]))]
        launches += [table.get(("traffic_level", level, shapes[level])) for level in (3, 4)]
    else:
        launches = [table.get(("traffic_level", level, shape)) for level, shape in enumerate(shapes)]
    if not launches or None in launches:
        return None
    return int(round(sum(launches) / len(launches)))


# Fused forward tiles (kFusedTileW1 x kFusedTileH1 = 26 x 32 level-(l+1)
# coefficients) of one 3840x2160 4:2:0 frame: luma levels 0+1 (960 x 544 per
# band at level 1), each chroma plane levels 1+2 (480 x 272).
FUSED_TILES_4K420 = -(-960 // 26) * -(-544 // 32) + 2 * (-(-480 // 26) * -(-272 // 32))
FUSED_MIN_TILES = 16384  # the library fuses launches of at least this many tiles


def fused_forward(batch):
    """Whether a batch of 4K 4:2:0 frames takes the fused levels-0+1 launch."""
    mode = int(os.environ.get("PYROWAVE_FUSED", "1"), 0)
    return mode >= 2 or (mode == 1 and FUSED_TILES_4K420 * batch >= FUSED_MIN_TILES)


# Fused inverse tiles of one 3840x2160 4:2:0 frame: 120 plane samples wide,
# 64 rows high in launches of at least 16384 such tiles, 16 rows in launches
# of fewer than 4096 32-row tiles, else 32 rows.
INV_FUSED_TILE_W = 120
INV_FUSED_TALL_H, INV_FUSED_SHORT_H, INV_FUSED_LONE_H = 64, 32, 16
INV_FUSED_TALL_TILES, INV_FUSED_LONE_TILES = 16384, 4096


def inv_fused_tiles(batch, tile_h):
    per_frame = sum(-(-pw_ // INV_FUSED_TILE_W) * -(-ph // tile_h) for pw_, ph in ((W, H), (W // 2, H // 2),
                                                                                  (W // 2, H // 2)))
    return per_frame * batch


def inv_fused_tile_h(batch):
    if inv_fused_tiles(batch, INV_FUSED_TALL_H) >= INV_FUSED_TALL_TILES:
        return INV_FUSED_TALL_H
    return INV_FUSED_LONE_H if inv_fused_tiles(batch, INV_FUSED_SHORT_H) < INV_FUSED_LONE_TILES else INV_FUSED_SHORT_H


def fused_inverse(batch):
    """Whether a batch of 4K 4:2:0 frames takes the fused inverse launch of
    each component's first two levels."""
    mode = int(os.environ.get("PYROWAVE_FUSED", "1"), 0)
    return mode >= 1


def dwt_algorithmic_bytes(fp16):
    """Compulsory forward-DWT traffic of one frame: input planes at 1 B/px plus
    every transmitted coefficient at its storage width (fp16 storage applies
    to the high bands of levels 0-1)."""
    aw, ah = max((W + 31) // 32 * 32, 128), max((H + 31) // 32 * 32, 128)
    total = W * H + 2 * (W // 2) * (H // 2)
    for c in range(3):
        for l in range(5):
            if c and l == 0:
                continue
            bw, bh = (aw >> 1) >> l, (ah >> 1) >> l
            for b in range(4):
                if b == 0 and l != 4:
                    continue
                width = 2 if (fp16 and l < 2 and b != 0) else 4
                total += bw * bh * width
    return total


def short_kernel_name(full):
    """`k_name<template args>` out of a demangled kernel name: the template
    arguments name the transform tile shape, so two shapes whose launches
    happen to have the same workgroup count stay apart."""
    import re
    m = re.search(r"(k_[a-z_0-9]+)(<[^>]*>)?", full)
    if not m:
        return None
    return m.group(1) + (m.group(2) or "").replace(" ", "")


def trace_summary(path, top=250):
    """Per-(kernel, workgroup count) dispatch statistics of a rocprofv3
    kernel-trace CSV.  The workgroup count separates launches of one kernel
    over different batch sizes and transform levels."""
    import collections
    import csv
    agg = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            name = short_kernel_name(r["Kernel_Name"]) or r["Kernel_Name"][:48]
            wgs = (int(r["Grid_Size_X"]) // max(int(r["Workgroup_Size_X"]), 1)) * \
                  (int(r["Grid_Size_Y"]) // max(int(r["Workgroup_Size_Y"]), 1)) * \
                  (int(r["Grid_Size_Z"]) // max(int(r["Workgroup_Size_Z"]), 1))
            agg[(name, wgs)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    total = sum(sum(v) for v in agg.values())
    print(f"{'kernel':48s} {'workgroups':>10s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'share':>6s}")
    for (name, wgs), v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        print(f"{name:48s} {wgs:10d} {len(v):6d} {sum(v) / len(v) / 1e3:9.2f} {min(v) / 1e3:9.2f} "
              f"{max(v) / 1e3:9.2f} {sum(v) / total:6.1%}")


def counter_table(paths):
    """Average of each counter per (kernel, workgroup count) over the
    dispatches of rocprofv3 --pmc CSV files."""
    import collections
    import csv
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        with open(path) as f:
            for r in csv.DictReader(f):
                name = short_kernel_name(r["Kernel_Name"])
                if not name:
                    continue
                wgs = int(r["Grid_Size"]) // max(int(r["Workgroup_Size"]), 1)
                agg[(name, wgs)][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for v in agg.values() for c in v})
    print(f"{'kernel':40s} {'workgroups':>10s} " + " ".join(f"{n:>22s}" for n in names))
    for key in sorted(agg, key=lambda k: (k[0], -k[1])):
        v = agg[key]
        print(f"{key[0]:40s} {key[1]:10d} " + " ".join(
            f"{sum(v[n]) / len(v[n]):22.4g}" if v[n] else f"{'-':>22s}" for n in names))
    # Per-wave figures of each kernel's largest launch.
    print()
    largest = {}
    for k, wgs in agg:
        largest[k] = max(largest.get(k, 0), wgs)
    for k in sorted(largest):
        v = agg[(k, largest[k])]
        waves = sum(v["SQ_WAVES"]) / len(v["SQ_WAVES"]) if v.get("SQ_WAVES") else 0
        if not waves:
            continue
        parts = []
        for n, label in (("SQ_INSTS_VALU", "VALU"), ("SQ_INSTS_LDS", "LDS"), ("SQ_INSTS_SALU", "SALU"),
                         ("SQ_INSTS_VMEM", "VMEM"), ("SQ_LDS_BANK_CONFLICT", "LDS bank-conflict cycles"),
                         ("SQ_WAIT_INST_LDS", "LDS wait cycles")):
            if v.get(n):
                parts.append(f"{sum(v[n]) / len(v[n]) / waves:.0f} {label}")
        print(f"{k} ({largest[k]} workgroups): " + ", ".join(parts) + " per wave")


def _counter_rows(path):
    import csv
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            short = short_kernel_name(name) or ("copy_neg" if "neg" in name else name[:40])
            grid = int(r["Grid_Size"]) if "Grid_Size" in r else int(r["Grid_Size_X"])
            wg = int(r["Workgroup_Size"]) if "Workgroup_Size" in r else int(r["Workgroup_Size_X"])
            yield short, grid // max(wg, 1), float(r["Counter_Value"])


def pmc_summary(fetch_csv, write_csv):
    """Average FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per (kernel,
    workgroups); the elementwise negation of --calibrate-copy reads and
    writes exactly 256 MiB, which gives each counter's scale factor."""
    import collections
    out = {}
    for label, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        agg = collections.defaultdict(list)
        for name, wgs, v in _counter_rows(path):
            agg[(name, wgs)].append(v)
        copy = [sum(v) / len(v) for (n, _), v in agg.items() if n == "copy_neg"]
        factor = (256 * 1024) / copy[0] if copy else 1.0
        out[label] = (agg, factor)
        print(f"{label}: copy kernel reads {copy[0] if copy else float('nan'):.0f} KiB for 262144 KiB "
              f"-> scale factor {factor:.4f}")
    fagg, ff = out["FETCH_SIZE"]
    wagg, wf = out["WRITE_SIZE"]
    print(f"{'kernel':40s} {'workgroups':>10s} {'calls':>6s} {'fetch_MB':>10s} {'write_MB':>10s} {'traffic_MB':>10s}")
    for key in sorted(fagg, key=lambda k: -sum(fagg[k])):
        fv = sum(fagg[key]) / len(fagg[key]) * ff * 1024
        wv = sum(wagg.get(key, [0])) / max(len(wagg.get(key, [0])), 1) * wf * 1024
        print(f"{key[0]:40s} {key[1]:10d} {len(fagg[key]):6d} {fv / 1e6:10.3f} {wv / 1e6:10.3f} {(fv + wv) / 1e6:10.3f}")
    # Per (direction, level, shape) of the headline batch: the row of the
    # launch with that shape's workgroup count at that level.
    print()
    for direction, shapes in TRANSFORM_SHAPES.items():
        for level in range(LEVELS):
            for shape, (kernel, tw, th, per_wg) in enumerate(shapes):
                wgs = -(-transform_tiles(level, tw, th, 32) // per_wg)
                key = (kernel, wgs)
                if key not in fagg:
                    print(f"# missing {direction} level {level} shape {shape} ({kernel}, {wgs} workgroups)")
                    continue
                fv = sum(fagg[key]) / len(fagg[key]) * ff * 1024
                wv = sum(wagg.get(key, [0])) / max(len(wagg.get(key, [0])), 1) * wf * 1024
                print(f"traffic_level {direction} {level} {shape} {int(round(fv + wv, -3))}")

    def row(key):
        fv = sum(fagg[key]) / len(fagg[key]) * ff * 1024
        wv = sum(wagg.get(key, [0])) / max(len(wagg.get(key, [0])), 1) * wf * 1024
        return int(round(fv + wv, -3))

    # The fused schedules of the headline batch: each component's first two
    # levels in one launch, luma level 2 alone, then levels 3 and 4 as above.
    th = inv_fused_tile_h(32)
    fused_keys = {"forward": ("k_dwt_forward_fused", -(-FUSED_TILES_4K420 * 32 // 4)),
                  "inverse": (f"k_dwt_inverse_fused<{th}>", -(-inv_fused_tiles(32, th) // 4))}
    for direction, key in fused_keys.items():
        if key in fagg:
            print(f"traffic_fused {direction} {row(key)}")
        else:
            print(f"# missing fused {direction} ({key[0]}, {key[1]} workgroups)")
        for shape, (kernel, tw, th_, per_wg) in enumerate(TRANSFORM_SHAPES[direction]):
            key = (kernel, -(-transform_tiles(2, tw, th_, 32, planes=1) // per_wg))
            if key in fagg:
                print(f"traffic_fused_level {direction} 2 {shape} {row(key)}")


# Transform tile shapes by index: (kernel, tile width, tile height, tiles per workgroup).
TRANSFORM_SHAPES = {
    "forward": [("k_dwt_forward<128,56,true>", 128, 56, 1), ("k_dwt_forward<32,32,false>", 32, 32, 1),
                ("k_dwt_forward<32,16,false>", 32, 16, 1), ("k_dwt_forward<128,40,false>", 128, 40, 1),
                ("k_dwt_forward_stream", 120, 32, 4)],
    "inverse": [("k_dwt_inverse<64,64,256,16,false>", 64, 64, 1), ("k_dwt_inverse<32,32,128,16,false>", 32, 32, 1),
                ("k_dwt_inverse<16,16,64,16,false>", 16, 16, 1), ("k_dwt_inverse<64,48,256,24,true>", 64, 48, 1),
                ("k_dwt_inverse_stream", 120, 32, 4)],
}


def transform_tiles(level, tw, th, batch, planes=None):
    """Tiles of one transform launch over a batch of 4K 4:2:0 frames: level 0
    is the luma plane alone; launch l > 0 runs luma level l with chroma level
    l - 1, three planes of the same (aligned) size (`planes` overrides)."""
    w, h = 3840 >> level, 2176 >> level
    if planes is None:
        planes = 1 if level == 0 else 3
    return planes * -(-w // tw) * -(-h // th) * batch


SERVING_SHAPES = ((1280, 720, 0), (1920, 1080, 0), (2560, 1440, 1), (3840, 2160, 0))


def reference_cpu_probe():
    """Whether the reference itself could run on these host cores: it needs a
    Vulkan loader with Mesa's lavapipe ICD and a Granite checkout to build.
    Returns None when all are there, else the reason it cannot."""
    import ctypes.util
    import glob
    import shutil
    missing = []
    if not ctypes.util.find_library("vulkan"):
        missing.append("no Vulkan loader (libvulkan)")
    icds = []
    for d in ("/usr/share/vulkan/icd.d", "/etc/vulkan/icd.d", "/usr/local/share/vulkan/icd.d"):
        icds += glob.glob(os.path.join(d, "lvp_icd*.json"))
    if not icds:
        missing.append("no lavapipe ICD (lvp_icd*.json)")
    if not shutil.which("vulkaninfo"):
        missing.append("no vulkaninfo")
    if not glob.glob("/**/Granite/vulkan", recursive=False) and not os.environ.get("GRANITE_DIR"):
        missing.append("no Granite checkout (the reference build fetches it from the network)")
    return "; ".join(missing) if missing else None


def cpu_baseline(host_frames, W, H, budget, n, threads):
    """The oracle's C restatement of the codec (not the reference) on the host
    cores: encode + packetize + decode of n frames of the sequence, one frame
    per thread at a time (frames are independent), wall time over all."""
    import concurrent.futures as cf
    from oracle import oracle as O
    if threads <= 0:
        threads = min(16, len(os.sched_getaffinity(0)))
    L = O.layout(W, H, 0)

    def one(i):
        bs, meta, _ = O.encode(L, host_frames[i % len(host_frames)], budget, sequence=1)
        out, pk = O.packetize(L, meta, bs, 1, 1 << 30)
        O.decode(L, [out[o:o + s].tobytes() for o, s in pk])

    t = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, range(n)))
    cpu_s = time.perf_counter() - t
    probe = reference_cpu_probe()
    return {"value": round(cpu_s / n * 1e6, 1), "unit": "us/frame", "cores": threads, "kind": "port",
            "sample": f"{n} frames of the 4K 4:2:0 sequence at {budget} B, encode+packetize+decode by the "
                      f"oracle's C restatement (not the reference), {threads} threads, one frame each at a time "
                      f"({cpu_s:.1f} s wall)",
            "reference_cpu": "reference CPU timing unavailable: " + probe if probe else "lavapipe present"}


def serving_leg(pw, torch, device, dev_frames, seconds, sessions=64, seed=4):
    """The serving loop: `sessions` sessions cycling through 720p/1080p 4:2:0,
    1440p 4:4:4 and 4K 4:2:0 at 60 or 120 fps, Poisson arrivals with 1 ms
    jitter, 1.2-2.0 bits/pixel, admitted by footprint, encode+decode per frame.
    Two runs: at cadence (latency p50/p99 of submit -> bitstream on the host)
    and saturated (every session keeps two frames queued: frames/s)."""
    rng = np.random.default_rng(seed)
    plan = []
    for i in range(sessions):
        w, h, c = SERVING_SHAPES[i % 4]
        plan.append(dict(shape=(w, h, c), fps=(60, 120)[(i // 4) % 2],
                         budget=int(w * h * rng.uniform(1.2, 2.0) / 8)))

    # Content: crops of the resident 4K sequence (4:4:4 chroma planes are luma
    # crops of neighbouring frames); every session gets its own decode target.
    def frame_of(shape, k):
        w, h, c = shape
        if c:
            planes = [dev_frames[(k + j) % len(dev_frames)][0][:h, :w] for j in range(3)]
        else:
            f = dev_frames[k % len(dev_frames)]
            planes = [f[0][:h, :w], f[1][:h // 2, :w // 2], f[2][:h // 2, :w // 2]]
        return pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in planes])

    frames = {s: [frame_of(s, k) for k in range(8)] for s in SERVING_SHAPES}
    targets = []
    for s in plan:
        w, h, c = s["shape"]
        ch, cw = (h, w) if c else (h // 2, w // 2)
        outs = [torch.empty((h, w), dtype=torch.uint8, device="cuda")] + \
               [torch.empty((ch, cw), dtype=torch.uint8, device="cuda") for _ in range(2)]
        targets.append((outs, pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs])))

    probe = pw.Server(device, 1 << 40)
    footprint = sum(probe.footprint(*s["shape"]) for s in plan)
    probe.close()
    server = pw.Server(device, footprint, streams=4, precision=pw.PRECISION_FP16, packet_boundary=8192)
    sids = [server.open_session(*s["shape"]) for s in plan]
    assert all(x is not None for x in sids)

    def run(saturated):
        lat, starts, count = [], set(), 0
        base = [rng.uniform(0, 1.0 / s["fps"]) for s in plan]
        next_t = list(base)
        k = [0] * len(plan)
        inflight = [0] * len(plan)
        t0 = time.perf_counter()
        end = t0 + seconds
        while True:
            now = time.perf_counter()
            if now < end:
                for i, s in enumerate(plan):
                    due = inflight[i] < 2 if saturated else (now - t0) >= next_t[i]
                    if due:
                        server.submit(sids[i], frames[s["shape"]][k[i] % 8], s["budget"], targets[i][1])
                        k[i] += 1
                        inflight[i] += 1
                        base[i] += rng.exponential(1.0 / s["fps"])
                        next_t[i] = base[i] + rng.normal(0, 1e-3)
            elif sum(inflight) == 0:
                break
            for c in server.poll(256, timeout_us=0 if saturated else 50):
                assert c.status == pw.SUCCESS, c.status
                server.take_bitstream(c)
                i = sids.index(c.session)
                inflight[i] -= 1
                count += 1
                lat.append((c.done_ns - c.submit_ns) / 1e3)
                starts.add(c.start_ns)
        wall = time.perf_counter() - t0
        lat = np.array(lat)
        return {"frames": count, "seconds": round(wall, 3), "frames_per_s": round(count / wall, 1),
                "latency_us": {"p50": round(float(np.percentile(lat, 50)), 1),
                               "p99": round(float(np.percentile(lat, 99)), 1)},
                "launches": len(starts), "mean_frames_per_launch": round(count / max(len(starts), 1), 2)}

    run(True)  # warm the lanes' table storage
    cadence = run(False)
    saturated = run(True)
    for sid in sids:
        server.close_session(sid)
    server.close()
    offered = sum(s["fps"] for s in plan)
    return {"sessions": sessions, "mix": "720p/1080p 4:2:0, 1440p 4:4:4, 4K 4:2:0 at 60/120 fps, 1.2-2.0 bpp",
            "streams": 4, "footprint_bytes": int(footprint), "offered_frames_per_s": offered,
            "cadence": cadence, "saturated": saturated}


def main():
    args = parse()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.counter_table:
        counter_table(args.counter_table)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(frames shard across GPUs with no exchange; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch  # first: the library binds to torch's HIP runtime
    import pyrowave_amd as pw
    from pyrowave_amd import synth

    t0 = time.time()
    gen = synth.SceneGenerator(W, H, True, seed=2024, frames=FRAMES, pan=3, redraw=0.05)
    host_frames = [gen.frame(t) for t in range(FRAMES)]
    print(f"[bench] generated {FRAMES} frames in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev_frames = [[torch.from_numpy(p).cuda() for p in f] for f in host_frames]
    torch.cuda.synchronize()
    if args.calibrate_copy:
        # 64 Mi floats: the fill writes 256 MiB, the negation reads 256 MiB and
        # writes 256 MiB.
        a = torch.ones(64 << 20, device="cuda")
        b = torch.empty_like(a)
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        del a, b

    device = pw.Device()
    # An explicit stream: the library launches on it, and the latency events
    # below are recorded on the same stream.
    stream = torch.cuda.Stream()
    sptr = stream.cuda_stream

    def make(prec, n):
        encs = [pw.Encoder(device, W, H, pw.CHROMA_420, precision=prec) for _ in range(n)]
        decs = [pw.Decoder(device, W, H, pw.CHROMA_420, precision=prec) for _ in range(n)]
        outs = [[torch.empty_like(p) for p in dev_frames[0]] for _ in range(n)]
        tgts = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in out]) for out in outs]
        return encs, decs, outs, tgts

    def frame_desc(i):
        return pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in dev_frames[i % FRAMES]])

    def run_steps(groups, steps, start, latencies=None):
        """Each step: every group (B encoders + decoders on its own stream)
        encodes and decodes its next B frames."""
        for s in range(steps):
            for g, (encs, decs, outs, tgts, st) in enumerate(groups):
                B = len(encs)
                fr = [frame_desc(start + (s * len(groups) + g) * B + i) for i in range(B)]
                if latencies is not None:
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(st)
                pw.encode_batch(encs, fr, [BUDGET] * B, stream=st.cuda_stream)
                pw.decode_encoded_batch(decs, encs, tgts, stream=st.cuda_stream)
                if latencies is not None:
                    e1.record(st)
                    latencies.append((e0, e1))

    def timed(prec, B, steps, warmup, nstreams):
        """Throughput with `nstreams` batches of B frames in flight on their own
        streams; then, on one stream alone, the transform stage events for the
        rooflines (one-stream throughput) and the per-stage times."""
        groups = [make(prec, B) + ((stream if g == 0 else torch.cuda.Stream()),) for g in range(nstreams)]
        device.set_timing(False)
        # Transform tile shapes per level, timed on this batch (the other
        # groups' plans start from the same choice: same level shapes).
        encs0, decs0, _, tgts0, st0 = groups[0]
        tuned = pw.autotune_batch(encs0, [frame_desc(i) for i in range(B)], [BUDGET] * B, decs0, tgts0,
                                  repetitions=5, stream=st0.cuda_stream)
        run_steps(groups, warmup, 0)
        torch.cuda.synchronize()
        lat = []
        t = time.perf_counter()
        run_steps(groups, steps, warmup * B * nstreams, lat)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        lat_ms = np.array([a.elapsed_time(b) for a, b in lat])
        # One stream: HIP events on the launch stream around the forward and
        # inverse transform stages only (four event records per step).
        one = groups[:1]
        device.set_timing_stages(ROOF_STAGES)
        device.stage_times(reset=True)
        t = time.perf_counter()
        run_steps(one, steps, 0)
        torch.cuda.synchronize()
        wall_one = time.perf_counter() - t
        roof = device.stage_times(reset=True)
        # Stage times (events on the launch stream) over a separate pass.
        device.set_timing(True)
        device.stage_times(reset=True)
        run_steps(one, max(2, steps // 3), 0)
        torch.cuda.synchronize()
        stages = device.stage_times(reset=True)
        device.set_timing(False)
        encs, decs, outs = groups[0][0], groups[0][1], groups[0][2]
        extra = [x for g in groups[1:] for x in g[0] + g[1]]
        return dict(encs=encs, decs=decs, outs=outs, extra=extra, wall=wall, wall_one=wall_one, lat_ms=lat_ms,
                    stages=stages, roof=roof, B=B, nstreams=nstreams, tuned=tuned)

    B = args.batch
    # Launch-bound case: one 1920x1080 4:2:0 frame per call (crop of the
    # resident sequence), FP16, eager launches vs a replayed hipGraph.
    def single_1080p(graphs, steps=60, warmup=5):
        d2 = pw.Device()
        d2.set_graphs(graphs)
        d2.set_timing(False)
        e = pw.Encoder(d2, 1920, 1080, pw.CHROMA_420, precision=pw.PRECISION_FP16)
        dd = pw.Decoder(d2, 1920, 1080, pw.CHROMA_420, precision=pw.PRECISION_FP16)
        out = [torch.empty((1080, 1920), dtype=torch.uint8, device="cuda"),
               torch.empty((540, 960), dtype=torch.uint8, device="cuda"),
               torch.empty((540, 960), dtype=torch.uint8, device="cuda")]
        tg = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in out])]
        frs = [pw.hip_frame([(f[0].data_ptr(), f[0].stride(0), 1), (f[1].data_ptr(), f[1].stride(0), 1),
                             (f[2].data_ptr(), f[2].stride(0), 1)]) for f in dev_frames[:8]]
        for k in range(warmup):
            pw.encode_batch([e], [frs[k % 8]], [400000], stream=sptr)
            pw.decode_encoded_batch([dd], [e], tg, stream=sptr)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(steps):
            pw.encode_batch([e], [frs[k % 8]], [400000], stream=sptr)
            pw.decode_encoded_batch([dd], [e], tg, stream=sptr)
        torch.cuda.synchronize()
        us = (time.perf_counter() - t) / steps * 1e6
        e.close()
        dd.close()
        d2.close()
        return round(us, 2)

    if args.lone_only:
        print(json.dumps({"single_1080p_fp16_us_per_frame": {"eager": single_1080p(False),
                                                              "graph": single_1080p(True)}}), flush=True)
        return

    fp16 = timed(pw.PRECISION_FP16, B, args.steps, args.warmup, args.streams)
    print(f"[bench] fp16 batched done", file=sys.stderr, flush=True)
    # A lone frame takes ~0.15 ms: enough steps that the wall clock is not
    # dominated by the first launches.
    fp32_steps = max(args.steps, 60)
    fp32 = timed(pw.PRECISION_FP32, 1, fp32_steps, max(args.warmup, 10), 1)
    print(f"[bench] fp32 single done", file=sys.stderr, flush=True)
    # The reference's precision 1 storage exactly (the LL of levels 0-1 in
    # fp16 too): less traffic, lower PSNR; batched like the headline.
    fp16_ll = None
    if not args.no_fp16_ll:
        fp16_ll = timed(pw.PRECISION_FP16_LL, B, args.steps, args.warmup, args.streams)
        print(f"[bench] fp16_ll batched done", file=sys.stderr, flush=True)

    # FP32 batched like the headline (separates batching from storage
    # precision) and packed-FP16 lifting (the reference's precision 0).
    fp32b = fp16m = None
    if not args.no_extra_precisions:
        fp32b = timed(pw.PRECISION_FP32, B, args.steps, args.warmup, args.streams)
        print(f"[bench] fp32 batched done", file=sys.stderr, flush=True)
        fp16m = timed(pw.PRECISION_FP16_MATH, B, args.steps, args.warmup, args.streams)
        print(f"[bench] fp16_math batched done", file=sys.stderr, flush=True)

    us_fp16 = fp16["wall"] / (args.steps * B * args.streams) * 1e6
    us_fp16_one = fp16["wall_one"] / (args.steps * B) * 1e6
    us_fp32 = fp32["wall"] / fp32_steps * 1e6

    # Quality and budget on one batch of frames: both paths decode frames 0..B-1.
    psnr = {}
    max_bytes = 0
    legs = [(n, r) for n, r in (("fp16", fp16), ("fp32", fp32), ("fp16_ll", fp16_ll), ("fp32_batched", fp32b),
                                ("fp16_math", fp16m)) if r is not None]
    for name, r in legs:
        encs, decs, outs = r["encs"], r["decs"], r["outs"]
        n = len(encs)
        vals = []
        for s in range(max(1, 4 // n)):
            fr = [frame_desc(s * n + i) for i in range(n)]
            tg = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in out]) for out in outs]
            pw.encode_batch(encs, fr, [BUDGET] * n, stream=sptr)
            pw.decode_encoded_batch(decs, encs, tg, stream=sptr)
            torch.cuda.synchronize()
            for i in range(min(n, 4)):
                src = host_frames[(s * n + i) % FRAMES]
                dec = [o.cpu().numpy() for o in outs[i]]
                vals.append(synth.psnr(src[0], dec[0]))
                out, pk = encs[i].packetize(1 << 30, BUDGET + 64)
                max_bytes = max(max_bytes, sum(sz for _, sz in pk))
        psnr[name] = float(np.mean(vals[:4]))

    # Rooflines of the transform kernels from the events of the timed region.
    # One batch runs one launch per level (level by level) or, with a fused
    # schedule, four: each component's first two levels, luma level 2,
    # levels 3 and 4.  A launch's algorithmic bytes are the batch's
    # compulsory bytes / launches, its duration the stage time / launches,
    # so achieved = compulsory bytes / stage time.
    fused = fused_forward(B)
    fused_inv = fused_inverse(B)

    def kernels_of(base, shapes, fused_first=False):
        """The kernels one transform stage launches: shape 4 is the streamed kernel."""
        names = {base + "_stream" if v == 4 else base for v in (shapes[2:] if fused_first else shapes)}
        if fused_first:
            names.add(base + "_fused")
        return "+".join(sorted(names))

    def roofline(stage, kernel, direction, shapes, fused_first=False):
        ms, frames = fp16["roof"][stage]
        per_batch = LEVELS - 1 if fused_first else LEVELS
        launches = args.steps * per_batch
        per_frame = dwt_algorithmic_bytes(True)
        bytes_per_launch = per_frame * frames / max(launches, 1)
        us_per_launch = ms * 1e3 / max(launches, 1)
        achieved = bytes_per_launch / (us_per_launch * 1e-6) / 1e9
        return {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(direction, shapes, args.batch, fused_first),
                "algorithmic_bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(us_per_launch, 2),
                "launches_timed": launches, "launches_per_batch": per_batch}

    single = f1080 = uneven = None
    if not args.headline_only:
        single = {"eager": single_1080p(False), "graph": single_1080p(True)}
        print(f"[bench] single-frame graphs done", file=sys.stderr, flush=True)

    # 1920x1080 4:2:0 (crops of the resident sequence) at 400,000 B per frame:
    # FP16 batched like the headline (batches of B frames on `streams`
    # streams) and the plain FP32 single-frame path, µs per frame.
    # Ten warmup steps: the leg's 64 fresh encoder/decoder slots run slower at
    # first (measured: the same leg run twice in one process, 29.6 then 22.8
    # us/frame with three warmup steps).
    def frames_1080p(prec, n, nstreams, steps=20, warmup=10):
        d2 = pw.Device()
        d2.set_timing(False)
        groups = []
        for g in range(nstreams):
            encs = [pw.Encoder(d2, 1920, 1080, pw.CHROMA_420, precision=prec) for _ in range(n)]
            decs = [pw.Decoder(d2, 1920, 1080, pw.CHROMA_420, precision=prec) for _ in range(n)]
            outs = [[torch.empty((1080, 1920), dtype=torch.uint8, device="cuda"),
                     torch.empty((540, 960), dtype=torch.uint8, device="cuda"),
                     torch.empty((540, 960), dtype=torch.uint8, device="cuda")] for _ in range(n)]
            tgs = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in out]) for out in outs]
            groups.append((encs, decs, outs, tgs, stream if g == 0 else torch.cuda.Stream()))
        frs = [pw.hip_frame([(f[0].data_ptr(), f[0].stride(0), 1), (f[1].data_ptr(), f[1].stride(0), 1),
                             (f[2].data_ptr(), f[2].stride(0), 1)]) for f in dev_frames]

        def run(k0, steps_):
            for s in range(steps_):
                for g, (encs, decs, _, tgs, st) in enumerate(groups):
                    fr = [frs[(k0 + (s * nstreams + g) * n + i) % FRAMES] for i in range(n)]
                    pw.encode_batch(encs, fr, [400000] * n, stream=st.cuda_stream)
                    pw.decode_encoded_batch(decs, encs, tgs, stream=st.cuda_stream)

        run(0, warmup)
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(warmup * n * nstreams, steps)
        torch.cuda.synchronize()
        us = (time.perf_counter() - t) / (steps * n * nstreams) * 1e6
        for encs, decs, _, _, _ in groups:
            for x in encs + decs:
                x.close()
        d2.close()
        return round(us, 2)

    if not args.headline_only:
        f1080 = {"fp16_batched": frames_1080p(pw.PRECISION_FP16, B, args.streams),
                 "fp32_single": frames_1080p(pw.PRECISION_FP32, 1, 1, steps=60, warmup=5)}
        f1080["speedup"] = round(f1080["fp32_single"] / f1080["fp16_batched"], 3)
        print(f"[bench] 1080p done", file=sys.stderr, flush=True)

    # One batch of 32 uneven frames (8 each of 1280x720 4:2:0, 1920x1080 4:2:0,
    # 2560x1440 4:4:4, 3840x2160 4:2:0; crops of the resident sequence, the
    # 4:4:4 chroma planes luma crops of neighbouring frames), 1.6 bits/pixel,
    # FP16, tile shapes autotuned on the batch: eager launches vs a replayed
    # hipGraph, µs per batch.
    def uneven_batch(graphs, steps=20, warmup=3):
        shapes = [(1280, 720, pw.CHROMA_420), (1920, 1080, pw.CHROMA_420), (2560, 1440, pw.CHROMA_444),
                  (3840, 2160, pw.CHROMA_420)]
        d2 = pw.Device()
        d2.set_timing(False)
        encs, decs, frs, tgs, budgets, keep = [], [], [], [], [], []
        for si, (w, h, c) in enumerate(shapes):
            for k in range(8):
                encs.append(pw.Encoder(d2, w, h, c, precision=pw.PRECISION_FP16))
                decs.append(pw.Decoder(d2, w, h, c, precision=pw.PRECISION_FP16))
                f = dev_frames[(8 * si + k) % FRAMES]
                if c == pw.CHROMA_444:
                    planes = [dev_frames[(8 * si + k + j) % FRAMES][0][:h, :w] for j in range(3)]
                else:
                    planes = [f[0][:h, :w], f[1][:h // 2, :w // 2], f[2][:h // 2, :w // 2]]
                frs.append(pw.hip_frame([(p.data_ptr(), p.stride(0), 1) for p in planes]))
                ch, cw = (h, w) if c == pw.CHROMA_444 else (h // 2, w // 2)
                outs = [torch.empty((h, w), dtype=torch.uint8, device="cuda")] + \
                       [torch.empty((ch, cw), dtype=torch.uint8, device="cuda") for _ in range(2)]
                keep.append(outs)
                tgs.append(pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs]))
                budgets.append(int(w * h * 1.6 / 8))
        pw.autotune_batch(encs, frs, budgets, decs, tgs, repetitions=5, stream=sptr)
        d2.set_graphs(graphs)
        for _ in range(warmup):
            pw.encode_batch(encs, frs, budgets, stream=sptr)
            pw.decode_encoded_batch(decs, encs, tgs, stream=sptr)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            pw.encode_batch(encs, frs, budgets, stream=sptr)
            pw.decode_encoded_batch(decs, encs, tgs, stream=sptr)
        torch.cuda.synchronize()
        us = (time.perf_counter() - t) / steps * 1e6
        for x in encs + decs:
            x.close()
        d2.close()
        return round(us, 1)

    if not args.headline_only:
        uneven = {"eager": uneven_batch(False), "graph": uneven_batch(True)}
        uneven["frames_per_s_eager"] = round(32 / uneven["eager"] * 1e6, 1)
        print(f"[bench] uneven batch done", file=sys.stderr, flush=True)

    serving = None
    if args.serving_seconds > 0 and not args.headline_only:
        serving = serving_leg(pw, torch, device, dev_frames, args.serving_seconds)
        print(f"[bench] serving done", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(host_frames, W, H, BUDGET, args.cpu_frames, args.cpu_threads)
        print(f"[bench] cpu baseline done", file=sys.stderr, flush=True)

    stage_us = {k: round(v[0] * 1e3 / max(v[1], 1), 2) for k, v in fp16["stages"].items()}
    stage_us32 = {k: round(v[0] * 1e3 / max(v[1], 1), 2) for k, v in fp32["stages"].items()}
    lat = fp16["lat_ms"] * 1e3
    line = {
        "metric": "encode+decode us/frame 4K 4:2:0 (FP16-storage, batched)",
        "value": round(us_fp16, 2),
        "unit": "us/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp16["wall"] / args.steps * 1e3, 3),
        "fp16_one_stream_us_per_frame": round(us_fp16_one, 2),
        "autotuned_tile_shapes": {"forward_levels_0_4": fp16["tuned"][0], "inverse_levels_0_4": fp16["tuned"][1],
                                  "shapes": "forward 0: 128x56, 1: 32x32, 2: 32x16, 3: 128x40; inverse 0: 64x64, "
                                            "1: 32x32, 2: 16x16, 3: 64x48; both 4: 120x32 streamed"},
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3840x2160 4:2:0 60-frame seeded sequence, {BUDGET} B/frame, "
                               f"{args.streams} batches of {B} frames in flight per step (one stream each), "
                               f"encode+decode device-resident",
                   "precision": "fp16 band storage (levels 0-1 high bands), f32 lifting",
                   "batch": B, "streams": args.streams, "budget_bytes": BUDGET},
        "fp32_single_us_per_frame": round(us_fp32, 2),
        "speedup_vs_fp32_single": round(us_fp32 / us_fp16, 3),
        "psnr_y_db": {"fp16": round(psnr["fp16"], 4), "fp32": round(psnr["fp32"], 4)},
        "delta_psnr_db": round(psnr["fp32"] - psnr["fp16"], 4),
        "max_frame_bytes": int(max_bytes),
        "within_budget": bool(max_bytes <= BUDGET),
        "batch_latency_us": {"p50": round(float(np.percentile(lat, 50)), 1),
                             "p99": round(float(np.percentile(lat, 99)), 1)},
        "stage_us_per_frame_fp16": stage_us,
        "stage_us_per_frame_fp32_single": stage_us32,
        "roofline": roofline("iDWT", kernels_of("k_dwt_inverse", fp16["tuned"][1], fused_inv), "inverse",
                             fp16["tuned"][1], fused_inv),
        "roofline_forward": roofline("DWT", kernels_of("k_dwt_forward", fp16["tuned"][0], fused), "forward",
                                     fp16["tuned"][0], fused),
        "fp16_ll_reference_storage": None if fp16_ll is None else {
            "us_per_frame": round(fp16_ll["wall"] / (args.steps * B * args.streams) * 1e6, 2),
            "psnr_y_db": round(psnr["fp16_ll"], 4), "delta_psnr_db": round(psnr["fp32"] - psnr["fp16_ll"], 4),
            "stage_us_per_frame": {k: round(v[0] * 1e3 / max(v[1], 1), 2) for k, v in fp16_ll["stages"].items()}},
        "fp32_batched": None if fp32b is None else {
            "us_per_frame": round(fp32b["wall"] / (args.steps * B * args.streams) * 1e6, 2),
            "psnr_y_db": round(psnr["fp32_batched"], 4),
            "stage_us_per_frame": {k: round(v[0] * 1e3 / max(v[1], 1), 2) for k, v in fp32b["stages"].items()}},
        "fp16_math_reference_precision0": None if fp16m is None else {
            "us_per_frame": round(fp16m["wall"] / (args.steps * B * args.streams) * 1e6, 2),
            "psnr_y_db": round(psnr["fp16_math"], 4), "delta_psnr_db": round(psnr["fp32"] - psnr["fp16_math"], 4),
            "stage_us_per_frame": {k: round(v[0] * 1e3 / max(v[1], 1), 2) for k, v in fp16m["stages"].items()}},
        "frames_1080p_us_per_frame": f1080,
        "single_1080p_fp16_us_per_frame": single,
        "uneven_batch_32_fp16_us_per_batch": uneven,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    for _, r in legs:
        for x in r["encs"] + r["decs"] + r["extra"]:
            x.close()
    device.close()


if __name__ == "__main__":
    main()
