"""SSIMULACRA2 throughput on one MI355X.

Default workload: 240 frame pairs of 1920x1080 YUV420P8 (BT.709, limited
range), synthetic (seeded blob field translating 2 px/frame; distorted by
8x8 DCT quantisation and a 3x3 box blur on half the blocks). A step is one
pass of the hot path over the batch of 240 pairs already resident in HBM:
convert + pyramid + XYB, blur/SSIM/edge maps, FP64 norms and scores.

Prints one JSON line (rank 0). Besides throughput it reports the map
kernel's roofline position (HIP events on the library's stream), the
per-frame path, p50/p99 step latency, and the CPU baseline: the FP32 CPU
restatement timed on the host cores on a bounded sample, plus |dscore| of
the GPU against the FP64 oracle on a sample of the same pairs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (w, h, bits, npairs, fp16)
    "1080p": (1920, 1080, 8, 240, False),
    "4k10": (3840, 2160, 10, 120, True),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="1080p", choices=sorted(CONFIGS))
    ap.add_argument("--map-th", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-frame", action="store_true", help="skip the one-pair-per-launch measurement")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run device-to-device copies of this many MiB (HBM counter calibration)")
    ap.add_argument("--oracle-pairs", type=int, default=3)
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; every box here has one GPU, "
                                   "and the path shards by frame pair with no inter-GPU traffic"}))
        sys.exit(2)
    import vship_amd
    from vship_amd import api, synth

    w, h, bits, npairs, fp16 = CONFIGS[a.config]
    fmt = api.PF_YUV420_U8 if bits == 8 else api.PF_YUV420_U16
    matrix = api.MATRIX_BT709 if bits == 8 else api.MATRIX_BT2020NCL
    t0 = time.time()
    clip = synth.video_clip(w, h, bits, npairs, seed=1 if bits == 8 else 2)
    pairs = [clip.frame(f) for f in range(npairs)]
    t_gen = time.time() - t0
    vship_amd.gpu_full_check(0)
    if a.copy_calibration:
        n = a.copy_calibration << 20
        src_b, dst_b = api.DeviceBuffer(n), api.DeviceBuffer(n)
        for _ in range(3):
            dst_b.copy_from(src_b, n)
        print(f"# calibration: 3 device copies of {n} bytes", flush=True)
        src_b.free()
        dst_b.free()
    dp = api.DevicePairs(pairs, fmt, bits, matrix)
    print(f"# generated {npairs} pairs {w}x{h} in {t_gen:.1f}s, {dp.nbytes / 1e9:.2f} GB on device", flush=True)

    batch = api.SSIMU2Batch(npairs, w, h, fp16=fp16, map_tile_height=a.map_th, persistent_grid=a.grid,
                            profile=True)
    for _ in range(a.warmup):
        scores = batch.compute(dp.descs)
    step_ms, map_ms, map64_ms, prep_ms, fin_ms = [], [], [], [], []
    t_start = time.perf_counter()
    for _ in range(a.steps):
        t1 = time.perf_counter()
        scores = batch.compute(dp.descs)
        step_ms.append((time.perf_counter() - t1) * 1e3)
        st = batch.stats()
        map_ms.append(st["map_fp32_ms"])
        prep_ms.append(st["prep_ms"])
        map64_ms.append(st["map_fp64_ms"])
        fin_ms.append(st["finalize_ms"])
    elapsed = time.perf_counter() - t_start
    st = batch.stats()
    value = npairs * a.steps / elapsed
    map_avg = float(np.mean(map_ms))
    achieved = st["map_bytes_fp32"] / (map_avg * 1e-3) / 1e9

    # per-frame path: one pair per launch sequence, no batching, same device inputs
    per_frame = None
    if not a.no_per_frame:
        single = api.SSIMU2Batch(1, w, h, fp16=fp16)
        nsingle = min(npairs, 48)
        for i in range(3):
            single.compute(dp.descs, 1)
        t1 = time.perf_counter()
        one = (api.Pair * 1)()
        for i in range(nsingle):
            one[0] = dp.descs[i]
            single.compute(one, 1)
        per_frame = nsingle / (time.perf_counter() - t1)
        single.close()

    out = {
        "metric": "SSIMU2 frame-pairs/s",
        "value": round(value, 2),
        "unit": "frame-pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not fp16 else "f16-storage/f32",
        "data": "synthetic",
        "config": {"workload": f"ssimulacra2 {w}x{h} YUV420P{bits} {'BT.709' if bits == 8 else 'BT.2020'} "
                               f"limited, {npairs} pairs per step, inputs resident in HBM",
                   "pairs_per_step": npairs, "map_tile_height": st["map_tile_height"]},
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 3),
                       "p99": round(float(np.percentile(step_ms, 99)), 3),
                       "per": f"step of {npairs} pairs"},
        "kernels_ms": {"prep": round(float(np.mean(prep_ms)), 3), "map_fp32": round(map_avg, 3),
                       "map_fp64_small_segments": round(float(np.mean(map64_ms)), 3),
                       "finalize": round(float(np.mean(fin_ms)), 3)},
        "per_frame_path": {"value": round(per_frame, 2) if per_frame else None, "unit": "frame-pairs/s",
                           "note": "one pair per launch sequence, eager, same device inputs"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"map_kernel<{st['map_tile_height']},float,float>", "algorithmic_bytes_per_launch": st["map_bytes_fp32"],
                     "note": "map kernel is FP32-VALU bound; algorithmic bytes = 2 images x 3 planes x "
                             f"{2 if fp16 else 4} B x sum of scale pixels"},
    }

    if not a.no_cpu:
        out["cpu_baseline"], out["score_vs_fp64_oracle"] = cpu_baseline(pairs, scores, bits, a)
    print(json.dumps(out), flush=True)
    batch.close()
    dp.free()


def cpu_baseline(pairs, gpu_scores, bits, a):
    """FP32 CPU restatement timed on the host cores (bounded sample) and the FP64
    oracle on a few pairs for |dscore|."""
    from tests import oracle_lib as ol

    threads = min(os.cpu_count() or 1, 16)
    ol.set_threads(threads)
    matrix = ol.SO_MAT_BT709 if bits == 8 else ol.SO_MAT_BT2020NCL
    n, t0 = 0, time.perf_counter()
    while n < len(pairs) and time.perf_counter() - t0 < a.cpu_seconds:
        r, d = pairs[n]
        ol.score_yuv420(r, d, bits, matrix, "f32")
        n += 1
    cpu_rate = n / (time.perf_counter() - t0)
    idx = np.linspace(0, len(pairs) - 1, a.oracle_pairs).astype(int)
    diffs = []
    for i in idx:
        r, d = pairs[i]
        s64, _ = ol.score_yuv420(r, d, bits, matrix, "f64")
        diffs.append(abs(float(gpu_scores[i]) - s64))
    base = {"value": round(cpu_rate, 4), "unit": "frame-pairs/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pairs of the workload, FP32 CPU restatement (OpenMP over rows), "
                      f"{a.cpu_seconds:.0f}s budget"}
    delta = {"pairs": [int(i) for i in idx], "max_abs": float(np.max(diffs)), "mean_abs": float(np.mean(diffs))}
    return base, delta


if __name__ == "__main__":
    main()
