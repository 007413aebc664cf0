"""SSIMULACRA2 throughput on one MI355X.

Workloads (--config), all synthetic (seeded blob field translating 2 px per
frame; distorted by 8x8 DCT quantisation and a 3x3 box blur on half the
blocks; vship_amd/synth/synth.c holds the generator):

  1080p   (default) 240 pairs 1920x1080 YUV420P8 BT.709 limited, FP32. A step
          is one pass of the hot path over the 240 pairs already resident in
          HBM: convert + pyramid + XYB, blur/SSIM/edge maps, FP64 scores.
  rgb256  one 256x256 planar RGB8 full-range pair (the CPU-runnable case).
  4k10    120 pairs 3840x2160 YUV420P10 BT.2020 limited, XYB scales 0-1
          stored in FP16, scales 2-5 in FP32 (FP32 math, FP64 sums);
          per-frame |dscore| vs the FP64 oracle.
  mixed   64 pairs per step, resolutions drawn by seed from 640x360..3840x2160,
          YUV420P8; launch shape from the autotuner, graph-captured launch.
  served  Poisson clip arrivals (200 clips/s), 8..64 pairs per clip at the
          mixed resolutions, 8/10-bit 50/50, continuous-batching server under a
          200 GB budget; frames/s and p50/p99 clip latency.
  butter1080p  Butteraugli (not SSIMULACRA2): 16 pairs 1920x1080 YUV420P8
          BT.709 limited per step, resident in HBM; norms Q=2, 3, inf at
          203 nits; stage times, the dominant kernel's roofline position and
          the FP64 CPU oracle as baseline.

Prints one JSON line. Besides throughput: the dominant kernel's roofline
position (HIP events on the library's stream; vector-ALU issue from the
committed counter record over the live duration, with the HBM position
beside it), the per-frame path, p50/p99 latency, and the CPU baseline (the
FP32 CPU restatement timed on the host's CPU share on a bounded sample, and
on one thread, with the host's CPU count and model) with |dscore| against the
FP64 oracle on a sample.

The default 1080p line also carries `sub_records` for the 4k10, mixed and
served workloads (each with its own throughput, latency, kernel times and
|dscore| against the FP64 oracle on sampled pairs).
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CONFIGS = ["1080p", "rgb256", "4k10", "mixed", "served", "butter1080p"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="1080p", choices=CONFIGS)
    ap.add_argument("--map-th", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-frame", action="store_true", help="skip the one-pair-per-launch measurement")
    ap.add_argument("--oracle-pairs", type=int, default=3)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run device-to-device copies of this many MiB (HBM counter calibration)")
    ap.add_argument("--arrival-seconds", type=float, default=2.0, help="served: length of the arrival window")
    ap.add_argument("--rate", type=float, default=200.0, help="served: clip arrival rate (clips/s)")
    ap.add_argument("--no-sub", action="store_true",
                    help="1080p: skip the 4K, mixed, served and host-input sub-records")
    ap.add_argument("--sub-steps", type=int, default=20, help="timed steps of each sub-record")
    return ap.parse_args()


def fmt_of(bits):
    from vship_amd import api

    if bits == 8:
        return api.PF_YUV420_U8, api.MATRIX_BT709
    return api.PF_YUV420_U16, api.MATRIX_BT2020NCL


def video_pairs(w, h, bits, n, seed):
    from vship_amd import synth

    clip = synth.video_clip(w, h, bits, n, seed)
    pairs = [clip.frame(f) for f in range(n)]
    clip.close()
    return pairs


def copy_calibration(mib):
    from vship_amd import api

    n = mib << 20
    src_b, dst_b = api.DeviceBuffer(n), api.DeviceBuffer(n)
    for _ in range(3):
        dst_b.copy_from(src_b, n)
    print(f"# calibration: 3 device copies of {n} bytes", flush=True)
    src_b.free()
    dst_b.free()


def timed_steps(batch, descs, n, steps, warmup):
    """`steps` passes over the batch with two in flight: step i + 1 is
    submitted before step i is waited for (the engine holds two outstanding
    batches, in stream order), so that consecutive steps run back to back on
    the GPU. Per-kernel times are each step's HIP events. Then the
    submit-to-result latency of single steps, one batch in flight, over
    min(steps, 50) more steps."""
    scores = None
    for _ in range(warmup):
        scores = batch.compute(descs, n)
    k = {"prep": [], "map": [], "map64": [], "fin": []}
    t_start = time.perf_counter()
    batch.submit(descs, n)
    for i in range(steps):
        if i + 1 < steps:
            batch.submit(descs, n)
        scores = batch.wait()
        st = batch.stats()
        k["prep"].append(st["prep_ms"])
        k["map"].append(st["map_fp32_ms"])
        k["map64"].append(st["map_fp64_ms"])
        k["fin"].append(st["finalize_ms"])
    elapsed = time.perf_counter() - t_start
    step_ms = []
    for _ in range(min(steps, 50)):
        t1 = time.perf_counter()
        one = batch.compute(descs, n)
        step_ms.append((time.perf_counter() - t1) * 1e3)
        if not np.array_equal(one, scores):
            raise RuntimeError("a step alone scored differently from the pipelined steps")
    return scores, elapsed, step_ms, k, batch.stats()


def per_frame_rate(descs, n, w, h, fp16):
    """One pair per launch sequence, eager, same device inputs."""
    from vship_amd import api

    single = api.SSIMU2Batch(1, w, h, fp16=fp16)
    nsingle = min(n, 48)
    one = (api.Pair * 1)()
    for i in range(3):
        one[0] = descs[i % n]
        single.compute(one, 1)
    t1 = time.perf_counter()
    for i in range(nsingle):
        one[0] = descs[i]
        single.compute(one, 1)
    rate = nsingle / (time.perf_counter() - t1)
    single.close()
    return rate


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_issue.json")
# MI355X vector ALUs: 256 CUs x 4 SIMDs x 16 lanes at the 2.4 GHz peak engine
# clock; a wave64 VALU instruction is 64 lane-ops (4 cycles of its SIMD)
VALU_SIMDS, VALU_CLOCK_HZ = 256 * 4, 2.4e9
VALU_PEAK_TOPS = VALU_SIMDS * 16 * VALU_CLOCK_HZ / 1e12


def _record(path, config, kernel):
    try:
        with open(path) as f:
            return json.load(f)["workloads"][config][kernel]
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic(config, kernel):
    """HBM bytes per launch of `kernel` on workload `config` from the committed
    record of the FETCH_SIZE/WRITE_SIZE runs (calibrated on a device copy), or
    None where no run of this kernel on this workload is recorded."""
    rec = _record(TRAFFIC_FILE, config, kernel)
    return None if rec is None else rec["traffic_bytes"]


def valu_position(config, kernel, ms):
    """Vector-ALU issue of one launch of `kernel`: its lane-ops (from the
    committed SQ_INSTS_VALU / SQ_INSTS_VALU_MFMA_F16 record of this workload;
    every VALU instruction, transcendentals included, counted once) over the
    launch's live duration, against the peak issue rate. None where no counter
    run of this kernel on this workload exists."""
    rec = _record(VALU_FILE, config, kernel)
    if rec is None:
        return None
    lane_ops = (rec["valu_insts"] - rec["mfma_insts"]) * 64
    achieved = lane_ops / (ms * 1e-3) / 1e12
    out = {"achieved": round(achieved, 2), "peak": round(VALU_PEAK_TOPS, 2), "unit": "T VALU lane-ops/s",
           "frac": round(achieved / VALU_PEAK_TOPS, 4), "lane_ops_per_launch": int(lane_ops),
           "transcendental_share": round(rec["trans_insts"] / max(1, rec["valu_insts"] - rec["mfma_insts"]), 4),
           "issue_of_active_cycles": rec.get("valu_issue_of_active_cycles")}
    if "valu_quad_cycles" in rec:
        # quad-cycles the VALU instructions occupy (half-rate ones twice), x 4
        # cycles, over every SIMD's cycles at the peak clock for the live time
        out["issue_cycles_frac"] = round(rec["valu_quad_cycles"] * 4 / (VALU_SIMDS * VALU_CLOCK_HZ * ms * 1e-3), 4)
        out["issue_cycles_of_active_cycles"] = rec.get("valu_cycles_of_active_cycles")
    return out


def map_kernel_name(st, fp16):
    # with FP16 storage the FP32 class is two launches, one per plane group
    # (FP16 scales 0-1, FP32 scales 2-5); the time and the counters cover both
    st_name = "__half+float" if fp16 else "float"
    if st["map_tile_height"] == 1:
        kernel = f"map_mx_kernel<{st_name}, SplitF16x2>"
        return kernel, f"{kernel} (matrix cores, tiles of 64 columns x 128 rows)"
    kernel = f"map_kernel<{st['map_tile_height']}, {st_name}, float>"
    return kernel, kernel


def roofline(st, map_ms, prep_ms, fp16, config):
    """The dominant kernel's position. The map kernel is bound by vector-ALU
    issue: `bound` is "valu" when its counter record exists, with the HBM
    position (algorithmic bytes over the live duration) beside it."""
    achieved = st["map_bytes_fp32"] / (map_ms * 1e-3) / 1e9
    elem = "2 B (scales 0-1) or 4 B (scales 2-5)" if fp16 else "4 B"
    kernel, label = map_kernel_name(st, fp16)
    traffic = measured_traffic(config, kernel)
    hbm = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4)}
    valu = valu_position(config, kernel, map_ms)
    out = {"bound": "valu" if valu else "hbm"}
    out.update(valu if valu else hbm)
    out.update({"traffic": traffic,
                "traffic_source": ("profiles/hbm_traffic.json: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch"
                                   if traffic is not None else "no PMC run of this kernel on this workload"),
                "kernel": label, "hbm": hbm,
                "algorithmic_bytes_per_launch": st["map_bytes_fp32"],
                "note": f"2 images x 3 XYB planes x {elem} read per pixel of every scale in the FP32 class; "
                        "the kernel is bound by vector-ALU issue (operand splitting and the maps), not HBM"})
    if valu:
        out["valu_source"] = "profiles/valu_issue.json: rocprofv3 SQ_INSTS_VALU minus SQ_INSTS_VALU_MFMA_F16 per " \
                             "launch over the live HIP-event duration"
    # the other large kernel of the step
    prep_kernel = f"prep_kernel<{'__half' if fp16 else 'float'}, false>"
    pv = valu_position(config, prep_kernel, prep_ms)
    if pv:
        out["prep"] = {"kernel": prep_kernel, "valu": pv}
    return out


def record(a, value, elapsed, steps, dtype, workload, extra_cfg):
    return {
        "metric": "SSIMU2 frame-pairs/s",
        "value": round(value, 2),
        "unit": "frame-pairs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed * 1e3 / max(steps, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": dict({"workload": workload}, **extra_cfg),
    }


def run_batched(a, pairs, bits, fp16, workload, tune=False, graph=False, steps=None, config=None, per_frame=True):
    """Times `steps` (default --steps) passes over the device-resident pairs;
    returns (line, scores)."""
    from vship_amd import api

    steps = a.steps if steps is None else steps
    config = a.config if config is None else config

    npairs = len(pairs)
    w = max(p[0][0].shape[1] for p in pairs)
    h = max(p[0][0].shape[0] for p in pairs)
    if bits == 0:  # planar RGB8 full range
        dp = api.DevicePairs(pairs, api.PF_RGB_U8, 8, 0)
    else:
        fmt, matrix = fmt_of(bits)
        dp = api.DevicePairs(pairs, fmt, bits, matrix)
    print(f"# {npairs} pairs up to {w}x{h}, {dp.nbytes / 1e9:.2f} GB of inputs on device", flush=True)
    th, grid, tuned = a.map_th, a.grid, None
    if tune:
        tuned = api.autotune(dp.descs, npairs, fp16=fp16, repeats=3)
        th, grid, graph = tuned["map_tile_height"], tuned["persistent_grid"], bool(tuned["use_graph"])
        print("# autotune best: " + json.dumps({k: tuned[k] for k in
                                                ("map_tile_height", "persistent_grid", "use_graph", "best_ms")}),
              flush=True)
    # eager handle with per-kernel HIP events
    batch = api.SSIMU2Batch(npairs, w, h, fp16=fp16, map_tile_height=th, persistent_grid=grid, profile=True)
    scores, elapsed, step_ms, k, st = timed_steps(batch, dp.descs, npairs, steps, a.warmup)
    value, graph_used = npairs * steps / elapsed, False
    out_extra = {}
    if graph:
        # the same shape captured once in a hipGraph and replayed each step
        gb = api.SSIMU2Batch(npairs, w, h, fp16=fp16, map_tile_height=th, persistent_grid=grid, use_graph=True)
        gscores, gel, gstep, _, _ = timed_steps(gb, dp.descs, npairs, steps, a.warmup)
        gb.close()
        if not np.array_equal(gscores, scores):
            raise RuntimeError("graph replay scores differ from eager launches")
        gval = npairs * steps / gel
        out_extra["eager_value"] = round(value, 2)
        out_extra["graph_value"] = round(gval, 2)
        value, elapsed, step_ms, graph_used = gval, gel, gstep, True
    map_ms = float(np.mean(k["map"]))
    out = record(a, value, elapsed, steps, "f16-storage/f32" if fp16 else "f32", workload,
                 {"pairs_per_step": npairs, "map_tile_height": st["map_tile_height"],
                  "persistent_grid": grid, "hip_graph": graph_used, "steps_in_flight": 2})
    out.update(out_extra)
    out["latency_ms"] = {"p50": round(float(np.percentile(step_ms, 50)), 3),
                         "p99": round(float(np.percentile(step_ms, 99)), 3),
                         "per": f"step of {npairs} pairs, submit to result, one step in flight"}
    out["kernels_ms"] = {"prep": round(float(np.mean(k["prep"])), 3), "map_fp32": round(map_ms, 3),
                         "map_fp64_small_segments": round(float(np.mean(k["map64"])), 3),
                         "finalize": round(float(np.mean(k["fin"])), 3)}
    if tuned is not None:
        out["autotune"] = tuned
    if per_frame and not a.no_per_frame:
        pf = per_frame_rate(dp.descs, npairs, w, h, fp16)
        out["per_frame_path"] = {"value": round(pf, 2), "unit": "frame-pairs/s",
                                 "note": "one pair per launch sequence, eager, same device inputs",
                                 "batched_over_per_frame": round(value / pf, 2)}
    out["roofline"] = roofline(st, map_ms, float(np.mean(k["prep"])), fp16, config)
    batch.close()
    dp.free()
    return out, scores


def run_served(a):
    """Poisson clip stream through the continuous-batching server.
    Returns (line, host pairs of the frame pools, their scores alone)."""
    from vship_amd import api, synth

    rng = np.random.default_rng(4)
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / a.rate)
        if t > a.arrival_seconds:
            break
        w, h = synth.RESOLUTIONS[int(rng.integers(0, len(synth.RESOLUTIONS)))]
        bits = 8 if rng.random() < 0.5 else 10
        arrivals.append((t, (w, h, bits), int(rng.integers(8, 65))))
    # device-resident pools of 8 distinct pairs per (resolution, depth); a clip
    # of n pairs cycles through its pool
    pools, host_pool = {}, {}
    for (w, h) in synth.RESOLUTIONS:
        for bits in (8, 10):
            fmt, matrix = fmt_of(bits)
            pp = video_pairs(w, h, bits, 8, seed=w * 7 + bits)
            host_pool[(w, h, bits)] = pp
            pools[(w, h, bits)] = api.DevicePairs(pp, fmt, bits, matrix)
    clips = []
    for (t, key, n) in arrivals:
        pool = pools[key]
        d = (api.Pair * n)()
        for i in range(n):
            d[i] = pool.descs[i % pool.n]
        clips.append((t, key, d, n))
    total_pairs = sum(c[3] for c in clips)
    print(f"# {len(clips)} clips, {total_pairs} pairs over {a.arrival_seconds:.1f}s of arrivals", flush=True)
    budget = 200 * 10 ** 9
    srv = api.SSIMU2Server(budget, 64, 3840, 2160, max_inflight_clips=64)
    # warm-up clip (first launches, code-object load) before the clock starts
    srv.wait(srv.submit(clips[0][2], clips[0][3]))
    ids = [None] * len(clips)
    t0 = srv.clock_ms()

    def feeder():
        for k, (t, _, d, n) in enumerate(clips):
            target = t0 + t * 1e3
            now = srv.clock_ms()
            if target > now:
                time.sleep((target - now) / 1e3)
            ids[k] = srv.submit(d, n, arrival_ms=target)

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    lat, results = [], []
    for cid in ids:
        s, l = srv.wait(cid)
        lat.append(l)
        results.append(s)
    t_end = srv.clock_ms()
    st = srv.stats()
    srv.close()
    elapsed = (t_end - t0) / 1e3
    value = total_pairs / elapsed
    # each clip's scores equal its pool's pairs scored alone
    alone = {}
    for key, pool in pools.items():
        with api.SSIMU2Batch(pool.n, key[0], key[1]) as b:
            alone[key] = b.compute(pool.descs, pool.n)
    same = sum(int(np.array_equal(np.array([alone[key][i % 8] for i in range(n)]), results[k]))
               for k, (_, key, _, n) in enumerate(clips))
    out = record(a, value, elapsed, 1, "f32",
                 "served ssimulacra2 clips: Poisson arrivals, 8-64 pairs per clip, 640x360..3840x2160 "
                 "YUV420 8/10-bit, inputs resident in HBM",
                 {"rate_clips_per_s": a.rate, "arrival_seconds": a.arrival_seconds, "clips": len(clips),
                  "pairs": total_pairs, "memory_budget_bytes": budget, "max_pairs_per_step": 64})
    out["steps"] = int(st["steps"])
    out["ms_per_step"] = round(elapsed * 1e3 / max(1, int(st["steps"])), 3)
    out["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 2),
                         "p99": round(float(np.percentile(lat, 99)), 2), "per": "clip, arrival to last score"}
    out["server"] = {k: st[k] for k in ("pairs_done", "steps", "clips_done")}
    out["clips_equal_alone"] = f"{same}/{len(clips)}"
    for p in pools.values():
        p.free()
    return out, host_pool, alone


BUTTER_KERNELS = {"input": "butter_input420_kernel<false>", "opsin": "butter_opsin_kernel", "low": "butter_low_kernel",
                  "mid": "butter_mid_kernel", "high": "butter_high_kernel", "diff": "butter_diff_kernel",
                  "mask": "butter_mask_kernel", "norm": "butter_norm_kernel"}
BUTTER_BYTES = {"input": "Y + 2 quarter-size chroma bytes read per image, 6 linear FP32 planes written at "
                         "full and quarter size",
                "opsin": "6 FP32 planes read, 6 written, per pixel of both scales",
                "low": "6 FP32 planes read, 12 written (low band as values, residual)",
                "mid": "6 read, 10 written (middle band, high residual)",
                "high": "4 read, 8 written (high, ultra-high bands)",
                "diff": "20 read (bands of both images), 4 written (AC, DC sums, 2 mask inputs)",
                "mask": "4 read, 1 written, plus the half-scale map read at full scale",
                "norm": "tile partials"}


def run_butter(a):
    """Butteraugli: 16 1920x1080 YUV420P8 pairs per step, device-resident."""
    from vship_amd import api

    w, h, npairs = 1920, 1080, 16
    pairs = video_pairs(w, h, 8, npairs, seed=4)
    dp = api.ButterDevicePairs(pairs, fmt=api.PF_YUV420_U8, bits=8, matrix=api.MATRIX_BT709)
    print(f"# butteraugli: {npairs} pairs {w}x{h} YUV420P8 on device, "
          f"{api.butter_pair_footprint(w, h) * npairs / 1e9:.1f} GB of work planes", flush=True)
    batch = api.ButteraugliBatch(w, h, npairs)
    norms = None
    for _ in range(a.warmup):
        norms = batch.compute(dp)
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t1 = time.perf_counter()
        norms = batch.compute(dp)
        step_ms.append((time.perf_counter() - t1) * 1e3)
    elapsed = time.perf_counter() - t0
    batch.close()
    # stage times: the same batch launched stage by stage between HIP events
    prof = api.ButteraugliBatch(w, h, npairs, profile=True)
    stages = {k: [] for k in api.BUTTER_STAGES}
    for _ in range(5):
        pn = prof.compute(dp)
        st = prof.stats()
        for k in stages:
            stages[k].append(st[k + "_ms"])
    if not np.array_equal(pn, norms):
        raise RuntimeError("stage-ordered launches give other norms than the fused order")
    stage_ms = {k: float(np.mean(v[1:])) for k, v in stages.items()}
    stage_bytes = st["stage_bytes"]
    prof.close()
    value = npairs * a.steps / elapsed
    out = record(a, value, elapsed, a.steps, "f32", "butteraugli 1920x1080 YUV420P8 BT.709 limited, 16 pairs per "
                 "step, inputs resident in HBM, qnorm 2, 203 nits", {"pairs_per_step": npairs})
    out["metric"] = "Butteraugli frame-pairs/s"
    out["latency_ms"] = {"p50": round(float(np.percentile(step_ms, 50)), 3),
                         "p99": round(float(np.percentile(step_ms, 99)), 3), "per": f"step of {npairs} pairs"}
    out["kernels_ms"] = {k: round(v, 3) for k, v in stage_ms.items()}
    top = max((k for k in stage_ms if k != "norm"), key=lambda k: stage_ms[k])
    achieved = stage_bytes[top] / (stage_ms[top] * 1e-3) / 1e9
    kernel = BUTTER_KERNELS[top]
    traffic = measured_traffic(a.config, kernel)
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                       "traffic_source": ("profiles/hbm_traffic.json: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 "
                                          "per launch" if traffic is not None
                                          else "no PMC run of this kernel on this workload"),
                       "kernel": f"{kernel} (both scales, {top} stage)",
                       "algorithmic_bytes_per_launch": int(stage_bytes[top]),
                       "note": BUTTER_BYTES[top] + "; FP32 stencil sums on the vector ALUs"}
    out["stage_bytes"] = {k: int(v) for k, v in stage_bytes.items()}
    return out, pairs, norms


def butter_cpu_baseline(pairs, norms, a):
    """The FP64 C oracle (single thread) on the first pairs within the budget,
    and the GPU norms of those pairs against it."""
    from tests import butter_lib as bl
    from tests import oracle_lib as ol

    n, rel, t0 = 0, [], time.perf_counter()
    while n < len(pairs) and (n == 0 or time.perf_counter() - t0 < a.cpu_seconds):
        r, d = pairs[n]
        rgb = [ol.yuv_to_rgb_any(im, 8, ol.SO_MAT_BT709, 1, 1, 0).astype(np.float32) for im in (r, d)]
        want = bl.butteraugli(list(rgb[0]), list(rgb[1]))
        rel.append(float(np.max(np.abs(norms[n] - want) / want)))
        n += 1
    rate = n / (time.perf_counter() - t0)
    base = {"value": round(rate, 4), "unit": "frame-pairs/s", "cores": 1, "kind": "port",
            "sample": f"first {n} pair(s) of the workload through the FP64 C oracle (YUV to R'G'B' and the "
                      f"metric), single thread, {a.cpu_seconds:.0f}s budget"}
    return base, {"pairs": list(range(n)), "max_rel_norm_diff": max(rel)}


def host_cpu():
    """The host the CPU baseline ran on: logical CPUs and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def _score_fn(pairs, bits):
    from tests import oracle_lib as ol

    bl = bits if isinstance(bits, list) else [bits] * len(pairs)

    def score(i, prec):
        r, d = pairs[i]
        if bl[i] == 0:
            return ol.score_rgb(ol.SO_RGB_U8, r, d, prec)[0]
        m = ol.SO_MAT_BT709 if bl[i] == 8 else ol.SO_MAT_BT2020NCL
        return ol.score_yuv420(r, d, bl[i], m, prec)[0]

    return score


def oracle_delta(pairs, bits, gpu_scores, idx, threads):
    """|GPU score - FP64 oracle score| on the pairs `idx`."""
    from tests import oracle_lib as ol

    ol.set_threads(threads)
    score = _score_fn(pairs, bits)
    diffs = [abs(float(gpu_scores[i]) - score(i, "f64")) for i in idx]
    return {"pairs": [int(i) for i in idx], "abs": [float(x) for x in diffs], "max_abs": float(np.max(diffs)),
            "mean_abs": float(np.mean(diffs))}


def cpu_threads():
    # the CPU share this process is given (16 threads per GPU on the MI355X
    # boxes, set through OMP_NUM_THREADS), else every logical CPU
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def cpu_baseline(pairs, bits, gpu_scores, a):
    """FP32 CPU restatement timed on the host's CPU share (bounded sample):
    one pair on one thread first, then pairs scored concurrently, one per
    thread (frame pairs are independent); and the FP64 oracle on a few pairs
    for |dscore|. bits: 0 for RGB8, else per-pair list or int."""
    import concurrent.futures as cf

    from tests import oracle_lib as ol

    threads = cpu_threads()
    score = _score_fn(pairs, bits)
    ol.set_threads(1)
    t1 = time.perf_counter()
    score(0, "f32")
    one_rate = 1.0 / (time.perf_counter() - t1)
    # enough pairs for about cpu_seconds of work on every thread, at least one per thread
    n = int(min(len(pairs), max(threads, round(one_rate * threads * a.cpu_seconds))))
    def one(i):
        ol.set_threads(1)  # per-thread OpenMP setting: no nested teams
        return score(i, "f32")

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, range(n)))
    cpu_rate = n / (time.perf_counter() - t0)
    idx = sorted(set(np.linspace(0, len(pairs) - 1, a.oracle_pairs).astype(int).tolist()))
    delta = oracle_delta(pairs, bits, gpu_scores, idx, threads)
    base = {"value": round(cpu_rate, 4), "unit": "frame-pairs/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pairs of the workload, FP32 CPU restatement, one pair per thread on "
                      f"{threads} threads",
            "single_thread": {"value": round(one_rate, 4), "unit": "frame-pairs/s", "sample": "pair 0, 1 thread"},
            "host": dict(host_cpu(), threads_used=threads)}
    return base, delta


def sub_4k10(a):
    """3840x2160 YUV420P10 BT.2020, FP16 planes for scales 0-1, 120 pairs."""
    pairs = video_pairs(3840, 2160, 10, 120, seed=2)
    out, scores = run_batched(a, pairs, 10, True, "ssimulacra2 3840x2160 YUV420P10 BT.2020 limited, FP16 planes "
                              "for scales 0-1, 120 pairs per step, inputs resident in HBM", steps=a.sub_steps,
                              config="4k10")
    out["score_vs_fp64_oracle"] = oracle_delta(pairs, 10, scores, [0, 119], cpu_threads())
    return out


def mixed_pairs():
    from vship_amd import synth

    res = synth.batch_resolutions(64, seed=3)
    clips, pairs = {}, []
    for i, (w, h) in enumerate(res):
        if (w, h) not in clips:
            clips[(w, h)] = synth.video_clip(w, h, 8, 64, seed=w + h)
        pairs.append(clips[(w, h)].frame(i))
    for c in clips.values():
        c.close()
    return res, pairs


def sub_mixed(a, steps=None, per_frame=False):
    res, pairs = mixed_pairs()
    out, scores = run_batched(a, pairs, 8, False, "ssimulacra2 uneven batch of 64 pairs, 640x360..3840x2160 "
                              "YUV420P8 BT.709 limited, inputs resident in HBM", tune=True, graph=True,
                              steps=steps, config="mixed", per_frame=per_frame)
    out["config"]["resolutions"] = {f"{w}x{h}": res.count((w, h)) for (w, h) in sorted(set(res))}
    idx = sorted({0, res.index((3840, 2160)), 63})
    out["score_vs_fp64_oracle"] = oracle_delta(pairs, 8, scores, idx, cpu_threads())
    return out, pairs, scores


def sub_served(a):
    out, host_pool, alone = run_served(a)
    pairs, scores, bits = [], [], []
    for key, pp in host_pool.items():
        pairs += pp[:1]
        scores.append(alone[key][0])
        bits.append(key[2])
    # the first pair of the 1920x1080 8-bit, 3840x2160 10-bit and 640x360 10-bit pools
    keys = list(host_pool)
    idx = [keys.index(k) for k in ((1920, 1080, 8), (3840, 2160, 10), (640, 360, 10)) if k in keys]
    out["score_vs_fp64_oracle"] = oracle_delta(pairs, bits, scores, idx, cpu_threads())
    out["score_vs_fp64_oracle"]["note"] = "first pair of the 1080p 8-bit, 4K 10-bit and 640x360 10-bit pools, " \
                                          "scores equal to the served ones (clips_equal_alone)"
    return out


def h2d_copy_rate(nbytes):
    """GB/s of one synchronous page-locked host-to-device copy of nbytes (best of 3)."""
    from vship_amd import api

    src, dst = api.PinnedBuffer(nbytes), api.DeviceBuffer(nbytes)
    src.array[:] = 1
    best = 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        api._check(api.lib().Vship_MemcpyHtoD(dst.ptr, src.ptr, nbytes), "Vship_MemcpyHtoD")
        best = max(best, nbytes / (time.perf_counter() - t0) / 1e9)
    src.free()
    dst.free()
    return best


def sub_host_input(a, pairs, want):
    """The 1080p workload with its inputs in HOST memory: every step uploads
    the 240 pairs through Vship_SSIMU2BatchSubmitHost, two steps in flight,
    so that step k+1's upload runs while step k's kernels do. Page-locked
    planes go by DMA directly; pageable ones through the staging ring."""
    from vship_amd import api

    n, steps = len(pairs), a.sub_steps
    fmt, matrix = fmt_of(8)
    out = {"metric": "ssimulacra2 frame-pairs/s, host input", "unit": "frame-pairs/s",
           "config": {"workload": "ssimulacra2 1920x1080 YUV420P8 BT.709 limited, 240 pairs per step, planes in "
                                  "host memory, uploaded every step (host-to-device copies inside the timed region)"}}
    for kind in ("page_locked", "pageable"):
        hp = api.HostPairs(pairs, fmt, 8, matrix, pinned=kind == "page_locked")
        b = api.SSIMU2Batch(n, 1920, 1080, map_tile_height=a.map_th, persistent_grid=a.grid)
        for _ in range(max(1, a.warmup)):
            b.submit_host(hp.descs)
            got = b.wait()
        if not np.array_equal(got, want):
            raise RuntimeError(f"host-input ({kind}) scores differ from the device-input ones")
        t0 = time.perf_counter()
        b.submit_host(hp.descs)
        for k in range(steps):
            if k + 1 < steps:
                b.submit_host(hp.descs)
            got = b.wait()
        el = time.perf_counter() - t0
        if not np.array_equal(got, want):
            raise RuntimeError(f"host-input ({kind}) scores differ from the device-input ones")
        out[kind] = {"value": round(n * steps / el, 2), "ms_per_step": round(el / steps * 1e3, 3),
                     "upload_GBps": round(hp.nbytes * steps / el / 1e9, 2), "steps": steps}
        if kind == "pageable":
            out[kind]["copy_threads"] = 4
        b.close()
        hp.free()
    nbytes = sum(p.nbytes for r, d in pairs for p in list(r) + list(d))
    rate = h2d_copy_rate(nbytes)
    out["input_bytes_per_step"] = nbytes
    out["h2d_copy_GBps"] = round(rate, 2)
    out["h2d_bound_pairs_per_s"] = round(n / (nbytes / (rate * 1e9)), 2)
    out["value"] = out["page_locked"]["value"]
    out["note"] = ("value: page-locked planes; h2d_bound_pairs_per_s: the rate a synchronous page-locked copy of "
                   "the same bytes allows; scores equal to the device-input ones bit for bit")
    return out


def rgb_float_pairs(w, h, n, seed):
    """Planar float RGB pairs in [0, 1] from the synthetic YUV420P8 clip
    (BT.709 limited range, chroma repeated 2x2)."""
    from vship_amd import synth

    clip = synth.video_clip(w, h, 8, n, seed)
    out = []
    for f in range(n):
        imgs = []
        for yuv in clip.frame(f):
            y = (yuv[0].astype(np.float32) - 16) / 219
            u = (np.repeat(np.repeat(yuv[1], 2, 0), 2, 1)[:h, :w].astype(np.float32) - 128) / 224
            v = (np.repeat(np.repeat(yuv[2], 2, 0), 2, 1)[:h, :w].astype(np.float32) - 128) / 224
            r = y + 1.5748 * v
            g = y - 0.1873 * u - 0.4681 * v
            b = y + 1.8556 * u
            imgs.append([np.ascontiguousarray(np.clip(c, 0, 1), np.float32) for c in (r, g, b)])
        out.append(tuple(imgs))
    clip.close()
    return out


def sub_vs_filter(a):
    """vship.SSIMULACRA2 through the mock VapourSynth host: 48 frames of
    1920x1080 float RGB in pageable memory (16 distinct pairs, each requested
    three times), 8 request threads, 1/3/8 engines (numStream)."""
    from vship_amd import api
    from tests import vs_mock

    if not os.path.exists(vs_mock.MOCK):
        return {"value": None, "note": "mock host not built"}
    w, h, uniq, reps = 1920, 1080, 16, 3
    pairs = rgb_float_pairs(w, h, uniq, seed=9)
    ref = [pairs[i % uniq][0] for i in range(uniq * reps)]
    dist = [pairs[i % uniq][1] for i in range(uniq * reps)]
    with api.SSIMU2Handler(w, h) as hd:
        direct = np.array([hd.compute_float(r, d) for r, d in pairs])
    per = {}
    for ns in (1, 3, 8):
        rc, scores, matches, err, ms = vs_mock.run_ssimu2(ref, dist, w, h, w * 4, num_stream=ns, gpu_id=0,
                                                          core_threads=8, nthreads=8, timed=True)
        if rc != 0 or matches != uniq * reps or not np.array_equal(scores, np.tile(direct, reps)):
            raise RuntimeError(f"vs filter numStream {ns}: rc {rc} {err}")
        per[str(ns)] = round(uniq * reps / ms * 1e3, 2)
    best = max(per, key=lambda k: per[k])
    return {"metric": "ssimulacra2 frame-pairs/s through vship.SSIMULACRA2", "unit": "frame-pairs/s",
            "value": per[best], "num_stream": int(best), "per_num_stream": per,
            "config": {"workload": f"{uniq * reps} frames {w}x{h} float RGB (RGBS) in pageable host memory, "
                                   "8 request threads, mock VapourSynth host"},
            "upload_GBps": round(per[best] * w * h * 4 * 6 / 1e9, 2),
            "note": "frame requests only (filter creation excluded); scores equal to Vship_ComputeSSIMU2Floatv2's"}


def brief(rec):
    keep = ("metric", "value", "unit", "steps", "ms_per_step", "dtype", "config", "latency_ms", "kernels_ms",
            "per_frame_path", "roofline", "score_vs_fp64_oracle", "clips_equal_alone", "server", "eager_value",
            "graph_value", "page_locked", "pageable", "input_bytes_per_step", "h2d_copy_GBps",
            "h2d_bound_pairs_per_s", "per_num_stream", "num_stream", "upload_GBps", "note")
    return {k: rec[k] for k in keep if k in rec}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box here has one GPU, and the path "
                                   "shards by frame pair with no inter-GPU traffic"}))
        sys.exit(2)
    import vship_amd
    from vship_amd import synth

    vship_amd.gpu_full_check(0)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    t0 = time.time()
    if a.config == "1080p":
        pairs, bits = video_pairs(1920, 1080, 8, 240, seed=1), 8
        print(f"# generated in {time.time() - t0:.1f}s", flush=True)
        out, scores = run_batched(a, pairs, 8, False, "ssimulacra2 1920x1080 YUV420P8 BT.709 limited, "
                                  "240 pairs per step, inputs resident in HBM")
        if not a.no_cpu:
            out["cpu_baseline"], out["score_vs_fp64_oracle"] = cpu_baseline(pairs, bits, scores, a)
        if not a.no_sub:
            subs = {}
            t1 = time.time()
            subs["host_input"] = brief(sub_host_input(a, pairs, scores))
            print(f"# sub-record host_input: {subs['host_input']['value']} pairs/s ({time.time() - t1:.0f}s)",
                  flush=True)
            del pairs
            t1 = time.time()
            subs["vs_filter"] = brief(sub_vs_filter(a))
            print(f"# sub-record vs_filter: {subs['vs_filter']['value']} pairs/s ({time.time() - t1:.0f}s)",
                  flush=True)
            for name, fn in (("4k10", sub_4k10), ("mixed", lambda x: sub_mixed(x, x.sub_steps)[0]),
                             ("served", sub_served)):
                t1 = time.time()
                subs[name] = brief(fn(a))
                print(f"# sub-record {name}: {subs[name]['value']} pairs/s ({time.time() - t1:.0f}s)", flush=True)
            out["sub_records"] = subs
        print(json.dumps(out), flush=True)
        return
    if a.config == "rgb256":
        pairs, bits = [synth.config0()], 0
        out, scores = run_batched(a, pairs, 0, False, "ssimulacra2 256x256 planar RGB8 full range, 1 pair per step")
    elif a.config == "4k10":
        pairs, bits = video_pairs(3840, 2160, 10, 120, seed=2), 10
        print(f"# generated in {time.time() - t0:.1f}s", flush=True)
        out, scores = run_batched(a, pairs, 10, True, "ssimulacra2 3840x2160 YUV420P10 BT.2020 limited, FP16 "
                                  "planes for scales 0-1, 120 pairs per step, inputs resident in HBM")
    elif a.config == "mixed":
        bits = 8
        out, pairs, scores = sub_mixed(a, per_frame=True)
    elif a.config == "butter1080p":
        out, pairs, norms = run_butter(a)
        if not a.no_cpu:
            out["cpu_baseline"], out["norms_vs_fp64_oracle"] = butter_cpu_baseline(pairs, norms, a)
        print(json.dumps(out), flush=True)
        return
    else:
        out, host_pool, alone = run_served(a)
        pairs, scores, bits = [], [], []
        for key, pp in host_pool.items():
            pairs += pp
            scores += list(alone[key])
            bits += [key[2]] * len(pp)
    if not a.no_cpu:
        out["cpu_baseline"], out["score_vs_fp64_oracle"] = cpu_baseline(pairs, bits, scores, a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
