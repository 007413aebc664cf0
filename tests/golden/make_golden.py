"""Writes the golden fixtures of this directory.

    python tests/golden/make_golden.py oracle [DIR]      # CPU: oracle_vectors.json
    python tests/golden/make_golden.py reference [DIR]   # GPU: reference_hip.json
    python tests/golden/make_golden.py reference-butter [DIR]   # GPU: reference_butter.json

(DIR defaults to this directory.)

oracle_vectors.json: the FP64 oracle's score and 108 features on the
256x256 RGB8 pair of vship_amd.synth.config0(), per-scale sizes, the
Gaussian taps and a YUV 4:2:0 -> R'G'B' sample (8x6, 10-bit, BT.2020).
A regression record: tests/test_golden.py recomputes every entry.

reference_hip.json: score and 108 features of the reference's own HIP
implementation (built from its sources by `make -C oracle ref`, run through
oracle/_ref/libvship_ref.so on a GPU) on the inputs of REF_CASES; the CPU
test suite then checks the oracle against the reference's output without a
GPU. Inputs are regenerated from their seeds (synth.config0 and
case_inputs below are deterministic).

reference_butter.json: Butteraugli norms (Q-norm, 3-norm, inf-norm), the
sum of the distortion map and 64 map samples of the reference's own HIP
Butteraugli (oracle/_ref/libvship_ref_butter.so) on BUTTER_CASES, for the
CPU checks of the Butteraugli oracle.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# (name, w, h, format name, seed); seed None: synth.config0()
REF_CASES = [
    ("config0", 256, 256, "u8", None),
    ("u16_333x217", 333, 217, "u16", 3),
    ("f32_640x360", 640, 360, "f32", 4),
    ("f16_97x61", 97, 61, "f16", 5),
    ("u8_64x48", 64, 48, "u8", 6),
]
FMT_ORACLE = {"u8": 0, "u16": 1, "f16": 2, "f32": 3}


def case_inputs(w, h, fmt, seed):
    """Planar RGB (ref, dist) of one case, regenerated from its seed."""
    from vship_amd import synth

    if seed is None:
        return synth.config0()
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    base = []
    for _ in range(3):
        f = np.zeros((h, w))
        for _ in range(4):
            cx, cy, s, a = rng.random() * w, rng.random() * h, 2 + rng.random() * max(w, h) / 3, rng.random()
            f += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
        f = 0.1 + 0.8 * f / max(f.max(), 1e-9) + 0.05 * rng.random((h, w))
        base.append(np.clip(f, 0, 1))
    dist = [np.clip(p + 6.0 / 255.0 * rng.standard_normal(p.shape), 0, 1) for p in base]

    def conv(p):
        if fmt == "u8":
            return np.round(p * 255).astype(np.uint8)
        if fmt == "u16":
            return np.round(p * 65535).astype(np.uint16)
        if fmt == "f16":
            return p.astype(np.float16)
        return p.astype(np.float32)

    return [conv(p) for p in base], [conv(p) for p in dist]


def yuv_sample():
    """8x6 10-bit 4:2:0 planes (fixed values) for the front-end record."""
    rng = np.random.default_rng(99)
    y = rng.integers(64, 941, (6, 8)).astype(np.uint16)
    u = rng.integers(64, 961, (3, 4)).astype(np.uint16)
    v = rng.integers(64, 961, (3, 4)).astype(np.uint16)
    return [y, u, v]


def make_oracle(outdir=HERE):
    from tests import oracle_lib as ol
    from vship_amd import synth

    ref, dist = synth.config0()
    s, f = ol.score_rgb(ol.SO_RGB_U8, ref, dist, "f64")
    planes = yuv_sample()
    rgb = ol.yuv_to_rgb(planes, 10, ol.SO_MAT_BT2020NCL)
    out = {
        "config0": {"score": s, "features": f.tolist(),
                    "input_sha": [int(np.sum(p.astype(np.uint64) * (np.arange(p.size, dtype=np.uint64) % 251 + 1)
                                         .reshape(p.shape))) for p in ref + dist]},
        "scale_dims_1920x1080": ol.scale_dims(1920, 1080),
        "gaussian_taps": ol.gaussian_taps().tolist(),
        "yuv420p10_bt2020_8x6": {"planes": [p.tolist() for p in planes], "rgb": rgb.tolist()},
    }
    path = os.path.join(outdir, "oracle_vectors.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(path)


def make_reference(outdir=HERE):
    from tests import oracle_lib as ol

    if not ol.ref_available():
        raise SystemExit("oracle/_ref/libvship_ref.so not built (make -C oracle ref)")
    out = {"source": "reference src/ssimu2 HIP code via oracle/ref_shim.hip, FP32", "cases": []}
    for name, w, h, fmt, seed in REF_CASES:
        ref, dist = case_inputs(w, h, fmt, seed)
        s, f = ol.ref_score_rgb(FMT_ORACLE[fmt], ref, dist)
        out["cases"].append({"name": name, "w": w, "h": h, "format": fmt, "seed": seed, "score": s,
                             "features": [float(x) for x in f]})
        print(f"{name}: reference score {s:.6f}", flush=True)
    path = os.path.join(outdir, "reference_hip.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(path)


# Butteraugli cases: (name, w, h, format, seed, qnorm); inputs from case_inputs
BUTTER_CASES = [
    ("f32_64x48", 64, 48, "f32", 11, 2),
    ("u16_75x47", 75, 47, "u16", 12, 2),
    ("f32_160x90_q3", 160, 90, "f32", 13, 3),
    ("f32_config0", 256, 256, "f32", None, 2),
]


def butter_inputs(w, h, fmt, seed):
    if seed is None:  # the 256x256 RGB8 pair of config0, as float
        ref, dist = case_inputs(w, h, "u8", None)
        return [p.astype(np.float32) / 255.0 for p in ref], [p.astype(np.float32) / 255.0 for p in dist]
    return case_inputs(w, h, fmt, seed)


def make_reference_butter(outdir=HERE):
    from tests import butter_lib as bl

    if not bl.ref_available():
        raise SystemExit("oracle/_ref/libvship_ref_butter.so not built (make -C oracle ref)")
    out = {"source": "reference src/butter HIP code via oracle/ref_butter_shim.hip, FP32, intensity 203",
           "cases": []}
    for name, w, h, fmt, seed, q in BUTTER_CASES:
        ref, dist = butter_inputs(w, h, fmt, seed)
        print(f"case {name}: {w}x{h} {fmt} qnorm {q}", flush=True)
        norms, dm = bl.ref_butteraugli(ref, dist, q, 203.0)
        pts = bl.sample_positions(w, h)
        out["cases"].append({"name": name, "w": w, "h": h, "format": fmt, "seed": seed, "qnorm": q,
                             "norms": [float(x) for x in norms], "map_sum": float(dm.sum()),
                             "map_samples": [[x, y, float(dm[y, x])] for x, y in pts]})
        print(f"{name}: reference norms {norms}", flush=True)
    path = os.path.join(outdir, "reference_butter.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(path)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "oracle"
    outdir = sys.argv[2] if len(sys.argv) > 2 else HERE
    os.makedirs(outdir, exist_ok=True)
    {"oracle": make_oracle, "reference": make_reference, "reference-butter": make_reference_butter}[what](outdir)
