"""Helpers shared by the GPU tests: synthetic inputs in every format and
one-call wrappers around the batched C ABI."""
import os
import sys

import numpy as np

import vship_amd
from vship_amd import api
from tests import oracle_lib as ol

FMT_NAMES = {api.PF_RGB_U8: "rgb_u8", api.PF_RGB_U16: "rgb_u16", api.PF_RGB_F16: "rgb_f16", api.PF_RGB_F32: "rgb_f32",
             api.PF_YUV420_U8: "yuv420_u8", api.PF_YUV420_U16: "yuv420_u16"}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rgb_pair(w, h, fmt, seed, noise=6.0):
    """Smooth random picture + noisy copy, in the given planar RGB format."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    base = []
    for c in range(3):
        f = np.zeros((h, w))
        for _ in range(4):
            cx, cy, s, a = rng.random() * w, rng.random() * h, 2 + rng.random() * max(w, h) / 3, rng.random()
            f += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
        f = 0.1 + 0.8 * f / max(f.max(), 1e-9) + 0.05 * rng.random((h, w))
        base.append(np.clip(f, 0, 1))
    dist = [np.clip(p + noise / 255.0 * rng.standard_normal(p.shape), 0, 1) for p in base]
    return [to_format(p, fmt) for p in base], [to_format(p, fmt) for p in dist]


def to_format(p, fmt):
    if fmt == api.PF_RGB_U8:
        return np.round(p * 255).astype(np.uint8)
    if fmt == api.PF_RGB_U16:
        return np.round(p * 65535).astype(np.uint16)
    if fmt == api.PF_RGB_F16:
        return p.astype(np.float16)
    return p.astype(np.float32)


ORACLE_FMT = {api.PF_RGB_U8: ol.SO_RGB_U8, api.PF_RGB_U16: ol.SO_RGB_U16, api.PF_RGB_F16: ol.SO_RGB_F16,
              api.PF_RGB_F32: ol.SO_RGB_F32}


def oracle_rgb(ref, dist, fmt):
    return ol.score_rgb(ORACLE_FMT[fmt], ref, dist, "f64")


def gpu_batch(pairs, fmt, bits=8, matrix=api.MATRIX_BT709, fp16=False, th=0, graph=False, features=True,
              persistent_grid=0, chroma_location=api.CHROMA_LEFT):
    """Scores (and features) of a list of (ref, dist) host pairs through the batched C ABI."""
    dp = api.DevicePairs(pairs, fmt, bits, matrix, chroma_location)
    mw = max(p[0][0].shape[1] for p in pairs)
    mh = max(p[0][0].shape[0] for p in pairs)
    with api.SSIMU2Batch(len(pairs), mw, mh, fp16=fp16, map_tile_height=th, use_graph=graph,
                         persistent_grid=persistent_grid) as b:
        out = b.compute(dp.descs, features=features)
    dp.free()
    return out
