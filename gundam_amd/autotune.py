"""Launch-shape autotuner for the fill kernel.

The tile record capacity fixes the device layout, so every candidate is built
afresh on a prefix of the caller's events (same dial mix, smaller count) and
timed with the library's own kernel timers (HIP events around the fill
kernel).  The candidate with the smallest fill time per evaluation wins.
"""
from __future__ import annotations

import itertools
import time
from typing import Iterable, Optional

import numpy as np

from . import _native as N
from .cache_manager import CacheManager
from .tables import Tables

LDS_BYTES_PER_CU = 160 * 1024


def mean_record_bytes(t: Tables, precision: str = "fp64") -> float:
    """Mean device record bytes per dial (gcm.hip record_words)."""
    fp32 = precision == "fp32"
    ty = t.dial_type.astype(np.int64)
    ln = np.diff(t.data_begin).astype(np.int64)
    w = ln.copy()  # splines and graphs: the payload length
    w[ty == N.DIAL_NORM] = 0
    w[ty == N.DIAL_SHIFT] = 1
    w[ty == N.DIAL_TABULATED] = 2 if fp32 else 1
    kr = ty == N.DIAL_KRIGED
    w[kr] = ((ln[kr] - 1) // 2) * (2 if fp32 else 1)
    grid = (ty == N.DIAL_BILINEAR) | (ty == N.DIAL_BICUBIC)
    w[grid] = 1 + ln[grid]
    n = max(1, w.size)
    return max(float(w.sum()) * (4 if fp32 else 8) / n, 1.0) if w.size else 8.0


def max_dials_for(tile_words: int, mean_bytes: float = 8.0, most: int = 0) -> int:
    """The library's dial slots per tile for a tile size (gcm.hip build_layout);
    most = the largest binned event's dial count."""
    want = max(int(np.ceil(1.12 * (tile_words * 8.0) / mean_bytes)), int(most))
    return min(1024, max(256, (want + 127) // 128 * 128))


def lds_per_block(tile_words: int, n_inputs: int, k_points: int = 1,
                  ring_buffers: int = 2, blocks_per_cu: int = 1,
                  mean_bytes: float = 8.0, most: int = 0) -> int:
    """LDS of one fill block (mirrors the host computation in gcm.hip: the
    inputs' x values are staged only when they are small and still let
    blocks_per_cu blocks share the CU; the 512 B of input-class records are
    always staged)."""
    max_dials = max_dials_for(tile_words, mean_bytes, most)
    ring = ring_buffers * (tile_words * 8 + max_dials * 8 + 64 * 16 + 32) + 2 * max_dials * 8 + 16
    in_bytes = k_points * n_inputs * 8
    classes = 16 * 32
    staged = (in_bytes + classes + 1023) // 1024 * 1024
    if in_bytes <= 16384 and (staged + ring) * blocks_per_cu <= LDS_BYTES_PER_CU:
        return staged + ring
    return (classes + 1023) // 1024 * 1024 + ring


def head(t: Tables, n_events: int) -> Tables:
    """The first n_events events of t with their dials (views, no copy)."""
    n = min(n_events, t.n_events)
    nd = int(t.dial_begin[n])
    return Tables(base_weight=t.base_weight[:n], bin=t.bin[:n],
                  dial_begin=t.dial_begin[:n + 1], dial_type=t.dial_type[:nd],
                  dial_input=t.dial_input[:nd], data_begin=t.data_begin[:nd + 1],
                  dial_data=t.dial_data, inputs=t.inputs,
                  table_begin=t.table_begin, n_parameters=t.n_parameters,
                  n_bins=t.n_bins, reweight_cap=t.reweight_cap,
                  dial_input2=None if t.dial_input2 is None else t.dial_input2[:nd],
                  _keep=[t])


def candidates(n_inputs: int, precision: str = "fp64", k_points: int = 1,
               mean_bytes: float = 8.0, most: int = 0) -> list:
    out = []
    for nb, tile, block, bpc in itertools.product((2, 3), (1024, 1536, 2048, 2560, 3072, 3584,
                                                         3840, 4096),
                                                  (256, 320, 384, 448, 512, 768),
                                                  (1, 2, 3, 4, 5, 6)):
        lds = lds_per_block(tile, n_inputs, k_points, ring_buffers=nb, blocks_per_cu=bpc,
                            mean_bytes=mean_bytes, most=most)
        # fill the CU's LDS: only the largest block count that fits
        if lds * bpc > LDS_BYTES_PER_CU or lds * (bpc + 1) <= LDS_BYTES_PER_CU:
            continue
        if (block + 64 * nb) * bpc > 2048:  # 32 waves per CU
            continue
        out.append({"tile_words": tile, "block_threads": block, "blocks_per_cu": bpc,
                    "ring_buffers": nb})
    return out


def autotune(tables: Tables, theta: np.ndarray, table_values: Optional[np.ndarray],
             precision: str = "fp64", n_events: int = 1_000_000, reps: int = 5,
             cands: Optional[Iterable[dict]] = None, device: int = 0,
             log=None) -> dict:
    """theta may hold K points ([K, P]): the shapes are then timed on K-point
    launches."""
    theta = np.atleast_2d(theta)
    k_points = theta.shape[0]
    sub = head(tables, n_events)
    binned = sub.bin >= 0
    most = int(np.diff(sub.dial_begin)[binned].max()) if binned.any() else 0
    cands = list(cands) if cands is not None else candidates(
        int(tables.inputs.shape[0]), precision, k_points, mean_record_bytes(sub, precision), most)
    results = []
    t_start = time.time()
    for c in cands:
        cm = CacheManager(device, precision=precision, **c)
        try:
            cm.build(sub, max_points=k_points)
            cm.propagate_parameters(theta, table_values)
            cm.enable_timing(True)
            for _ in range(reps):
                cm.fill()
            cm.sync()
            tm = cm.timing()
            ms = tm["fill_ms"] / max(1, tm["count"])
        finally:
            cm.close()
        results.append((ms, c))
        if log:
            log(f"autotune {c}: fill {ms:.4f} ms")
    results.sort(key=lambda r: r[0])
    best = dict(results[0][1])
    best["fill_ms_subset"] = results[0][0]
    best["subset_events"] = sub.n_events
    best["candidates"] = len(results)
    best["tune_seconds"] = time.time() - t_start
    return best
