"""Autotuner for batched frames: picks the tile size, the number of
(triangle, tile) pairs per visibility work item and the quad work from
which a work item's region is cut smaller (together: how many workgroups
share a dense tile) by timing the batch itself.

Every choice draws bit-identical frames (the rasteriser's arithmetic is
invariant to the tile size and the visibility merge is order-independent),
so the tuner only trades speed.  No counterpart exists in the reference,
whose tile size is a viewer parameter (node/gpu.cxx:117).
"""
from __future__ import annotations

import dataclasses
import time

TILE_CHOICES = ((8, 8), (4, 4))          # 64 and 32 pixel tiles (tile_blocks x 8)
SEG_CHOICES = (256, 512, 1024, 2048, 4096)
SPLIT_CHOICES = (0, 512, 1024, 2048)


def _record(ex_frames, scene_list, device, tile_blocks):
    from .rgl import DeviceScene, GPU
    frames = []
    for sc, old in zip(scene_list, ex_frames or [None] * len(scene_list)):
        if old is not None:
            old[0].close()
        g = GPU(0)
        ds = DeviceScene(dataclasses.replace(sc, tile_blocks=tile_blocks), device, prim=True)
        ds.record(g)
        frames.append((g, ds))
    return frames


def time_batch(ex, frames, stream, steps: int, warmup: int = 3) -> float:
    import torch
    recs = [g for g, _ in frames]
    for _ in range(warmup):
        ex.run_batch_and_wait(recs, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ex.run_batch(recs, stream)
    ex.wait()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def autotune(ex, scene_list, device="cuda:0", steps: int = 10, tiles=TILE_CHOICES, segs=SEG_CHOICES,
             splits=SPLIT_CHOICES, log=None):
    """Times every (tile size, pairs per item, split threshold) choice on
    `ex` (graph replay on) and leaves the fastest set: returns (frames,
    choice, table) where frames are the recorders of the chosen tile size."""
    import torch
    stream = torch.cuda.current_stream(device).cuda_stream
    table = []
    best = None
    frames = None
    for tb in tiles:
        frames = _record(frames, scene_list, device, tb)
        for seg in segs:
            for split in splits:
                ex.set_raster_params(seg, split)
                ms = time_batch(ex, frames, stream, steps)
                table.append({"tile_px": tb[0] * 8, "seg_pairs": seg, "split_quads": split, "ms": round(ms, 4)})
                if log:
                    log(table[-1])
                if best is None or ms < best[0]:
                    best = (ms, tb, seg, split)
    _, tb, seg, split = best
    if frames is None or tb != tiles[-1]:
        frames = _record(frames, scene_list, device, tb)
    ex.set_raster_params(seg, split)
    return frames, {"tile_px": tb[0] * 8, "seg_pairs": seg, "split_quads": split}, table
