"""Times the oracle's CPU path on a scene like the one the reference was
timed on in this container (1920x1080, 327,690 small random triangles in
33 DrawElements calls, rglv::BaseProgram, RB_RGBF32 + RB_F32, 64 px tiles,
sRGB store): the serial bin and the tile draw + store at 1, 4 and 8
threads, best of N frames.  Test infrastructure: it checks that the CPU
baseline bench.py reports runs at the reference's speed.

  python oracle/calibrate_cpu.py [--frames 10] [--threads 1,4,8]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from rsr_amd import scenes  # noqa: E402
from rsr_amd.scenes import (GL_COLOR_ATTACHMENT0, GL_COLOR_BUFFER_BIT, GL_DEPTH_ATTACHMENT,  # noqa: E402
                            GL_DEPTH_BUFFER_BIT, GL_TRIANGLES, GL_UNSIGNED_SHORT, RB_F32, RB_RGBF32)

DRAWS, TRIS_PER_DRAW = 33, 9930


def calibration_scene(width=1920, height=1080, seed=12, spread_px=8.0):
    """Small triangles scattered over the screen at random depths, drawn
    with identity matrices (positions are NDC), no culling."""
    rng = np.random.default_rng(seed)
    arrays, calls = {}, []
    sx, sy = 2.0 * spread_px / width, 2.0 * spread_px / height
    for d in range(DRAWS):
        c = np.c_[rng.uniform(-1, 1, TRIS_PER_DRAW), rng.uniform(-1, 1, TRIS_PER_DRAW), rng.uniform(-0.9, 0.9, TRIS_PER_DRAW)]
        v = np.repeat(c, 3, axis=0)
        v[:, 0] += rng.uniform(-sx, sx, len(v))
        v[:, 1] += rng.uniform(-sy, sy, len(v))
        for k in range(3):
            arrays[f"m{d}.{k}"] = scenes.pad4(v[:, k].astype(np.float32))
        arrays[f"m{d}.idx"] = np.arange(3 * TRIS_PER_DRAW, dtype=np.uint16)
        calls += [("use_buffer", s, f"@m{d}.{s}" if s < 3 else None) for s in range(16)]
        calls.append(("draw_elements", GL_TRIANGLES, 3 * TRIS_PER_DRAW, GL_UNSIGNED_SHORT, f"@m{d}.idx", 0))
    eye = scenes.mat4_rows(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)
    sc = scenes.CallScene(name=f"calibration_{DRAWS * TRIS_PER_DRAW}_{width}x{height}", width=width, height=height,
                          tile_blocks=(8, 8), arrays=arrays)
    sc.calls = [("renderbuffer_type", GL_COLOR_ATTACHMENT0, RB_RGBF32), ("renderbuffer_type", GL_DEPTH_ATTACHMENT, RB_F32),
                ("clear_color", 0.0, 0.0, 0.0), ("clear_depth", 1.0), ("view_matrix", eye), ("projection_matrix", eye),
                ("clear", GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT), ("use_program", scenes.PGM_BASE)] + calls + \
        [("use_program", scenes.PGM_DEFAULT_POST), ("store_color_tc", ("out", 0), True)]
    sc.outputs = ["color"]
    return sc


def time_frames(sc, threads, frames):
    """(bin ms, draw ms, frame ms, fragments) per frame."""
    import ctypes
    from oracle_backend import OracleGPU
    from rsr_amd.scenes import render_calls
    g = OracleGPU(threads)
    g.L.rglo_gpu_times.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    out = [np.zeros((sc.height, sc.width), np.uint32)]
    rows = []
    t = np.zeros(2, np.float64)
    for _ in range(frames):
        render_calls(g, sc, out)
        g.run()
        g.L.rglo_gpu_times(g.h, t.ctypes.data)
        st = g.stats()
        rows.append((t[0] * 1e3, t[1] * 1e3, (t[0] + t[1]) * 1e3, st["fragments"]))
        g.L.rglo_gl_reset(g.h)
    g.close()
    return np.array(rows)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=10)
    ap.add_argument("--threads", default="1,4,8")
    a = ap.parse_args()
    sc = calibration_scene()
    print(f"{sc.name}: {DRAWS} draws x {TRIS_PER_DRAW} triangles, BaseProgram, 64 px tiles", flush=True)
    for n in (int(x) for x in a.threads.split(",")):
        r = time_frames(sc, n, a.frames)
        best = r[np.argmin(r[:, 2])]
        print(f"threads {n}: bin {r[:, 0].min():.1f} ms  draw {r[:, 1].min():.1f} ms  frame {best[2]:.1f} ms  "
              f"fragments {int(best[3])}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
