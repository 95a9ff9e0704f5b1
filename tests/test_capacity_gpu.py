"""Per-frame capacity growth: a frame whose (triangle, tile) pairs exceed
the list capacity sized from its triangle count (2 per triangle + 4096)
reports RGL_E_CAPACITY at the wait, grows the capacity, and draws the
oracle's buffers bit for bit when run again; single- and double-buffered
(where the grown capacity reaches every executor)."""
import numpy as np
import pytest
import torch

import inorder_scenes as S
from rsr_amd import rgl, scenes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def scene_many_grounds(n=100):
    # the ground quad crosses the near plane: its clipped fans span about
    # 800 tiles of 32 px; drawn n times, then the sphere (clipped list)
    sc, pro, _ = S.base(width=1280, height=720, tile_blocks=(4, 4), name=f"ground_x{n}_1280x720_tile32")
    sc.calls = pro + [("clear", S.CB | S.DB)] + S.lambert(sc)
    for _ in range(n):
        sc.calls += S.draw(sc, "ground")
    sc.calls += S.draw(sc, "sphere") + S.stores()
    sc.outputs = S.OUT3
    return sc


def scene_many_walls(n=100):
    # a quad standing inside the view (about 90 tiles of 32 px per
    # triangle, nothing clipped) drawn n times, then the sphere (inline list)
    sc, pro, _ = S.base(width=1280, height=720, tile_blocks=(4, 4), name=f"wall_x{n}_1280x720_tile32")
    pos = np.array([(-1.5, -0.2, 0.0), (1.5, -0.2, 0.0), (1.5, 2.5, 0.0), (-1.5, 2.5, 0.0)], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    uv = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float32)
    bufs = {k: scenes.pad4(pos[:, k]) for k in range(3)}
    bufs.update({3 + k: scenes.pad4(nrm[:, k]) for k in range(3)})
    bufs[9], bufs[10] = scenes.pad4(uv[:, 0]), scenes.pad4(uv[:, 1])
    sc.arrays.update(scenes.mesh_arrays("wall", scenes.Mesh(bufs, np.array([0, 1, 2, 0, 2, 3], np.uint16))))
    sc.calls = pro + [("clear", S.CB | S.DB), ("disable", S.GL_CULL_FACE)] + S.lambert(sc)
    for _ in range(n):
        sc.calls += S.draw(sc, "wall")
    sc.calls += [("enable", S.GL_CULL_FACE)] + S.draw(sc, "sphere") + S.stores()
    sc.outputs = S.OUT3
    return sc


@pytest.mark.parametrize("double", [False, True], ids=["single", "double_buffered"])
@pytest.mark.parametrize("build", [scene_many_walls, scene_many_grounds], ids=["inline_list", "clipped_list"])
def test_list_overflow_grows_and_redraws(build, double):
    from oracle_backend import render_calls as oracle_calls
    from rsr_amd.scenes import output_array, render_calls
    sc = build()
    ref, rst = oracle_calls(sc, threads=8)
    print(f"[{sc.name}] oracle {rst}", flush=True)
    # more pairs than the first capacity of the list concerned
    assert rst["tile_pairs"] > 2 * rst["submitted"] + 4096 or rst["clipped_tile_pairs"] > 65536
    g = rgl.GPU(0)
    if double:
        g.set_double_buffer(True)
    arrs = {k: torch.from_numpy(np.ascontiguousarray(a)).to(DEV) for k, a in sc.arrays.items()}
    ptr_of = {id(sc.arrays[k]): t.data_ptr() for k, t in arrs.items()}
    host = [output_array(kind, sc.width, sc.height) for kind in sc.outputs]
    outs = [torch.from_numpy(h.view(np.int32) if h.dtype == np.uint32 else h).to(DEV) for h in host]
    render_calls(g, sc, [o.data_ptr() for o in outs], ptr=lambda a: ptr_of[id(a)])
    stream = torch.cuda.current_stream(DEV).cuda_stream
    g.run(stream)
    assert g.L.rgl_gpu_wait(g.h) == rgl.RGL_E_CAPACITY
    g.run_and_wait(stream, retries=1)
    st = g.stats()
    assert st["overflow"] == 0 and st["tile_pairs"] == rst["tile_pairs"], (st, rst)
    assert st["clipped_tile_pairs"] == rst["clipped_tile_pairs"] and st["fragments"] == rst["fragments"], (st, rst)
    for k, (r, o, h) in enumerate(zip(ref, outs, host)):
        rb, gb = r.view(np.uint32), o.cpu().numpy().view(h.dtype).view(np.uint32)
        bad = np.argwhere(rb != gb)
        assert bad.size == 0, f"{sc.name}: output {k} differs at {bad[:8].tolist()}"
    if double:
        # a second run of the same frame on the other executor draws without
        # an overflow too
        g.run(stream)
        g.run(stream)
        assert g.L.rgl_gpu_wait(g.h) == 0
    g.close()
