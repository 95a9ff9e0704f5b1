"""GPU tests of batches (many frames, one launch sequence) and of graph
replay: every frame of a batch must equal the frame drawn alone and the
oracle's frame bit for bit; graph replays must equal eager runs."""
import dataclasses

import numpy as np
import pytest
import torch

from rsr_amd import rgl, scenes
from test_parity_gpu import _compare

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _recorders(scene_list):
    out = []
    for sc in scene_list:
        g = rgl.GPU(0)
        ds = rgl.DeviceScene(sc, DEV, prim=True)
        ds.record(g)
        out.append((g, ds))
    return out


def _same(name, a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{name}: buffer {k} differs"


def test_batch_matches_single_and_oracle():
    from oracle_backend import render as oracle_render
    mix = [
        scenes.scene_sphere_ground(),
        scenes.scene_icospheres(640, 360, instances=64, seed=11, tex_dim=256),
        scenes.scene_clip_floor(320, 200, 16),
        dataclasses.replace(scenes.scene_icospheres(488, 272, instances=40, seed=5, tex_dim=256),
                            tile_blocks=(4, 4), name="icospheres40_488x272_tile32"),
        dataclasses.replace(scenes.scene_sphere_ground(320, 176), cull=False, depth_func=scenes.GL_LEQUAL,
                            name="sphere_ground_leq_320x176"),
        scenes.scene_batch(3, seed=21)[2],
    ]
    stream = torch.cuda.current_stream().cuda_stream
    recs = _recorders(mix)
    ex = rgl.GPU(0)
    ex.run_batch_and_wait([g for g, _ in recs], stream)
    bst = ex.stats()
    batched = [ds.outputs() for _, ds in recs]
    frags = 0
    for (g, ds), sc, out in zip(recs, mix, batched):
        g.run_and_wait(stream)
        alone = ds.outputs()
        st = g.stats()
        frags += st["fragments"]
        _same(sc.name, out, alone)
        _compare(sc.name, oracle_render(sc, threads=8), (*out, st))
    assert bst["fragments"] == frags
    assert bst["tiles"] == sum(g.stats()["tiles"] for g, _ in recs)
    ex.close()
    for g, _ in recs:
        g.close()


def test_batched_config3_64_scenes():
    from oracle_backend import render as oracle_render
    batch = scenes.scene_batch()
    stream = torch.cuda.current_stream().cuda_stream
    recs = _recorders(batch)
    ex = rgl.GPU(0)
    ex.set_graph(True)
    for _ in range(3):           # capture, then replays
        ex.run_batch_and_wait([g for g, _ in recs], stream)
    st = ex.stats()
    print(f"64-scene batch: {st}", flush=True)
    assert st["primitives_submitted"] == sum(s.triangles for s in batch)
    batched = [ds.outputs() for _, ds in recs]
    for k, ((g, ds), sc, out) in enumerate(zip(recs, batch, batched)):
        g.run_and_wait(stream)
        _same(sc.name, out, ds.outputs())
        _compare(sc.name, oracle_render(sc, threads=8), (*out, g.stats()))
    ex.close()
    for g, _ in recs:
        g.close()


def test_graph_replay_matches_eager():
    sc = scenes.scene_icospheres(960, 540, instances=96, seed=5, tex_dim=256)
    other = scenes.scene_clip_floor(320, 200, 16)
    stream = torch.cuda.current_stream().cuda_stream
    (g, ds), (g2, ds2) = _recorders([sc, other])
    g.run_and_wait(stream)
    ref = ds.outputs()
    g.set_graph(True)
    for k in range(5):
        ds.color.zero_(); ds.depth.zero_(); ds.prim.zero_()
        torch.cuda.synchronize()
        g.run_and_wait(stream)
        _same(f"replay {k}", ds.outputs(), ref)
        g2.run_and_wait(stream)          # eager work of the program between replays
        torch.cuda.synchronize()
    caps = g.graph_captures()
    print(f"graph captures: {caps}")
    assert 1 <= caps <= 2
    g.close(); g2.close()


def test_stress_4k_views_batched():
    # the 4K stress geometry at 800 instances (256k triangles per view), four
    # views in one batch, each against the oracle; the fixed guard band
    from oracle_backend import render as oracle_render
    views = [scenes.scene_stress(v, instances=800) for v in range(4)]
    stream = torch.cuda.current_stream().cuda_stream
    recs = _recorders(views)
    ex = rgl.GPU(0)
    ex.run_batch_and_wait([g for g, _ in recs], stream)
    for (g, ds), sc in zip(recs, views):
        out = ds.outputs()
        g.run_and_wait(stream)
        _same(sc.name, out, ds.outputs())
        _compare(sc.name, oracle_render(sc, threads=8), (*out, g.stats()))
    ex.close()
    for g, _ in recs:
        g.close()


def test_stress_4k_reference_guard_band():
    # the reference's guard band above 2048 pixels (factor 1/15 at 3840):
    # on-screen triangles beyond it are culled or clipped, as in the oracle
    from oracle_backend import render as oracle_render
    sc = scenes.scene_stress(2, instances=800, fixed_guard_band=False)
    ref = oracle_render(sc, threads=8)
    got = rgl.render(sc)
    _compare(sc.name, ref, got)
    assert got[3]["primitives_clipped"] > 0
    assert (got[1] < 1.0).mean() < 0.5 * (oracle_render(scenes.scene_stress(2, instances=800), threads=8)[1] < 1.0).mean()


def test_stress_2m_view_matches_oracle():
    # one full-size stress view (2M triangles, 3840x2160, fixed guard band)
    # against the oracle
    from oracle_backend import render as oracle_render
    sc = scenes.scene_stress(1)
    got = rgl.render(sc)
    _compare(sc.name, oracle_render(sc, threads=16), got)
    assert got[3]["primitives_submitted"] == 2_000_000


def test_stress_2m_batch_equals_single():
    # full size (2M triangles per view): the batch of four views equals each
    # view drawn alone
    views = [scenes.scene_stress(v) for v in range(4)]
    stream = torch.cuda.current_stream().cuda_stream
    recs = _recorders(views)
    ex = rgl.GPU(0)
    ex.run_batch_and_wait([g for g, _ in recs], stream)
    bst = ex.stats()
    outs = [ds.outputs() for _, ds in recs]
    frags = 0
    for (g, ds), sc, out in zip(recs, views, outs):
        g.run_and_wait(stream)
        _same(sc.name, out, ds.outputs())
        frags += g.stats()["fragments"]
    print(f"stress batch: {bst}", flush=True)
    assert bst["fragments"] == frags and bst["primitives_submitted"] == 8_000_000
    ex.close()
    for g, _ in recs:
        g.close()


def test_raster_params_draw_identical_frames():
    # every (pairs per work item, region split) setting draws the same
    # frame: single and multi-segment tiles, quadrants cut in two and four
    # (the split follows the previous frame's quad work, so each setting
    # runs twice)
    from oracle_backend import render as oracle_render
    mix = [scenes.scene_icospheres(960, 540, instances=128, seed=7, tex_dim=256),
           dataclasses.replace(scenes.scene_sphere_ground(320, 176), cull=False, depth_func=scenes.GL_LEQUAL,
                               name="sphere_ground_leq_320x176"),
           scenes.scene_clip_floor(320, 200, 16)]
    stream = torch.cuda.current_stream().cuda_stream
    recs = _recorders(mix)
    ref = None
    for seg, split in ((512, 0), (64, 0), (64, 1), (128, 16), (256, 64), (4096, 1), (4096, 0)):
        ex = rgl.GPU(0)
        ex.set_raster_params(seg, split)
        for rep in range(2):
            ex.run_batch_and_wait([g for g, _ in recs], stream)
            out = [ds.outputs() for _, ds in recs]
            if ref is None:
                ref = out
                for sc, o in zip(mix, out):
                    g, ds = recs[mix.index(sc)]
                    g.run_and_wait(stream)
                    _compare(sc.name, oracle_render(sc, threads=8), (*o, g.stats()))
            else:
                for sc, a, b in zip(mix, out, ref):
                    _same(f"{sc.name} seg {seg} split {split} run {rep}", a, b)
        ex.close()
    for g, _ in recs:
        g.close()
