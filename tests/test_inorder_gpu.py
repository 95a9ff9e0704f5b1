"""GPU parity of the in-order tile path and of the call sequences the
single-pass scenes leave out: every output of every store (colour, depth,
primitive id) against the oracle, bit for bit, and the fragment counts."""
import numpy as np
import pytest
import torch  # noqa: F401

import inorder_scenes as S

pytestmark = pytest.mark.gpu


def _check(sc):
    from oracle_backend import render_calls as oracle_calls
    from rsr_amd import rgl
    print(f"[{sc.name}] {sc.width}x{sc.height} tiles {sc.tile_blocks}", flush=True)
    ref, rst = oracle_calls(sc, threads=8)
    got, gst = rgl.render_calls(sc)
    for k, (kind, r, g) in enumerate(zip(sc.outputs, ref, got)):
        rb, gb = r.view(np.uint32), g.view(np.uint32)
        bad = np.argwhere(rb != gb)
        assert bad.size == 0, f"{sc.name}: output {k} ({kind}) differs at {bad[:8].tolist()} " \
                              f"(oracle {rb[tuple(bad[0])]:#x} gpu {gb[tuple(bad[0])]:#x})"
    assert rst["fragments"] == gst["fragments"], (rst, gst)
    return ref, rst, gst


@pytest.mark.parametrize("build", S.ALL, ids=lambda f: getattr(f, "__name__", "variant"))
def test_call_scene_matches_oracle(build):
    sc = build()
    ref, rst, gst = _check(sc)
    # every scene draws something into its first colour/depth output
    assert rst["fragments"] > 1000


def test_dense_in_order_tiles_sort_in_runs():
    # more than 4096 records in a tile: LDS-sorted runs merged in global memory
    sc = S.scene_dense_blend()
    ref, rst, gst = _check(sc)
    assert gst["tile_pairs"] > 20000


def test_in_order_and_key_frames_in_one_batch():
    # a batch mixing in-order frames and key-path frames: each frame equals
    # the oracle's
    from oracle_backend import render_calls as oracle_calls
    from rsr_amd import rgl
    dev = "cuda:0"
    mix = [S.scene_blend(), S.scene_color_depth(), S.scene_mid_frame_clears(), S.scene_mixed_programs(),
           S.scene_scissor(11, 7, 250, 151, 333, 181, (4, 4), name="scissor_odd_target_tile32")]
    recs, outs = [], []
    for sc in mix:
        g = rgl.GPU(0)
        dev_arrays = {k: torch.from_numpy(np.ascontiguousarray(a)).to(dev) for k, a in sc.arrays.items()}
        ptr_of = {id(sc.arrays[k]): t.data_ptr() for k, t in dev_arrays.items()}
        o = [torch.zeros((sc.height, sc.width), dtype=torch.float32 if k == "depth" else torch.int32, device=dev)
             for k in sc.outputs]
        from rsr_amd.scenes import render_calls
        render_calls(g, sc, [t.data_ptr() for t in o], ptr=lambda a, p=ptr_of: p[id(a)])
        recs.append((g, dev_arrays))
        outs.append(o)
    ex = rgl.GPU(0)
    ex.run_batch_and_wait([g for g, _ in recs], torch.cuda.current_stream().cuda_stream)
    for sc, o in zip(mix, outs):
        ref, _ = oracle_calls(sc, threads=8)
        for k, (r, g) in enumerate(zip(ref, o)):
            gb = g.cpu().numpy().view(np.uint32)
            assert np.array_equal(r.view(np.uint32), gb), f"{sc.name} output {k}"
    ex.close()
    for g, _ in recs:
        g.close()


def test_blended_icospheres_match_oracle():
    # icospheres with alpha blending on, every tile on the in-order path
    import dataclasses
    from oracle_backend import render as oracle_render
    from rsr_amd import rgl, scenes
    sc = scenes.scene_icospheres(640, 360, instances=48)
    sc = dataclasses.replace(sc, blend=True, name=sc.name + "_blend")
    print(f"[{sc.name}]", flush=True)
    rc, rd, rp, rst = oracle_render(sc, threads=8)
    gc, gd, gp, gst = rgl.render(sc)
    for name, r, g in (("colour", rc, gc), ("depth", rd, gd), ("prim", rp, gp)):
        assert np.array_equal(r.view(np.uint32), g.view(np.uint32)), name
    assert rst["fragments"] == gst["fragments"]


def test_blended_config2_equals_the_key_path_frame():
    # config 2 (327,680 triangles at 1920x1080) with alpha blending on takes
    # the in-order path in every tile; its program writes alpha 1, so the
    # blend keeps the source colour exactly and every buffer equals the key
    # path's frame
    import dataclasses
    from rsr_amd import rgl, scenes
    sc = scenes.scene_icospheres()
    print(f"[{sc.name}] key path and blended", flush=True)
    kc, kd, kp, kst = rgl.render(sc)
    bc, bd, bp, bst = rgl.render(dataclasses.replace(sc, blend=True, name=sc.name + "_blend"))
    for name, a, b in (("colour", kc, bc), ("depth", kd, bd), ("prim", kp, bp)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    assert kst["fragments"] == bst["fragments"]
