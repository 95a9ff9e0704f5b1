"""Double-buffered runs (rgl_gpu_set_double_buffer, the reference's
rglv::doubleBuffer, rglv_gpu.cxx:16,96-116): frames queued without a wait
in between alternate between two executors, and each draws the buffers
the oracle draws, bit for bit; frames that store into one target keep
submission order; the modes the double-buffered path does not run are
refused, and switching back gives single-buffered runs again."""
import numpy as np
import pytest
import torch

import inorder_scenes as S
from rsr_amd import scenes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _stream():
    return torch.cuda.current_stream(DEV).cuda_stream


@pytest.mark.parametrize("in_flight", [1, 3])
def test_call_scenes_queued_double_buffered_match_oracle(in_flight):
    # key-path and in-order frames (blend, mid-frame clears, several
    # programs, scissor) queued back to back, each into its own outputs;
    # in_flight 1: two executors, 3: three
    from oracle_backend import render_calls as oracle_calls
    from rsr_amd import rgl
    from rsr_amd.scenes import output_array, render_calls
    mix = [S.scene_color_depth(), S.scene_blend(), S.scene_mid_frame_clears(), S.scene_mixed_programs(),
           S.scene_scissor(11, 7, 250, 151, 333, 181, (4, 4), name="scissor_odd_target_tile32"), S.scene_color_depth()]
    g = rgl.GPU(0)
    g.set_double_buffer(in_flight)
    keep = []
    for sc in mix:
        print(f"[double-buffered] queue {sc.name} {sc.width}x{sc.height}", flush=True)
        arrs = {k: torch.from_numpy(np.ascontiguousarray(a)).to(DEV) for k, a in sc.arrays.items()}
        ptr_of = {id(sc.arrays[k]): t.data_ptr() for k, t in arrs.items()}
        host = [output_array(kind, sc.width, sc.height) for kind in sc.outputs]
        outs = [torch.from_numpy(h.view(np.int32) if h.dtype == np.uint32 else h).to(DEV) for h in host]
        render_calls(g, sc, [o.data_ptr() for o in outs], ptr=lambda a, p=ptr_of: p[id(a)])
        g.run(_stream())
        keep.append((sc, arrs, outs, host))
    g.wait()
    for sc, _, outs, host in keep:
        ref, _ = oracle_calls(sc, threads=8)
        for k, (r, o, h) in enumerate(zip(ref, outs, host)):
            rb, gb = r.view(np.uint32), o.cpu().numpy().view(h.dtype).view(np.uint32)
            bad = np.argwhere(rb != gb)
            assert bad.size == 0, f"{sc.name}: output {k} differs at {bad[:8].tolist()}"
    g.close()


def test_frames_into_one_target_store_in_submission_order():
    # a 327k-triangle frame, then a light one, into the same buffers, three
    # times over: the buffers end as the light frame drawn alone, and the
    # stats are the last frame's
    from rsr_amd import rgl
    heavy = scenes.scene_icospheres()
    light = scenes.scene_sphere_ground(width=1920, height=1080)
    g = rgl.GPU(0)
    g.set_double_buffer(True)
    dh = rgl.DeviceScene(heavy, DEV, prim=True)
    dl = rgl.DeviceScene(light, DEV, prim=True)
    dl.color, dl.depth, dl.prim = dh.color, dh.depth, dh.prim
    torch.cuda.synchronize()
    print("[double-buffered] heavy, light x3 into one target", flush=True)
    for _ in range(3):
        dh.record(g)
        g.run(_stream())
        dl.record(g)
        g.run(_stream())
    g.wait()
    got = dh.outputs()
    st = g.stats()
    one = rgl.GPU(0)
    ref = rgl.DeviceScene(light, DEV, prim=True)
    ref.record(one)
    one.run_and_wait(_stream())
    exp = ref.outputs()
    for name, a, b in zip(("colour", "depth", "prim"), exp, got):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{name}: the heavy frame's stores came last"
    est = one.stats()
    assert st["primitives_submitted"] == est["primitives_submitted"] == light.triangles
    assert st["fragments"] == est["fragments"]
    # the other order ends with the heavy frame
    dl.record(g)
    g.run(_stream())
    dh.record(g)
    g.run(_stream())
    g.wait()
    hv = rgl.DeviceScene(heavy, DEV, prim=True)
    hv.record(one)
    one.run_and_wait(_stream())
    for name, a, b in zip(("colour", "depth", "prim"), hv.outputs(), dh.outputs()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{name}: heavy frame last"
    one.close()
    g.close()


def test_refused_modes_and_switching_back():
    from rsr_amd import rgl
    sc = scenes.scene_sphere_ground()
    g = rgl.GPU(0)
    d = rgl.DeviceScene(sc, DEV, prim=True)
    d.record(g)
    g.run_and_wait(_stream())
    single = d.outputs()
    g.set_double_buffer(True)
    for on, off in ((g.set_timing, g.set_timing), (g.set_graph, g.set_graph)):
        on(True)
        with pytest.raises(rgl.RglError):
            g.run(_stream())
        off(False)
    with pytest.raises(rgl.RglError):
        g.run_batch([g], _stream())
    with pytest.raises(rgl.RglError):
        g.set_double_buffer(5)
    d.color.zero_()
    torch.cuda.synchronize()
    d.record(g)
    g.run(_stream())
    g.run(_stream())
    g.wait()
    for a, b in zip(single, d.outputs()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    g.set_double_buffer(False)
    d.color.zero_()
    torch.cuda.synchronize()
    g.run_and_wait(_stream())
    for a, b in zip(single, d.outputs()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    g.close()


def test_wait_previous_returns_the_frame_before_the_last():
    # two canvases in turn, as a viewer would: after each Run, wait for the
    # frame before it and check that frame's canvas and stats; the last
    # frame is checked after the final wait
    from rsr_amd import rgl
    light = scenes.scene_sphere_ground(width=1920, height=1080)
    heavy = scenes.scene_icospheres()
    g = rgl.GPU(0)
    g.set_double_buffer(True)
    ref = {}
    one = rgl.GPU(0)
    for sc in (light, heavy):
        r = rgl.DeviceScene(sc, DEV, prim=True)
        r.record(one)
        one.run_and_wait(_stream())
        ref[sc.name] = (r.outputs(), one.stats())
    one.close()
    canv = [rgl.DeviceScene(sc, DEV, prim=True) for sc in (light, heavy)]
    torch.cuda.synchronize()
    seq = [0, 1, 0, 1, 1, 0]
    for k, c in enumerate(seq):
        print(f"[double-buffered] frame {k}: {canv[c].scene.name}", flush=True)
        canv[c].record(g)
        g.run(_stream())
        if k == 0:
            continue
        g.wait_previous()
        prev = canv[seq[k - 1]]
        outs, st = ref[prev.scene.name]
        if seq[k - 1] != c:   # the previous frame's canvas is not being redrawn
            for a, b in zip(outs, prev.outputs()):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (k, prev.scene.name)
        assert g.stats()["primitives_submitted"] == st["primitives_submitted"], k
    g.wait()
    last = canv[seq[-1]]
    for a, b in zip(ref[last.scene.name][0], last.outputs()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    g.close()
