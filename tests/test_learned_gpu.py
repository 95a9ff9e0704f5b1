"""GPU tests of the learned shader: the MFMA forward and the training
gradients against the numpy restatement (bf16 rounding at the same
operands, float64 accumulation), training on frames of seeded scenes, and
the learned frame against the plain heavy shader (coverage, depth and ids
untouched; PSNR; fallback)."""
import numpy as np
import pytest
import torch

from rsr_amd import learned, rgl, scenes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _samples(n, seed):
    rng = np.random.RandomState(seed)
    s = np.zeros((n, learned.SAMPLE), np.float32)
    nv = rng.normal(size=(n, 3)); nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    vv = rng.normal(size=(n, 3)); vv /= np.linalg.norm(vv, axis=1, keepdims=True)
    s[:, 0:3], s[:, 3:6] = nv, vv
    s[:, 6:9] = np.array(learned.POS_CENTER) + rng.uniform(-25, 25, (n, 3))
    s[:, 9:12] = rng.uniform(0, 1, (n, 3))
    s[:, 12] = rng.uniform(0.05, 1, n)
    s[:, 13:16] = rng.uniform(0, 1, (n, 3))
    return s


def _params(seed):
    p = learned.init_params(seed)
    rng = np.random.RandomState(seed + 100)
    # nonzero biases so that every layer has both live and dead ReLUs
    for a, b in ((learned.B1, learned.W2), (learned.B2, learned.W3), (learned.B3, learned.W4)):
        p[a:b] = rng.uniform(-0.3, 0.3, b - a)
    return p


def test_forward_matches_reference():
    s, p = _samples(3000, 1), _params(3)
    gpu = rgl.GPU(0)
    ds, dp = torch.from_numpy(s).to(DEV), torch.from_numpy(p).to(DEV)
    out = torch.zeros((len(s), 3), dtype=torch.float32, device=DEV)
    gpu.mlp_forward(ds.data_ptr(), len(s), dp.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want, _ = learned.ref_forward(s, p)
    err = np.abs(got - want)
    print(f"forward: max abs err {err.max():.3e}, mean {err.mean():.3e}, output scale {np.abs(want).mean():.3f}")
    assert err.mean() < 2e-3 and err.max() < 3e-2
    gpu.close()


def test_gradient_matches_reference():
    s, p = _samples(1024, 2), _params(4)
    gpu = rgl.GPU(0)
    ds, dp = torch.from_numpy(s).to(DEV), torch.from_numpy(p).to(DEV)
    g = torch.zeros(learned.TRAINABLE, dtype=torch.float32, device=DEV)
    loss = gpu.mlp_gradient(ds.data_ptr(), len(s), dp.data_ptr(), g.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = g.cpu().numpy().astype(np.float64)
    ref_loss, want = learned.ref_gradient(s, p)
    print(f"loss gpu {loss:.6f} ref {ref_loss:.6f}")
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    names = ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")
    for name, a, b in zip(names, learned.layers(got), learned.layers(want)):
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"{name}: |g| {np.linalg.norm(b):.4e} rel err {rel:.2e}")
        assert rel < 2e-2, name
    # the padded inputs of layer 1 get no gradient
    assert np.all(learned.layers(got)[0][:, learned.ENC:] == 0)
    # a training step moves the loss down on this batch
    opt = torch.zeros(2 * learned.TRAINABLE, dtype=torch.float32, device=DEV)
    for k in range(20):
        gpu.mlp_train(ds.data_ptr(), len(s), dp.data_ptr(), opt.data_ptr(), 1, 1024, 3e-3, 7, k,
                      torch.cuda.current_stream().cuda_stream)
    loss2 = gpu.mlp_gradient(ds.data_ptr(), len(s), dp.data_ptr(), g.data_ptr(), torch.cuda.current_stream().cuda_stream)
    print(f"loss after 20 Adam steps {loss2:.6f}")
    assert loss2 < 0.8 * loss
    gpu.close()


def _frame(gpu, sc, params=None, max_rms=2.0 / 255.0):
    gpu.set_learned_shading(params.data_ptr() if params is not None else None, max_rms)
    c, d, p, st = rgl.render(sc, gpu=gpu)
    return c, d, p, st


def test_learned_frame_and_training_gain():
    gpu = rgl.GPU(0)
    w, h = 640, 360
    samples = learned.collect_samples(gpu, list(range(100, 108)), 30000, w, h, DEV)
    n = samples.shape[0]
    print(f"{n} training samples")
    assert n > 100000
    s_host = samples.cpu().numpy()
    assert np.all(np.isfinite(s_host)) and s_host[:, 13:16].min() >= 0 and s_host[:, 13:16].max() <= 1
    p0 = torch.from_numpy(learned.init_params(0)).to(DEV)
    params = p0.clone()
    opt = torch.zeros(2 * learned.TRAINABLE, dtype=torch.float32, device=DEV)
    g = torch.zeros(learned.TRAINABLE, dtype=torch.float32, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    loss0 = gpu.mlp_gradient(samples.data_ptr(), 8192, params.data_ptr(), g.data_ptr(), stream)
    learned.train(gpu, samples, params, opt, epochs=6, batch=4096, lr=3e-3, lr_end=2e-4)
    loss1 = gpu.mlp_gradient(samples.data_ptr(), 8192, params.data_ptr(), g.data_ptr(), stream)
    print(f"training loss {loss0:.5f} -> {loss1:.5f}")
    assert loss1 < loss0 / 10

    sc = scenes.scene_heavy(w, h)                      # evaluation layout, not a training seed
    pc, pd, pp, pst = _frame(gpu, sc)                   # plain heavy shader
    cov = pd < 1.0
    lc, ld, lp, lst = _frame(gpu, sc, params, max_rms=1.0)     # network everywhere (no fallback)
    assert np.array_equal(pd.view(np.uint32), ld.view(np.uint32)) and np.array_equal(pp, lp)
    assert lst["fallback_quads"] == 0
    uc, _, _, _ = _frame(gpu, sc, p0, max_rms=1.0)             # the untrained network
    fc, fd, fp, fst = _frame(gpu, sc, params)                  # with the 2/255 fallback
    assert np.array_equal(pd.view(np.uint32), fd.view(np.uint32)) and np.array_equal(pp, fp)
    q_trained, q_untrained, q_fallback = learned.psnr(pc, lc, cov), learned.psnr(pc, uc, cov), learned.psnr(pc, fc, cov)
    print(f"PSNR over covered pixels: untrained {q_untrained:.2f} dB, trained {q_trained:.2f} dB, "
          f"trained + fallback {q_fallback:.2f} dB ({fst['fallback_quads']} of {4 * fst['tiles']} tile quadrants re-shaded)")
    assert q_trained > q_untrained + 10.0
    assert q_fallback >= q_trained
    # background pixels are the clear colour in every variant
    assert np.array_equal(pc[~cov], lc[~cov]) and np.array_equal(pc[~cov], fc[~cov])
    # a zero error budget re-shades every tile whose probes saw any error
    zc, _, _, zst = _frame(gpu, sc, params, max_rms=0.0)
    same = sum(np.array_equal(zc[y0:y1, x0:x1], pc[y0:y1, x0:x1])
               for _, _, (x0, y0, x1, y1) in _quadrants(w, h, sc.tile_blocks[0] * 8, sc.tile_blocks[1] * 8))
    print(f"zero budget: {zst['fallback_quads']} tile quadrants re-shaded, {same} equal to the plain frame")
    assert zst["fallback_quads"] > 0 and same >= zst["fallback_quads"]
    gpu.set_learned_shading(None, 0.0)
    gpu.close()


def _quadrants(w, h, tw, th):
    # (tile, quadrant, pixel rect) of every tile quadrant of a w x h target:
    # a tile splits where its 8 x 8 probe cells do
    tiles_x = (w + tw - 1) // tw
    for ty in range((h + th - 1) // th):
        for tx in range(tiles_x):
            x0, y0 = tx * tw, ty * th
            x1, y1 = min(x0 + tw, w), min(y0 + th, h)
            cw, ch = (x1 - x0 + 7) // 8, (y1 - y0 + 7) // 8
            for q in range(4):
                qx0, qx1 = (x0 + 4 * cw, x1) if q & 1 else (x0, min(x0 + 4 * cw, x1))
                qy0, qy1 = (y0 + 4 * ch, y1) if q >> 1 else (y0, min(y0 + 4 * ch, y1))
                yield ty * tiles_x + tx, q, (qx0, qy0, qx1, qy1)


def _hash32(v):
    v &= 0xffffffff
    v ^= v >> 16; v = (v * 0x7feb352d) & 0xffffffff
    v ^= v >> 15; v = (v * 0x846ca68b) & 0xffffffff
    v ^= v >> 16
    return v


def _probe(x0, y0, x1, y1, tile, j):
    # probe j of a tile: one jittered pixel in each cell of an 8 x 8 grid
    cw, ch = (x1 - x0 + 7) // 8, (y1 - y0 + 7) // 8
    h = _hash32((((tile * 64 + j) * 0x9e3779b9) & 0xffffffff) + 17)
    x = x0 + (j & 7) * cw + h % cw
    y = y0 + (j >> 3) * ch + (h >> 16) % ch
    return (x, y) if x < x1 and y < y1 else None


@pytest.mark.parametrize("max_rms,tile", [(2.0 / 255.0, 8), (0.01, 8), (6.0 / 255.0, 8), (0.01, 4), (2.0 / 255.0, 4)])
def test_fallback_tiles_follow_the_probes(max_rms, tile):
    # every tile whose covered probes show a mean squared 8-bit error over
    # the limit (the network's colours against the plain shader's) is the
    # plain frame's tile, every other tile the network's, bit for bit; the
    # count is of re-shaded tile quadrants, four per tile
    import os
    wpath = os.path.join(os.path.dirname(learned.__file__), "ggx_mlp_weights.npy")
    params = torch.from_numpy(np.load(wpath, allow_pickle=False).astype(np.float32)).to(DEV)
    import dataclasses
    gpu = rgl.GPU(0)
    # 64 and 32 px tiles (360 rows leave 40 and 8 px tiles at the bottom,
    # whose quadrants differ in size)
    sc = dataclasses.replace(scenes.scene_heavy(640, 360), tile_blocks=(tile, tile))
    tw, th = sc.tile_blocks[0] * 8, sc.tile_blocks[1] * 8
    pc, pd, _, _ = _frame(gpu, sc)
    nc, _, _, _ = _frame(gpu, sc, params, max_rms=1.0)
    fc, _, _, fst = _frame(gpu, sc, params, max_rms=max_rms)
    lim = np.float32(max_rms) * np.float32(255.0)
    max_mse = lim * lim
    ch = lambda c: np.array([(int(c) >> s) & 255 for s in (0, 8, 16)], np.int64)
    tiles_x = (sc.width + tw - 1) // tw
    over_quads = 0
    for ty in range(0, (sc.height + th - 1) // th):
        for tx in range(tiles_x):
            t = ty * tiles_x + tx
            x0, y0 = tx * tw, ty * th
            x1, y1 = min(x0 + tw, sc.width), min(y0 + th, sc.height)
            se = cnt = 0
            for j in range(64):
                p = _probe(x0, y0, x1, y1, t, j)
                if p is None or not pd[p[1], p[0]] < 1.0:
                    continue
                se += int(((ch(pc[p[1], p[0]]) - ch(nc[p[1], p[0]])) ** 2).sum())
                cnt += 3
            over = cnt > 0 and np.float32(se) / np.float32(cnt) > max_mse
            over_quads += 4 * int(over)
            want = (pc if over else nc)[y0:y1, x0:x1]
            assert np.array_equal(fc[y0:y1, x0:x1], want), (tx, ty, over, se, cnt)
    print(f"limit {max_rms * 255:.2f}/255: {over_quads // 4} tiles re-shaded")
    assert fst["fallback_quads"] == over_quads
    gpu.set_learned_shading(None, 0.0)
    gpu.close()


def test_config4_1080p_shipped_weights():
    # the learned-shading workload at its full size with the shipped
    # network: coverage, depth and ids identical to the plain heavy frame,
    # and >= 40 dB over covered pixels with the 2/255 per-tile fallback
    import os
    wpath = os.path.join(os.path.dirname(learned.__file__), "ggx_mlp_weights.npy")
    params = torch.from_numpy(np.load(wpath, allow_pickle=False).astype(np.float32)).to(DEV)
    gpu = rgl.GPU(0)
    sc = scenes.scene_heavy()
    pc, pd, pp, pst = _frame(gpu, sc)
    fc, fd, fp, fst = _frame(gpu, sc, params)
    nc, nd, npid, nst = _frame(gpu, sc, params, max_rms=1.0)
    cov = pd < 1.0
    assert np.array_equal(pd.view(np.uint32), fd.view(np.uint32)) and np.array_equal(pp, fp)
    assert np.array_equal(pd.view(np.uint32), nd.view(np.uint32)) and np.array_equal(pp, npid)
    q, qn = learned.psnr(pc, fc, cov), learned.psnr(pc, nc, cov)
    print(f"1080p heavy: PSNR covered {q:.2f} dB with fallback ({fst['fallback_quads']} of {4 * fst['tiles']} tile quadrants), "
          f"network alone {qn:.2f} dB")
    assert q >= 40.0
    gpu.set_learned_shading(None, 0.0)
    gpu.close()


def test_config5_view_learned_shading():
    # the stress workload's learned leg on one full 3840x2160 view (2M
    # triangles, depth complexity ~8): coverage, depth and ids identical to
    # the plain heavy frame and >= 40 dB over covered pixels with fallback
    import os
    wpath = os.path.join(os.path.dirname(learned.__file__), "ggx_mlp_weights.npy")
    params = torch.from_numpy(np.load(wpath, allow_pickle=False).astype(np.float32)).to(DEV)
    gpu = rgl.GPU(0)
    sc = scenes.scene_stress(1)
    print(f"[{sc.name} learned]", flush=True)
    pc, pd, pp, pst = _frame(gpu, sc)
    fc, fd, fp, fst = _frame(gpu, sc, params)
    cov = pd < 1.0
    assert cov.sum() > 100000
    assert np.array_equal(pd.view(np.uint32), fd.view(np.uint32)) and np.array_equal(pp, fp)
    q = learned.psnr(pc, fc, cov)
    print(f"4K stress view: PSNR covered {q:.2f} dB ({fst['fallback_quads']} of {4 * fst['tiles']} tile quadrants re-shaded)")
    assert q >= 40.0
    gpu.set_learned_shading(None, 0.0)
    gpu.close()


def test_learned_shading_keeps_linear_and_exposure_stores_plain():
    # the network predicts the default post program's sRGB colour: frames
    # stored linear or through the Exposure post shade with the plain
    # shader even when learned shading is on
    import dataclasses
    import os
    wpath = os.path.join(os.path.dirname(learned.__file__), "ggx_mlp_weights.npy")
    params = torch.from_numpy(np.load(wpath, allow_pickle=False).astype(np.float32)).to(DEV)
    gpu = rgl.GPU(0)
    sc = dataclasses.replace(scenes.scene_heavy(640, 360), post_program=scenes.PGM_EXPOSURE_POST)
    sc.uniforms = sc.uniforms.copy()
    sc.uniforms[0] = 1.3          # the Exposure post reads uniforms[0] (here also the eye's x)
    plain = _frame(gpu, sc)
    assert (plain[1] < 1.0).sum() > 1000 and len(np.unique(plain[0])) > 100
    lrn = _frame(gpu, sc, params, max_rms=1.0)
    assert lrn[3]["fallback_quads"] == 0
    assert np.array_equal(plain[0], lrn[0])
    gpu.set_learned_shading(None, 0.0)
    gpu.close()


def test_network_call_on_another_stream_orders_against_frames():
    # a frame on stream A, a network forward on stream B, the frame again on
    # stream B: both frames equal the oracle's
    from oracle_backend import render as oracle_render
    sc = scenes.scene_heavy(640, 360)
    ref = oracle_render(sc, threads=8)
    gpu = rgl.GPU(0)
    ds = rgl.DeviceScene(sc, DEV, prim=True)
    ds.record(gpu)
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    gpu.run(a.cuda_stream)
    s, p = _samples(5000, 5), _params(6)
    dsm, dp = torch.from_numpy(s).to(DEV), torch.from_numpy(p).to(DEV)
    out = torch.zeros((len(s), 3), dtype=torch.float32, device=DEV)
    gpu.mlp_forward(dsm.data_ptr(), len(s), dp.data_ptr(), out.data_ptr(), b.cuda_stream)
    a.synchronize()
    first = [t.cpu().clone() for t in (ds.color, ds.depth, ds.prim)]
    gpu.run(b.cuda_stream)
    gpu.wait()
    torch.cuda.synchronize()
    for k, (x, y) in enumerate(zip(first, (ds.color, ds.depth, ds.prim))):
        assert np.array_equal(x.numpy().view(np.uint32), ref[k].view(np.uint32)), f"frame on stream A, buffer {k}"
        assert np.array_equal(y.cpu().numpy().view(np.uint32), ref[k].view(np.uint32)), f"frame on stream B, buffer {k}"
    want, _ = learned.ref_forward(s, p)
    assert np.abs(out.cpu().numpy() - want).mean() < 2e-3
    gpu.close()


def test_learned_frame_graph_replay_matches_eager():
    # a learned frame with the fallback forks its probes onto a second
    # stream and joins them before the fallback: replayed from a captured
    # graph (5 times, with eager frames of another program between) it
    # stores exactly the eager frame
    import os
    wpath = os.path.join(os.path.dirname(__file__), "..", "rsr_amd", "ggx_mlp_weights.npy")
    params = torch.from_numpy(np.load(wpath, allow_pickle=False).astype(np.float32)).to(DEV)
    sc = scenes.scene_heavy(640, 360)
    gpu = rgl.GPU(0)
    ec, ed, ep, est = _frame(gpu, sc, params)
    print(f"eager: {est['fallback_quads']} of {4 * est['tiles']} tile quadrants re-shaded", flush=True)
    assert 0 < est["fallback_quads"] < 4 * est["tiles"]
    other = rgl.GPU(0)
    oc, od, op, _ = rgl.render(scenes.scene_sphere_ground(), gpu=other)
    gpu.set_graph(True)
    ds = rgl.DeviceScene(sc, DEV, prim=True)
    ds.record(gpu)
    stream = torch.cuda.current_stream(DEV).cuda_stream
    for k in range(5):
        ds.color.zero_()
        torch.cuda.synchronize()
        gpu.run_and_wait(stream)
        c, d, p = ds.outputs()
        assert np.array_equal(c, ec) and np.array_equal(d.view(np.uint32), ed.view(np.uint32)) and np.array_equal(p, ep), k
        assert gpu.stats()["fallback_quads"] == est["fallback_quads"]
        c2, _, _, _ = rgl.render(scenes.scene_sphere_ground(), gpu=other)
        assert np.array_equal(c2, oc)
    assert gpu.graph_captures() == 1
    gpu.set_graph(False)
    gpu.set_learned_shading(None, 0.0)
    other.close()
    gpu.close()
