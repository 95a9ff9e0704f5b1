"""Trains the learned shader on the GPU and saves its parameter block.

    python -m rsr_amd.train_learned --out rsr_amd/ggx_mlp_weights.npy

Training data: covered fragments of seeded random scenes (instance layouts
100..100+scenes-1 under the heavy shader, 1920x1080), each fragment's 13
shading inputs and the plain heavy shader's sRGB colour, collected by the
pipeline itself (rgl_gl_store_training_samples).  The network is trained by
the library's own forward and backward kernels (rgl_mlp_train).  Prints
one progress line per chunk of steps, and the PSNR of the evaluation scene
(layout 2, not a training layout) against the plain shader at the end.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(ROOT, "rsr_amd", "ggx_mlp_weights.npy"))
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--per-scene", type=int, default=125000)
    ap.add_argument("--epochs", type=float, default=40.0)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lr-end", type=float, default=5e-5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    args = ap.parse_args()

    import torch
    from rsr_amd import learned, rgl, scenes
    dev = "cuda:0"
    gpu = rgl.GPU(0)
    t0 = time.time()
    samples = learned.collect_samples(gpu, list(range(100, 100 + args.scenes)), args.per_scene,
                                      args.width, args.height, dev)
    torch.cuda.synchronize()
    print(f"collected {samples.shape[0]} samples from {args.scenes} scenes in {time.time() - t0:.1f} s", flush=True)
    params = torch.from_numpy(learned.init_params(args.seed)).to(dev)
    opt = torch.zeros(2 * learned.TRAINABLE, dtype=torch.float32, device=dev)
    t1 = time.time()

    def log(step, total, loss):
        if step % 512 == 0 or step == total:
            rmse = float(np.sqrt(max(loss, 0.0))) * 255.0
            print(f"step {step}/{total} batch loss {loss:.3e} (rms {rmse:.2f}/255) {time.time() - t1:.1f} s", flush=True)

    learned.train(gpu, samples, params, opt, epochs=args.epochs, batch=args.batch, lr=args.lr, lr_end=args.lr_end,
                  seed=args.seed + 1, log=log)
    torch.cuda.synchronize()
    print(f"trained in {time.time() - t1:.1f} s", flush=True)
    p = params.cpu().numpy()
    np.save(args.out, p)
    # evaluation scene against the plain shader
    sc = scenes.scene_heavy(args.width, args.height)
    gpu.set_learned_shading(None, 0.0)
    pc, pd, _, _ = rgl.render(sc, gpu=gpu)
    gpu.set_learned_shading(params.data_ptr(), 1.0)
    lc, _, _, _ = rgl.render(sc, gpu=gpu)
    gpu.set_learned_shading(params.data_ptr(), 2.0 / 255.0)
    fc, _, _, st = rgl.render(sc, gpu=gpu)
    cov = pd < 1.0
    print(f"evaluation PSNR (covered pixels): network {learned.psnr(pc, lc, cov):.2f} dB, "
          f"with 2/255 fallback {learned.psnr(pc, fc, cov):.2f} dB ({st['fallback_quads']}/{4 * st['tiles']} tile quadrants); "
          f"saved {args.out}", flush=True)
    gpu.set_learned_shading(None, 0.0)
    gpu.close()


if __name__ == "__main__":
    main()
