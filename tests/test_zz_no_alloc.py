"""GPU test of rgl_mlp_train_noalloc, the training entry point that
allocates nothing: through the C ABI its parameters and optimiser state
equal rgl_mlp_train's bit for bit on two calls in a row, the second with a
smaller batch, and a batch beyond its scratch limit is refused."""
import ctypes

import numpy as np
import pytest
import torch

from rsr_amd import rgl

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _samples(n, seed):
    # 16 floats per sample: unit normal and view vector, position, albedo,
    # roughness, target colour
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 16), np.float32)
    nv = rng.normal(size=(n, 3)); nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    vv = rng.normal(size=(n, 3)); vv /= np.linalg.norm(vv, axis=1, keepdims=True)
    s[:, 0:3], s[:, 3:6] = nv, vv
    s[:, 6:9] = rng.uniform(-20, 20, (n, 3))
    s[:, 9:12] = rng.uniform(0, 1, (n, 3))
    s[:, 12] = rng.uniform(0.05, 1, n)
    s[:, 13:16] = rng.uniform(0, 1, (n, 3))
    return torch.from_numpy(s).to(DEV)


def test_noalloc_training_matches_allocating_training():
    L = rgl.lib()
    n = 12000
    samples = _samples(n, 3)
    P = L.rgl_mlp_param_count()
    T = L.rgl_mlp_trainable_count()
    rng = np.random.default_rng(4)
    p0 = (rng.normal(size=P) * 0.05).astype(np.float32)
    p0[T:T + 4] = (0.0, 0.0, 0.0, 1.0 / 20.0)        # position centre and 1 / half-extent
    pa = torch.from_numpy(p0.copy()).to(DEV)
    pb = torch.from_numpy(p0.copy()).to(DEV)
    oa = torch.zeros(2 * T, dtype=torch.float32, device=DEV)
    ob = torch.zeros(2 * T, dtype=torch.float32, device=DEV)
    loss_b = torch.zeros(1, dtype=torch.float32, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    ga, gb = rgl.GPU(0), rgl.GPU(0)
    for call, (steps, batch, step0) in enumerate(((3, 4096, 0), (2, 1024, 3))):
        loss_a = ga.mlp_train(samples.data_ptr(), n, pa.data_ptr(), oa.data_ptr(), steps, batch, 1e-3, 77, step0, stream)
        gb.mlp_train_noalloc(samples.data_ptr(), n, pb.data_ptr(), ob.data_ptr(), steps, batch, 1e-3, 77, step0,
                             loss_b.data_ptr(), stream)
        torch.cuda.synchronize()
        assert torch.equal(pa.view(torch.int32), pb.view(torch.int32)), f"call {call}: parameters differ"
        assert torch.equal(oa.view(torch.int32), ob.view(torch.int32)), f"call {call}: optimiser state differs"
        # the loss sums per-wave partials with atomics: equal to rounding
        assert abs(float(loss_b.item()) - loss_a) <= 1e-5 * max(1.0, abs(loss_a)), (call, loss_a, float(loss_b.item()))
        assert not torch.equal(pa, torch.from_numpy(p0).to(DEV))
    # a batch beyond the scratch array is refused before anything runs
    with pytest.raises(rgl.RglError, match="4096"):
        gb.mlp_train_noalloc(samples.data_ptr(), n, pb.data_ptr(), ob.data_ptr(), 1, 4097, 1e-3, 77, 5, None, stream)
    ga.close()
    gb.close()
