"""CPU checks of the learned shader's numpy restatement (the checker of the
GPU forward and gradient tests): its backpropagation against central finite
differences of the same network in float64, its bf16 rounding against the
definition, and the encoding against closed forms."""
import numpy as np

from rsr_amd import learned


def _samples(n, seed):
    rng = np.random.RandomState(seed)
    s = np.zeros((n, learned.SAMPLE), np.float32)
    s[:, :13] = rng.uniform(-1, 1, (n, 13))
    s[:, 6:9] = np.array(learned.POS_CENTER) + rng.uniform(-20, 20, (n, 3))
    s[:, 13:16] = rng.uniform(0, 1, (n, 3))
    return s


def test_backprop_matches_finite_differences():
    s = _samples(16, 1)
    p = learned.init_params(2).astype(np.float64)
    rng = np.random.RandomState(3)
    p[learned.B1:learned.W2] = rng.uniform(-0.2, 0.2, learned.W2 - learned.B1)
    g = learned.exact_gradient(s, p)
    idx = np.concatenate([rng.choice(np.arange(learned.W1, learned.B1).reshape(64, 128)[:, :117].ravel(), 40),
                          rng.choice(np.arange(learned.B1, learned.TRAINABLE), 60)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h; pm[i] -= h
        fd = (learned.exact_loss(s, pp) - learned.exact_loss(s, pm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 3e38], np.float32)
    got = learned.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert list(got[:4]) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0]
    assert got[4] == -2.5 and got[5] == 0.0
    assert np.all(np.abs(learned.bf16(np.linspace(-4, 4, 1001, dtype=np.float32)) - np.linspace(-4, 4, 1001)) <= 4 * 2 ** -9)


def test_encoding_closed_forms():
    s = _samples(64, 4)
    p = learned.init_params(0)
    e = learned.encode(s, p)
    assert e.shape == (64, 128) and np.all(e[:, 117:] == 0)
    x = s[:, 0].astype(np.float64)
    for k in range(4):
        assert np.allclose(e[:, 1 + 2 * k], np.sin(2 ** k * np.pi * x), atol=2e-5)
        assert np.allclose(e[:, 2 + 2 * k], np.cos(2 ** k * np.pi * x), atol=2e-5)
    pos = (s[:, 6:9] - np.array(learned.POS_CENTER)) / learned.POS_HALF_EXTENT
    assert np.allclose(e[:, [54, 63, 72]], pos, atol=1e-6)


def test_psnr_closed_form():
    a = np.full((4, 4), 0x808080, np.uint32)
    b = a.copy(); b[0, 0] = 0x818080
    # one channel of 48 off by one: mse = 1/48
    assert abs(learned.psnr(a, b) - 10 * np.log10(255 ** 2 * 48)) < 1e-9
    assert learned.psnr(a, a) == float("inf")


def test_rounded_gradient_tracks_exact_gradient():
    # the bf16-rounded restatement (the GPU tests' checker) differs from the
    # exact gradient only by rounding: a transposed or dropped term would not
    s = _samples(256, 5)
    p = learned.init_params(6)
    rng = np.random.RandomState(7)
    p[learned.B1:learned.W2] = rng.uniform(-0.2, 0.2, learned.W2 - learned.B1)
    loss_r, gr = learned.ref_gradient(s, p)
    ge = learned.exact_gradient(s, p.astype(np.float64))
    assert abs(loss_r - learned.exact_loss(s, p.astype(np.float64))) < 0.02 * loss_r
    for a, b in zip(learned.layers(gr), learned.layers(ge)):
        assert np.linalg.norm(a - b) <= 0.05 * np.linalg.norm(b) + 1e-9
