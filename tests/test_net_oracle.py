"""Pins the CPU restatement of the warm-start network (oracle/net_oracle.py):
its backward pass against central finite differences of its own forward pass
(exact float64 graph), the adjoint pairs it uses, and its BF16 rounding against
torch's conversion."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import net_oracle as NO  # noqa: E402


def _images(rng, shapes):
    out = []
    for R, C in shapes:
        x = rng.normal(size=(R, C, NO.NFEAT))
        x[:, :, 10:] = 0.0
        t = rng.normal(size=(R, C, 2))
        w = (rng.random((R, C)) > 0.3).astype(np.float64)
        out.append((x, t, w))
    return out


def _params(rng):
    parts = []
    for cin, cout, _, _ in NO.LAYERS:
        parts.append(rng.normal(scale=np.sqrt(2.0 / (9 * cin)), size=9 * cout * cin))
        parts.append(rng.normal(scale=0.05, size=cout))
    return np.concatenate(parts)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    imgs = _images(rng, [(9, 13), (6, 5)])
    p = _params(rng)
    loss, g = NO.loss_and_grads(p, imgs, bf=False)
    assert np.isfinite(loss) and loss > 0
    # a few parameters of every layer, weights and biases
    idx, k = [], 0
    for cin, cout, _, _ in NO.LAYERS:
        n = 9 * cout * cin
        idx += list(k + rng.choice(n, 3, replace=False)) + [k + n + int(rng.integers(cout))]
        k += n + cout
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (NO.loss_and_grads(pp, imgs, bf=False)[0] - NO.loss_and_grads(pm, imgs, bf=False)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-5 * max(abs(fd), 1e-3), (q, fd, g[q])


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(1)
    for R, C in [(7, 9), (8, 8), (1, 5), (3, 1)]:
        Rc, Cc = (R + 1) // 2, (C + 1) // 2
        x = rng.normal(size=(R, C, 3))
        y = rng.normal(size=(Rc, Cc, 3))
        assert np.isclose(np.sum(NO.pool(x) * y), np.sum(x * NO.pool_t(y, R, C)))
        assert np.isclose(np.sum(NO.up(y, R, C) * x), np.sum(y * NO.up_t(x, Rc, Cc)))


def test_transposed_weights_give_the_data_gradient():
    """<conv(x; w), g> = <x, conv(g; w^T)> (the data gradient as a convolution)."""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(6, 11, 16))
    g = rng.normal(size=(6, 11, 32))
    w = rng.normal(size=(9, 32, 16))
    assert np.isclose(np.sum(NO.conv(x, w, None) * g), np.sum(x * NO.conv(g, NO.transpose_weights(w), None)))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 4000),
                        np.float32([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, -0.0, 3.0e38])])
    ref = torch.from_numpy(a.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(NO.bf16(a), ref)
    assert NO.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0          # tie to even (down)
    assert NO.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6  # tie to even (up)


def test_package_network_matches_the_oracle_contract():
    """The layer table, parameter count and packed-parameter order of
    imau_vel.learned are the oracle's; the committed weights have that shape."""
    from imau_vel import learned
    assert [tuple(l) for l in learned.LAYERS] == [tuple(l) for l in NO.LAYERS]
    assert learned.n_params() == NO.n_params()
    p = learned.init_params(0)
    parts = NO.split_params(p)
    assert len(parts) == len(NO.LAYERS)
    assert np.all(parts[0][0][:, :, learned.USED_FEATURES:] == 0.0)   # unused inputs
    assert np.all(parts[-1][0][:, 2:, :] == 0.0)                       # unused outputs
    if os.path.exists(learned.WEIGHTS):
        with np.load(learned.WEIGHTS, allow_pickle=False) as z:
            assert z["params"].shape == (NO.n_params(),)
            assert np.array_equal(z["layers"], np.array(NO.LAYERS, np.int32))
            assert np.all(np.isfinite(z["params"]))


def test_conv_pool_up_match_torch_functional():
    """The forward operators against torch.nn.functional on ragged images
    (odd extents, 1-pixel rows and columns): conv = conv2d with zero padding 1
    (weights w[tap][cout][cin], tap = 3 (dr + 1) + (dc + 1), cross-correlation);
    pool = avg_pool2d(2) of the image zero-padded to even extents; up =
    nearest x2 upsampling cropped to the fine image."""
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(5)
    for R, C in [(7, 9), (8, 8), (1, 5), (3, 1), (13, 6)]:
        cin, cout = 5, 4
        x = rng.normal(size=(R, C, cin))
        w = rng.normal(size=(9, cout, cin))
        b = rng.normal(size=cout)
        xt = torch.from_numpy(x.transpose(2, 0, 1)[None].copy())
        wt = torch.from_numpy(w.reshape(3, 3, cout, cin).transpose(2, 3, 0, 1).copy())
        ref = F.conv2d(xt, wt, torch.from_numpy(b), padding=1)[0].numpy().transpose(1, 2, 0)
        np.testing.assert_allclose(NO.conv(x, w, b), ref, rtol=1e-12, atol=1e-12)
        xp = F.pad(xt, (0, C & 1, 0, R & 1))
        ref = F.avg_pool2d(xp, 2)[0].numpy().transpose(1, 2, 0)
        np.testing.assert_allclose(NO.pool(x), ref, rtol=1e-12, atol=1e-14)
        Rc, Cc = (R + 1) // 2, (C + 1) // 2
        c = rng.normal(size=(Rc, Cc, 3))
        ref = F.interpolate(torch.from_numpy(c.transpose(2, 0, 1)[None].copy()), scale_factor=2, mode="nearest")
        ref = ref[0, :, :R, :C].numpy().transpose(1, 2, 0)
        np.testing.assert_array_equal(NO.up(c, R, C), ref)
