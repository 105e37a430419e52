"""GPU parity of the warm-start network kernels (include/imau_net.h) against the
CPU restatement oracle/net_oracle.py, through the C ABI.

Tolerances: a BF16 output may differ from the oracle's by one BF16 step
(relative 2^-7) because the GPU accumulates in FP32 and the oracle in float64
before the same rounding; FP32 weight gradients of one layer agree to 1e-4
relative (summation order); the gradients of the whole network, where one-step
BF16 differences of intermediate tensors propagate, agree to 3e-2 relative per
layer (norm)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import net_oracle as NO  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (5, 37), (33, 17), (4, 16), (12, 9)]


@pytest.fixture(scope="module")
def env():
    import torch
    from imau_vel import learned, solver
    ctx = solver.Context()
    yield torch, learned, ctx
    ctx.close()


def bits(a):
    """float array -> BF16 bit patterns (int16), rounded to nearest even."""
    r = NO.bf16(a).astype(np.float32).view(np.uint32) >> 16
    return r.astype(np.uint16).view(np.int16)


def unbits(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def level_shapes(shapes, l):
    out = []
    for r, c in shapes:
        for _ in range(l):
            r, c = (r + 1) // 2, (c + 1) // 2
        out.append((r, c))
    return out


def pack_images(imgs):
    return np.concatenate([a.reshape(-1) for a in imgs])


def split_images(flat, shapes, C):
    out, k = [], 0
    for r, c in shapes:
        out.append(flat[k:k + r * c * C].reshape(r, c, C))
        k += r * c * C
    return out


def close_bf16(got, ref, what):
    tol = 2.0 ** -7 * np.abs(ref) + 1e-30
    bad = np.abs(got - ref) > tol
    assert not bad.any(), "%s: %d of %d differ by more than one BF16 step, worst %g vs %g" % (
        what, bad.sum(), bad.size, got[bad][0], ref[bad][0])


@pytest.mark.parametrize("cin,cout,level", [(16, 32, 0), (32, 32, 0), (32, 32, 2), (32, 16, 0)])
def test_conv_matches_oracle(env, cin, cout, level):
    torch, learned, ctx = env
    rng = np.random.default_rng(cin + cout + level)
    plan = learned.Plan(ctx, rows=[r for r, _ in SHAPES], cols=[c for _, c in SHAPES])
    shp = level_shapes(SHAPES, level)
    xs = [rng.normal(size=(r, c, cin)) for r, c in shp]
    w = rng.normal(scale=0.2, size=(9, cout, cin)).astype(np.float32)
    b = rng.normal(scale=0.1, size=cout).astype(np.float32)
    x_d = dev(torch, bits(pack_images(xs)))
    w_d, b_d = dev(torch, w.ravel()), dev(torch, b)
    wf = torch.zeros(w.size, dtype=torch.int16, device="cuda")
    wt = torch.zeros(w.size, dtype=torch.int16, device="cuda")
    L = learned._bind(ctx.L)
    ctx._check(L.imau_net_pack(ctx.h, cin, cout, learned._ptr(w_d), learned._ptr(wf), learned._ptr(wt)))
    npix = plan.npix[level]
    assert npix == sum(r * c for r, c in shp)
    wq = NO.bf16(w)
    if cout == 16:   # the last layer: FP32 output of channels 0, 1, no ReLU
        y = torch.zeros(npix * 2, dtype=torch.float32, device="cuda")
        ctx._check(L.imau_net_conv(ctx.h, plan.h, level, cin, cout, learned._ptr(x_d), learned._ptr(wf),
                                   learned._ptr(b_d), 0, learned._ptr(y), 1))
        ctx._check(L.imau_net_sync(ctx.h))
        got = split_images(y.cpu().numpy().astype(np.float64), shp, 2)
        for g, x in zip(got, xs):
            ref = NO.conv(NO.bf16(x), wq, b)[:, :, :2]
            assert np.allclose(g, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    else:
        y = torch.zeros(npix * cout, dtype=torch.int16, device="cuda")
        ctx._check(L.imau_net_conv(ctx.h, plan.h, level, cin, cout, learned._ptr(x_d), learned._ptr(wf),
                                   learned._ptr(b_d), 1, learned._ptr(y), 0))
        ctx._check(L.imau_net_sync(ctx.h))
        got = split_images(unbits(y), shp, cout)
        for g, x in zip(got, xs):
            close_bf16(g, NO.bf16(np.maximum(NO.conv(NO.bf16(x), wq, b), 0.0)), "conv %d->%d" % (cin, cout))
    # the packed transpose is the data-gradient weight set
    assert np.array_equal(unbits(wt).reshape(9, cin, cout), NO.transpose_weights(wq))
    plan.close()


@pytest.mark.parametrize("cin,cout,level", [(16, 32, 0), (32, 32, 1), (32, 16, 0)])
def test_conv_grad_w_matches_oracle(env, cin, cout, level):
    torch, learned, ctx = env
    rng = np.random.default_rng(7 + cin + cout)
    plan = learned.Plan(ctx, rows=[r for r, _ in SHAPES], cols=[c for _, c in SHAPES])
    shp = level_shapes(SHAPES, level)
    xs = [NO.bf16(rng.normal(size=(r, c, cin))) for r, c in shp]
    gs = [NO.bf16(rng.normal(size=(r, c, cout))) for r, c in shp]
    L = learned._bind(ctx.L)
    work = torch.zeros(int(L.imau_net_grad_w_work(plan.h, level, cin, cout)), dtype=torch.float32, device="cuda")
    dw = torch.zeros(9 * cout * cin, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    g_d, x_d = dev(torch, bits(pack_images(gs))), dev(torch, bits(pack_images(xs)))
    ctx._check(L.imau_net_conv_grad_w(ctx.h, plan.h, level, cin, cout, learned._ptr(g_d), learned._ptr(x_d),
                                      learned._ptr(dw), learned._ptr(db), learned._ptr(work)))
    ctx._check(L.imau_net_sync(ctx.h))
    rw, rb = np.zeros((9, cout, cin)), np.zeros(cout)
    for g, x in zip(gs, xs):
        a, b = NO.conv_grad_w(g, x)
        rw += a
        rb += b
    assert np.linalg.norm(dw.cpu().numpy().reshape(rw.shape) - rw) <= 1e-4 * np.linalg.norm(rw)
    assert np.allclose(db.cpu().numpy(), rb, rtol=1e-4, atol=1e-4 * np.abs(rb).max())
    plan.close()


def test_pool_up_and_their_gradients_match_oracle(env):
    torch, learned, ctx = env
    rng = np.random.default_rng(11)
    plan = learned.Plan(ctx, rows=[r for r, _ in SHAPES], cols=[c for _, c in SHAPES])
    L = learned._bind(ctx.L)
    P = learned._ptr
    for l in range(3):
        fs, cs = level_shapes(SHAPES, l), level_shapes(SHAPES, l + 1)
        nf, nc = plan.npix[l], plan.npix[l + 1]
        x = [NO.bf16(rng.normal(size=(r, c, 32))) for r, c in fs]
        b = [NO.bf16(rng.normal(size=(r, c, 32))) for r, c in cs]
        yf = [NO.bf16(rng.normal(size=(r, c, 32))) for r, c in fs]   # ReLU outputs: sign pattern only
        yc = [NO.bf16(rng.normal(size=(r, c, 32))) for r, c in cs]
        x_d, b_d = dev(torch, bits(pack_images(x))), dev(torch, bits(pack_images(b)))
        yf_d, yc_d = dev(torch, bits(pack_images(yf))), dev(torch, bits(pack_images(yc)))
        out_c = torch.zeros(nc * 32, dtype=torch.int16, device="cuda")
        out_f = torch.zeros(nf * 32, dtype=torch.int16, device="cuda")
        gp_f = torch.zeros(nf * 32, dtype=torch.int16, device="cuda")
        gu_c = torch.zeros(nc * 32, dtype=torch.int16, device="cuda")
        ctx._check(L.imau_net_pool(ctx.h, plan.h, l, P(x_d), P(out_c)))
        ctx._check(L.imau_net_upadd(ctx.h, plan.h, l, P(x_d), P(b_d), P(out_f)))
        ctx._check(L.imau_net_grad_pool(ctx.h, plan.h, l, P(x_d), P(b_d), P(yf_d), P(gp_f)))
        ctx._check(L.imau_net_grad_up(ctx.h, plan.h, l, P(x_d), P(yc_d), P(gu_c)))
        ctx._check(L.imau_net_sync(ctx.h))
        pc = split_images(unbits(out_c), cs, 32)
        uf = split_images(unbits(out_f), fs, 32)
        gpf = split_images(unbits(gp_f), fs, 32)
        guc = split_images(unbits(gu_c), cs, 32)
        for k, ((r, c), (rc, cc)) in enumerate(zip(fs, cs)):
            close_bf16(pc[k], NO.bf16(NO.pool(x[k])), "pool")
            close_bf16(uf[k], NO.bf16(x[k] + NO.up(b[k], r, c)), "upadd")
            close_bf16(gpf[k], NO.bf16(x[k] + NO.pool_t(b[k], r, c)) * (yf[k] > 0), "grad_pool")
            close_bf16(guc[k], NO.bf16(NO.up_t(x[k], rc, cc)) * (yc[k] > 0), "grad_up")
    plan.close()


def _images(rng, shapes):
    out = []
    for r, c in shapes:
        x = rng.normal(size=(r, c, 16))
        x[:, :, 10:] = 0.0
        out.append((NO.bf16(x), rng.normal(size=(r, c, 2)), (rng.random((r, c)) > 0.2).astype(np.float64)))
    return out


def test_network_forward_and_gradients_match_oracle(env):
    torch, learned, ctx = env
    rng = np.random.default_rng(5)
    shapes = [(37, 23), (16, 16), (9, 50)]
    imgs = _images(rng, shapes)
    net = learned.Net(ctx, seed=3)
    p = net.params_numpy()
    plan = learned.Plan(ctx, rows=[r for r, _ in shapes], cols=[c for _, c in shapes])
    x0 = dev(torch, bits(pack_images([x for x, _, _ in imgs])))
    t = dev(torch, pack_images([t for _, t, _ in imgs]).astype(np.float32))
    w = dev(torch, pack_images([w for _, _, w in imgs]).astype(np.float32))
    scale = 1.0 / sum(float(w.sum()) for _, _, w in imgs)
    out = net.forward(plan, x0, train=True)
    loss = net.backward(plan, x0, t, w, scale)
    net.sync()
    got_out = split_images(out.cpu().numpy().astype(np.float64), shapes, 2)
    params = NO.split_params(p)
    for g, (x, _, _) in zip(got_out, imgs):
        ref, _ = NO.forward(params, x, bf=True)
        assert np.linalg.norm(g - ref) <= 1e-2 * np.linalg.norm(ref)
    ref_loss, ref_g = NO.loss_and_grads(p, imgs, bf=True)
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss
    got_g = net.grad.cpu().numpy().astype(np.float64)
    k = 0
    for i, (cin, cout, _, _) in enumerate(NO.LAYERS):
        n = 9 * cout * cin + cout
        a, b = got_g[k:k + n], ref_g[k:k + n]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-12, (i, np.linalg.norm(a - b), np.linalg.norm(b))
        k += n
    net.release(plan)
    plan.close()


def test_adam_matches_oracle(env):
    torch, learned, ctx = env
    rng = np.random.default_rng(9)
    n = 1000
    p, g = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    m, v = (rng.normal(size=n) * 0.1).astype(np.float32), (rng.random(n) * 0.01).astype(np.float32)
    d = [dev(torch, a.copy()) for a in (p, g, m, v)]
    L = learned._bind(ctx.L)
    P = learned._ptr
    ctx._check(L.imau_net_adam(ctx.h, n, P(d[0]), P(d[1]), P(d[2]), P(d[3]), 1e-3, 0.9, 0.999, 1e-8, 3))
    ctx._check(L.imau_net_sync(ctx.h))
    rp, rm, rv = NO.adam(p, g, m, v, 1e-3, 0.9, 0.999, 1e-8, 3)
    assert np.allclose(d[0].cpu().numpy(), rp, rtol=1e-6, atol=1e-7)
    # moments: one FP32 rounding of the inputs' scale (the GPU fuses b1 m + (1 - b1) g)
    assert np.allclose(d[2].cpu().numpy(), rm, rtol=1e-6, atol=1e-7)
    assert np.allclose(d[3].cpu().numpy(), rv, rtol=1e-6, atol=2e-9)


def test_training_steps_reduce_the_loss(env):
    """Adam through the GPU kernels lowers the loss of a fixed batch."""
    torch, learned, ctx = env
    rng = np.random.default_rng(21)
    shapes = [(40, 40), (40, 40)]
    imgs = []
    for r, c in shapes:
        x = rng.normal(size=(r, c, 16))
        x[:, :, 10:] = 0.0
        t = np.stack([0.5 * x[:, :, 8] + 0.2, -0.3 * x[:, :, 9]], axis=-1)   # learnable from the input
        imgs.append((NO.bf16(x), t, np.ones((r, c))))
    net = learned.Net(ctx, seed=1)
    plan = learned.Plan(ctx, rows=[r for r, _ in shapes], cols=[c for _, c in shapes])
    x0 = dev(torch, bits(pack_images([x for x, _, _ in imgs])))
    t = dev(torch, pack_images([t for _, t, _ in imgs]).astype(np.float32))
    w = dev(torch, pack_images([w for _, _, w in imgs]).astype(np.float32))
    scale = 1.0 / (2 * 40 * 40)
    losses = []
    for _ in range(40):
        net.forward(plan, x0, train=True)
        losses.append(net.backward(plan, x0, t, w, scale))
        net.adam_step(3e-3)
    assert losses[-1] < 0.6 * losses[0], losses[::8]
    assert all(np.isfinite(losses))
    net.release(plan)
    plan.close()
