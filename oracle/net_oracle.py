"""CPU restatement of the warm-start network (numpy, float64 arithmetic).

TEST INFRASTRUCTURE ONLY.  This is the checker the network kernels of
libimauvel.so (include/imau_net.h) are compared against; nothing in the product
imports it.  Only tests/, __graft_entry__.smoke() and bench.py may use it.

The network has no counterpart in the reference (IMAU-ICE starts each Picard
loop from its previous velocities, src/ice_velocity_module.f90:249-374), so the
definition below is the contract itself:

  e0 = relu(conv(x0; W0))          level 0, 16 -> 32 channels
  e0b = relu(conv(e0; W1))         level 0
  e1 = relu(conv(pool(e0b); W2))   level 1
  e2 = relu(conv(pool(e1); W3))    level 2
  e3 = relu(conv(pool(e2); W4)); e3b = relu(conv(e3; W5))   level 3
  d2 = relu(conv(e2 + up(e3b); W6))    level 2
  d1 = relu(conv(e1 + up(d2); W7))     level 1
  d0 = relu(conv(e0b + up(d1); W8))    level 0
  out = conv(d0; W9)[:, :, 0:2]        level 0, 32 -> 16 channels, FP32

conv is a 3x3 convolution with zero padding, weights [tap][cout][cin] with
tap = (dr + 1) * 3 + (dc + 1), plus a bias; pool is the 2x2 mean with zeros
outside the image (extents halve, rounded up); up repeats each pixel 2x2.

With `bf16=True` every tensor the GPU stores in BF16 is rounded to BF16
(round to nearest, ties to even) at the same place: the input, the weights,
every activation after its bias and ReLU, pooled and summed tensors, and every
gradient tensor after the operation that produces it.  Weight gradients stay
unrounded (FP32 on the GPU).  With `bf16=False` the same graph is exact float64
arithmetic, which the finite-difference test pins.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, level, relu) of the ten convolutions
LAYERS = [(16, 32, 0, True), (32, 32, 0, True), (32, 32, 1, True), (32, 32, 2, True), (32, 32, 3, True),
          (32, 32, 3, True), (32, 32, 2, True), (32, 32, 1, True), (32, 32, 0, True), (32, 16, 0, False)]
NFEAT = 16
NOUT = 16


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _r(a, on):
    return bf16(a) if on else a


def conv(x, w, b):
    """3x3 convolution, zero padding: x [R,C,cin], w [9,cout,cin], b [cout] -> [R,C,cout]."""
    R, C, _ = x.shape
    xp = np.zeros((R + 2, C + 2, x.shape[2]))
    xp[1:-1, 1:-1] = x
    y = np.zeros((R, C, w.shape[1]))
    for tap in range(9):
        dr, dc = tap // 3 - 1, tap % 3 - 1
        y += xp[1 + dr:1 + dr + R, 1 + dc:1 + dc + C] @ w[tap].T
    return y + (b if b is not None else 0.0)


def conv_grad_w(g, x):
    """dw[tap][co][ci] = sum_p g[p][co] x[p + tap][ci]; db[co] = sum_p g[p][co]."""
    R, C, _ = x.shape
    xp = np.zeros((R + 2, C + 2, x.shape[2]))
    xp[1:-1, 1:-1] = x
    dw = np.zeros((9, g.shape[2], x.shape[2]))
    for tap in range(9):
        dr, dc = tap // 3 - 1, tap % 3 - 1
        xs = xp[1 + dr:1 + dr + R, 1 + dc:1 + dc + C]
        dw[tap] = g.reshape(-1, g.shape[2]).T @ xs.reshape(-1, x.shape[2])
    return dw, g.reshape(-1, g.shape[2]).sum(0)


def transpose_weights(w):
    """Weights of the data gradient: wt[tap][ci][co] = w[8 - tap][co][ci]."""
    return np.ascontiguousarray(w[::-1].transpose(0, 2, 1))


def pool(x):
    R, C, K = x.shape
    xp = np.zeros((R + (R & 1), C + (C & 1), K))
    xp[:R, :C] = x
    return 0.25 * (xp[0::2, 0::2] + xp[0::2, 1::2] + xp[1::2, 0::2] + xp[1::2, 1::2])


def up(b, R, C):
    return np.repeat(np.repeat(b, 2, axis=0), 2, axis=1)[:R, :C]


def pool_t(g, R, C):
    """Adjoint of pool: each fine pixel receives a quarter of its parent."""
    return 0.25 * up(g, R, C)


def up_t(a, Rc, Cc):
    """Adjoint of up: each coarse pixel sums its children inside the image."""
    R, C, K = a.shape
    ap = np.zeros((2 * Rc, 2 * Cc, K))
    ap[:R, :C] = a
    return ap[0::2, 0::2] + ap[0::2, 1::2] + ap[1::2, 0::2] + ap[1::2, 1::2]


def split_params(flat):
    """Flat FP32 parameter vector -> [(w [9,cout,cin], b [cout])] per layer (layer by layer: w then b)."""
    out, k = [], 0
    for cin, cout, _, _ in LAYERS:
        n = 9 * cout * cin
        out.append((np.asarray(flat[k:k + n], np.float64).reshape(9, cout, cin),
                    np.asarray(flat[k + n:k + n + cout], np.float64)))
        k += n + cout
    return out


def n_params():
    return sum(9 * ci * co + co for ci, co, _, _ in LAYERS)


def forward(params, x0, bf=True):
    """params from split_params; x0 [R,C,16].  Returns (out [R,C,2], activations)."""
    W = [(_r(w, bf), b) for w, b in params]
    A = {"x0": _r(x0, bf)}

    def cv(k, x):
        w, b = W[k]
        y = conv(x, w, b)
        if LAYERS[k][3]:
            y = np.maximum(y, 0.0)
        return _r(y, bf) if k < 9 else y

    A["e0"] = cv(0, A["x0"])
    A["e0b"] = cv(1, A["e0"])
    A["p1"] = _r(pool(A["e0b"]), bf)
    A["e1"] = cv(2, A["p1"])
    A["p2"] = _r(pool(A["e1"]), bf)
    A["e2"] = cv(3, A["p2"])
    A["p3"] = _r(pool(A["e2"]), bf)
    A["e3"] = cv(4, A["p3"])
    A["e3b"] = cv(5, A["e3"])
    A["s2"] = _r(A["e2"] + up(A["e3b"], *A["e2"].shape[:2]), bf)
    A["d2"] = cv(6, A["s2"])
    A["s1"] = _r(A["e1"] + up(A["d2"], *A["e1"].shape[:2]), bf)
    A["d1"] = cv(7, A["s1"])
    A["s0"] = _r(A["e0b"] + up(A["d1"], *A["e0b"].shape[:2]), bf)
    A["d0"] = cv(8, A["s0"])
    full = cv(9, A["d0"])
    A["out"] = full[:, :, :2]
    return A["out"], A


def loss_grad(out, t, w, scale, bf=True):
    """loss = scale sum_p w |out - t|^2 and its gradient [R,C,16] (channels 2.. zero)."""
    d = out - t
    loss = scale * float(np.sum(w[..., None] * d * d))
    g = np.zeros(out.shape[:2] + (NOUT,))
    g[:, :, :2] = _r(2.0 * scale * w[..., None] * d, bf)
    return loss, g


def backward(params, A, g9, bf=True):
    """Weight and bias gradients [(dw, db)] per layer for one image, given
    d loss / d out (g9, [R,C,16])."""
    W = [_r(w, bf) for w, _ in params]
    grads = [None] * 10

    def dx(k, g):
        return _r(conv(g, transpose_weights(W[k]), None), bf)

    def mask(s, y):
        return _r(s, bf) * (y > 0)

    grads[9] = conv_grad_w(g9, A["d0"])
    dd0 = dx(9, g9)
    G8 = mask(dd0, A["d0"])
    grads[8] = conv_grad_w(G8, A["s0"])
    ds0 = dx(8, G8)
    G7 = mask(up_t(ds0, *A["d1"].shape[:2]), A["d1"])
    grads[7] = conv_grad_w(G7, A["s1"])
    ds1 = dx(7, G7)
    G6 = mask(up_t(ds1, *A["d2"].shape[:2]), A["d2"])
    grads[6] = conv_grad_w(G6, A["s2"])
    ds2 = dx(6, G6)
    G5 = mask(up_t(ds2, *A["e3b"].shape[:2]), A["e3b"])
    grads[5] = conv_grad_w(G5, A["e3"])
    de3 = dx(5, G5)
    G4 = mask(de3, A["e3"])
    grads[4] = conv_grad_w(G4, A["p3"])
    dp3 = dx(4, G4)
    G3 = mask(ds2 + pool_t(dp3, *A["e2"].shape[:2]), A["e2"])
    grads[3] = conv_grad_w(G3, A["p2"])
    dp2 = dx(3, G3)
    G2 = mask(ds1 + pool_t(dp2, *A["e1"].shape[:2]), A["e1"])
    grads[2] = conv_grad_w(G2, A["p1"])
    dp1 = dx(2, G2)
    G1 = mask(ds0 + pool_t(dp1, *A["e0b"].shape[:2]), A["e0b"])
    grads[1] = conv_grad_w(G1, A["e0"])
    de0 = dx(1, G1)
    G0 = mask(de0, A["e0"])
    grads[0] = conv_grad_w(G0, A["x0"])
    return grads


def loss_and_grads(flat, images, bf=True):
    """images: list of (x0 [R,C,16], t [R,C,2], w [R,C]); loss normalised by sum w.
    Returns (loss, flat gradient) summed over the images."""
    params = split_params(flat)
    scale = 1.0 / max(sum(float(w.sum()) for _, _, w in images), 1e-30)
    total, gsum = 0.0, [None] * 10
    for x0, t, w in images:
        out, A = forward(params, x0, bf)
        loss, g9 = loss_grad(out, t, w, scale, bf)
        total += loss
        for k, (dw, db) in enumerate(backward(params, A, g9, bf)):
            gsum[k] = (dw, db) if gsum[k] is None else (gsum[k][0] + dw, gsum[k][1] + db)
    flat_g = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in gsum])
    return total, flat_g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    """One Adam step in float32 (the GPU's arithmetic), returns (p, m, v)."""
    f = np.float32
    p, g, m, v = (np.asarray(a, np.float32) for a in (p, g, m, v))
    b1, b2 = f(b1), f(b2)   # the kernel's arguments are FP32, 1 - b is formed in FP32
    m = b1 * m + (f(1.0) - b1) * g
    v = b2 * v + (f(1.0) - b2) * g * g
    bc1, bc2 = f(1.0) - b1 ** f(step), f(1.0) - b2 ** f(step)
    p = p - f(lr) * (m / bc1) / (np.sqrt(v / bc2) + f(eps))
    return p, m, v


# ---------------------------------------------------------------------------
# network I/O of a solver problem (image g = member g's a-grid: rows i (x),
# columns j (y)); the arrays are the (ny, nx) a-grid / (ny, nx-1) cx /
# (ny-1, nx) cy fields of one member
# ---------------------------------------------------------------------------
RHO_I, GRAV = 910.0, 9.81


def features(Hi, Hs, Hb, sx, sy, beta_sq, A, m=3.0, n=3.0, vel_max=5000.0):
    """[nx, ny, 16] features (before BF16 rounding) of one member."""
    gs = np.hypot(sx, sy)
    taud = RHO_I * GRAV * Hi * gs
    speed = np.zeros_like(Hi)
    ok = (Hi > 0) & (gs > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        sl = np.where(ok & (beta_sq > 0), (taud / np.where(beta_sq > 0, beta_sq, 1.0)) ** m, 0.0)
        de = np.where(ok & (A > 0), 2.0 * A * taud ** n * Hi / (n + 2.0), 0.0)
        speed = np.minimum(sl + de, vel_max)
        ex = np.where(gs > 0, -speed * sx / np.where(gs > 0, gs, 1.0), 0.0)
        ey = np.where(gs > 0, -speed * sy / np.where(gs > 0, gs, 1.0), 0.0)
        lb = np.where(beta_sq > 0, np.log10(np.where(beta_sq > 0, beta_sq, 1.0)) - 4.0, 0.0)
        la = np.where(A > 0, np.log10(np.where(A > 0, A, 1.0)) + 16.5, 0.0)
    f = np.zeros(Hi.shape + (NFEAT,))
    f[..., 0], f[..., 1], f[..., 2] = Hi * 1e-3, Hs * 1e-3, Hb * 1e-3
    f[..., 3], f[..., 4] = 100.0 * sx, 100.0 * sy
    f[..., 5], f[..., 6], f[..., 7] = lb, la, (Hi > 0)
    f[..., 8], f[..., 9] = np.arcsinh(ex / 10.0), np.arcsinh(ey / 10.0)
    return np.ascontiguousarray(f.transpose(1, 0, 2))


def targets(u_cx, v_cy, Hi):
    """[nx, ny, 2] asinh-scaled a-grid velocities and [nx, ny] weights."""
    ny, nx = Hi.shape
    ua = np.empty((ny, nx))
    ua[:, 0], ua[:, -1] = u_cx[:, 0], u_cx[:, -1]
    ua[:, 1:-1] = 0.5 * (u_cx[:, :-1] + u_cx[:, 1:])
    va = np.empty((ny, nx))
    va[0, :], va[-1, :] = v_cy[0, :], v_cy[-1, :]
    va[1:-1, :] = 0.5 * (v_cy[:-1, :] + v_cy[1:, :])
    t = np.stack([np.arcsinh(ua / 10.0), np.arcsinh(va / 10.0)], axis=-1).transpose(1, 0, 2)
    return np.ascontiguousarray(t), np.ascontiguousarray((Hi > 0).astype(np.float64).T)


def warm_velocities(o, vel_max=5000.0):
    """Network output o [nx, ny, 2] -> (u_cx [ny, nx-1], v_cy [ny-1, nx])."""
    ua, va = 10.0 * np.sinh(o[..., 0]).T, 10.0 * np.sinh(o[..., 1]).T
    u = np.clip(0.5 * (ua[:, :-1] + ua[:, 1:]), -vel_max, vel_max)
    v = np.clip(0.5 * (va[:-1, :] + va[1:, :]), -vel_max, vel_max)
    return u, v
