"""Learned warm start of the DIVA solve (include/imau_net.h).

The network (four levels, 3x3 convolutions with 32 channels, BF16 on the
matrix cores) maps ten per-cell features of the geometry, friction and rate
factor to the a-grid velocities.  Its forward and backward passes, the loss and
the Adam optimiser are kernels of libimauvel.so; this module only sequences
them and owns the device buffers (torch allocates them, gathers training
batches and sums the loss weights; the network's arithmetic is the
library's).

IMAU-ICE has no counterpart: its Picard loop starts from the previous call's
velocities (src/ice_velocity_module.f90:249-374).  For independent solves the
network supplies that start (`warm_start`), and the solve then runs with
`warm_visc_iters` viscosity passes to bring viscosity and basal stress to it.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np
# torch's HIP runtime must initialise before libimauvel.so's in the same
# process: import it before the library is loaded by any Context
import torch

from . import solver

NET_EXPORTED = (
    "imau_net_plan_create", "imau_net_plan_for_problem", "imau_net_plan_destroy", "imau_net_plan_pixels",
    "imau_net_grad_w_work", "imau_net_loss_work", "imau_net_conv", "imau_net_conv_grad_w", "imau_net_pool",
    "imau_net_upadd", "imau_net_grad_pool", "imau_net_grad_up", "imau_net_loss", "imau_net_adam", "imau_net_pack",
    "imau_net_features", "imau_net_targets", "imau_net_warm_start", "imau_net_sync",
)

# (cin, cout, level, relu) of the ten convolutions, in order
LAYERS = [(16, 32, 0, True), (32, 32, 0, True), (32, 32, 1, True), (32, 32, 2, True), (32, 32, 3, True),
          (32, 32, 3, True), (32, 32, 2, True), (32, 32, 1, True), (32, 32, 0, True), (32, 16, 0, False)]
NFEAT, NOUT, WIDTH, NLEVELS = 16, 16, 32, 4
USED_FEATURES = 10
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_net.npz")

_bound = False


def _bind(L):
    global _bound
    if _bound:
        return L
    P, vp, i32, i64, f32 = ctypes.POINTER, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    L.imau_net_plan_create.argtypes = [vp, i32, P(i32), P(i32), P(vp)]
    L.imau_net_plan_for_problem.argtypes = [vp, vp, P(vp)]
    L.imau_net_plan_destroy.argtypes = [vp]
    L.imau_net_plan_destroy.restype = None
    L.imau_net_plan_pixels.argtypes = [vp, i32]
    L.imau_net_plan_pixels.restype = i64
    L.imau_net_grad_w_work.argtypes = [vp, i32, i32, i32]
    L.imau_net_grad_w_work.restype = i64
    L.imau_net_loss_work.argtypes = [vp]
    L.imau_net_loss_work.restype = i64
    L.imau_net_conv.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, i32, vp, i32]
    L.imau_net_conv_grad_w.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp]
    L.imau_net_pool.argtypes = [vp, vp, i32, vp, vp]
    L.imau_net_upadd.argtypes = [vp, vp, i32, vp, vp, vp]
    L.imau_net_grad_pool.argtypes = [vp, vp, i32, vp, vp, vp, vp]
    L.imau_net_grad_up.argtypes = [vp, vp, i32, vp, vp, vp]
    L.imau_net_loss.argtypes = [vp, vp, vp, vp, vp, f32, vp, vp, P(ctypes.c_double)]
    L.imau_net_adam.argtypes = [vp, i64, vp, vp, vp, vp, f32, f32, f32, f32, i32]
    L.imau_net_pack.argtypes = [vp, i32, i32, vp, vp, vp]
    L.imau_net_features.argtypes = [vp, vp, P(solver.ImauCfg), vp, vp]
    L.imau_net_targets.argtypes = [vp, vp, vp, vp, vp]
    L.imau_net_warm_start.argtypes = [vp, vp, P(solver.ImauCfg), vp, vp]
    L.imau_net_sync.argtypes = [vp]
    _bound = True
    return L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def n_params():
    return sum(9 * ci * co + co for ci, co, _, _ in LAYERS)


def init_params(seed=0):
    """He-normal weights, zero biases; unused input channels and output
    channels start (and stay) at zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, (cin, cout, _, _) in enumerate(LAYERS):
        fan = 9 * (USED_FEATURES if k == 0 else cin)
        w = rng.normal(scale=np.sqrt(2.0 / fan), size=(9, cout, cin))
        if k == 0:
            w[:, :, USED_FEATURES:] = 0.0
        if k == len(LAYERS) - 1:
            w *= 0.1
            w[:, 2:, :] = 0.0
        parts += [w.ravel(), np.zeros(cout)]
    return np.concatenate(parts).astype(np.float32)


class Plan:
    """Image layout of a batch at every level (imau_net_plan)."""

    def __init__(self, ctx, rows=None, cols=None, problem=None):
        self.ctx, self.L = ctx, _bind(ctx.L)
        self.h = ctypes.c_void_p()
        if problem is not None:
            ctx._check(self.L.imau_net_plan_for_problem(ctx.h, problem.h, ctypes.byref(self.h)))
        else:
            n = len(rows)
            r = (ctypes.c_int32 * n)(*rows)
            c = (ctypes.c_int32 * n)(*cols)
            ctx._check(self.L.imau_net_plan_create(ctx.h, n, r, c, ctypes.byref(self.h)))
        self.npix = [int(self.L.imau_net_plan_pixels(self.h, l)) for l in range(NLEVELS)]

    def close(self):
        if self.h:
            self.L.imau_net_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Net:
    """Parameters (FP32 master copy, BF16 packed copies, Adam moments) and the
    sequenced forward / backward passes.  `buffers(plan)` holds the
    activations of one plan, reused across calls."""

    def __init__(self, ctx, params=None, seed=0):
        import torch
        self.torch = torch
        self.ctx, self.L = ctx, _bind(ctx.L)
        dev = torch.device("cuda")
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.size != n_params():
            raise ValueError("parameter vector has %d entries, the network %d" % (p.size, n_params()))
        self.p = torch.from_numpy(p.copy()).to(dev)
        self.grad = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step = 0
        self.offs = []
        k = 0
        for cin, cout, _, _ in LAYERS:
            n = 9 * cout * cin
            self.offs.append((k, n, cout))
            k += n + cout
        self.wf = [torch.zeros(9 * ci * co, dtype=torch.int16, device=dev) for ci, co, _, _ in LAYERS]
        self.wt = [torch.zeros(9 * ci * co, dtype=torch.int16, device=dev) for ci, co, _, _ in LAYERS]
        self._bufs = {}
        self.pack()

    # -- parameters --------------------------------------------------------
    def w(self, k):
        o, n, _ = self.offs[k]
        return self.p[o:o + n]

    def b(self, k):
        o, n, c = self.offs[k]
        return self.p[o + n:o + n + c]

    def pack(self):
        self.torch.cuda.synchronize()
        for k, (cin, cout, _, _) in enumerate(LAYERS):
            self.ctx._check(self.L.imau_net_pack(self.ctx.h, cin, cout, _ptr(self.w(k)), _ptr(self.wf[k]),
                                                 _ptr(self.wt[k])))

    def params_numpy(self):
        self.sync()
        return self.p.cpu().numpy().copy()

    def save(self, path, **meta):
        np.savez(path, params=self.params_numpy(), layers=np.array(LAYERS, np.int32), **meta)

    @classmethod
    def load(cls, ctx, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            if not np.array_equal(z["layers"], np.array(LAYERS, np.int32)):
                raise ValueError("%s holds a different network" % path)
            return cls(ctx, params=z["params"])

    def sync(self):
        self.ctx._check(self.L.imau_net_sync(self.ctx.h))

    # -- buffers -----------------------------------------------------------
    def buffers(self, plan, train=False):
        key = (id(plan), train)
        if key in self._bufs:
            return self._bufs[key]
        torch = self.torch
        dev = torch.device("cuda")
        n = plan.npix

        def act(l, c=WIDTH):
            return torch.empty(n[l] * c, dtype=torch.int16, device=dev)

        B = dict(out=torch.empty(n[0] * 2, dtype=torch.float32, device=dev))
        for name, l in (("e0", 0), ("e0b", 0), ("p1", 1), ("e1", 1), ("p2", 2), ("e2", 2), ("p3", 3), ("e3", 3),
                        ("e3b", 3), ("s2", 2), ("d2", 2), ("s1", 1), ("d1", 1), ("s0", 0), ("d0", 0)):
            B[name] = act(l)
        if train:
            B["g9"] = act(0, NOUT)
            for name, l in (("ga", 0), ("gb", 0), ("gc", 0), ("h1a", 1), ("h1b", 1), ("h1c", 1), ("h2a", 2),
                            ("h2b", 2), ("h2c", 2), ("h3a", 3), ("h3b", 3)):
                B[name] = act(l)
            B["loss_work"] = torch.empty(int(self.L.imau_net_loss_work(plan.h)), dtype=torch.float64, device=dev)
            wmax = max(int(self.L.imau_net_grad_w_work(plan.h, l, ci, co)) for ci, co, l, _ in LAYERS)
            B["dw_work"] = torch.empty(wmax, dtype=torch.float32, device=dev)
        self._bufs[key] = (plan, B)  # the plan is kept alive with its buffers
        return self._bufs[key]

    def release(self, plan):
        for key in [k for k in self._bufs if k[0] == id(plan)]:
            del self._bufs[key]

    # -- passes ------------------------------------------------------------
    def _conv(self, plan, k, x, y, y_f32=False):
        cin, cout, l, relu = LAYERS[k]
        self.ctx._check(self.L.imau_net_conv(self.ctx.h, plan.h, l, cin, cout, _ptr(x), _ptr(self.wf[k]),
                                             _ptr(self.b(k)), 1 if relu else 0, _ptr(y), 1 if y_f32 else 0))

    def _convT(self, plan, k, g, y):
        cin, cout, l, _ = LAYERS[k]
        self.ctx._check(self.L.imau_net_conv(self.ctx.h, plan.h, l, cout, cin, _ptr(g), _ptr(self.wt[k]), None, 0,
                                             _ptr(y), 0))

    def _dw(self, plan, k, g, x, B):
        cin, cout, l, _ = LAYERS[k]
        o, n, c = self.offs[k]
        self.ctx._check(self.L.imau_net_conv_grad_w(self.ctx.h, plan.h, l, cin, cout, _ptr(g), _ptr(x),
                                                     _ptr(self.grad[o:o + n]), _ptr(self.grad[o + n:o + n + c]),
                                                     _ptr(B["dw_work"])))

    def forward(self, plan, x0, train=False):
        """x0: BF16 features [pix][16] at level 0 (int16 tensor).  Returns the
        FP32 output [pix][2] (a view of the plan's buffer)."""
        _, B = self.buffers(plan, train)
        h, L = self.ctx.h, self.L
        ck = self.ctx._check
        self._conv(plan, 0, x0, B["e0"])
        self._conv(plan, 1, B["e0"], B["e0b"])
        ck(L.imau_net_pool(h, plan.h, 0, _ptr(B["e0b"]), _ptr(B["p1"])))
        self._conv(plan, 2, B["p1"], B["e1"])
        ck(L.imau_net_pool(h, plan.h, 1, _ptr(B["e1"]), _ptr(B["p2"])))
        self._conv(plan, 3, B["p2"], B["e2"])
        ck(L.imau_net_pool(h, plan.h, 2, _ptr(B["e2"]), _ptr(B["p3"])))
        self._conv(plan, 4, B["p3"], B["e3"])
        self._conv(plan, 5, B["e3"], B["e3b"])
        ck(L.imau_net_upadd(h, plan.h, 2, _ptr(B["e2"]), _ptr(B["e3b"]), _ptr(B["s2"])))
        self._conv(plan, 6, B["s2"], B["d2"])
        ck(L.imau_net_upadd(h, plan.h, 1, _ptr(B["e1"]), _ptr(B["d2"]), _ptr(B["s1"])))
        self._conv(plan, 7, B["s1"], B["d1"])
        ck(L.imau_net_upadd(h, plan.h, 0, _ptr(B["e0b"]), _ptr(B["d1"]), _ptr(B["s0"])))
        self._conv(plan, 8, B["s0"], B["d0"])
        self._conv(plan, 9, B["d0"], B["out"], y_f32=True)
        return B["out"]

    def backward(self, plan, x0, target, weight, scale):
        """Loss of the last forward pass and the gradient of every parameter
        (into self.grad).  Returns the loss (waits for the stream)."""
        _, B = self.buffers(plan, True)
        h, L = self.ctx.h, self.L
        ck = self.ctx._check
        loss = ctypes.c_double()
        ck(L.imau_net_loss(h, plan.h, _ptr(B["out"]), _ptr(target), _ptr(weight), float(scale), _ptr(B["g9"]),
                           _ptr(B["loss_work"]), ctypes.byref(loss)))
        gp = lambda l, a, pc, y, g: ck(L.imau_net_grad_pool(h, plan.h, l, _ptr(a), _ptr(pc), _ptr(y), _ptr(g)))
        gu = lambda l, a, y, g: ck(L.imau_net_grad_up(h, plan.h, l, _ptr(a), _ptr(y), _ptr(g)))
        self._dw(plan, 9, B["g9"], B["d0"], B)
        self._convT(plan, 9, B["g9"], B["ga"])                      # dd0
        gp(0, B["ga"], None, B["d0"], B["gb"])                      # G8
        self._dw(plan, 8, B["gb"], B["s0"], B)
        self._convT(plan, 8, B["gb"], B["gc"])                      # ds0 (kept for the skip into e0b)
        gu(0, B["gc"], B["d1"], B["h1a"])                           # G7
        self._dw(plan, 7, B["h1a"], B["s1"], B)
        self._convT(plan, 7, B["h1a"], B["h1b"])                    # ds1 (kept for the skip into e1)
        gu(1, B["h1b"], B["d2"], B["h2a"])                          # G6
        self._dw(plan, 6, B["h2a"], B["s2"], B)
        self._convT(plan, 6, B["h2a"], B["h2b"])                    # ds2 (kept for the skip into e2)
        gu(2, B["h2b"], B["e3b"], B["h3a"])                         # G5
        self._dw(plan, 5, B["h3a"], B["e3"], B)
        self._convT(plan, 5, B["h3a"], B["h3b"])                    # de3
        gp(3, B["h3b"], None, B["e3"], B["h3a"])                    # G4
        self._dw(plan, 4, B["h3a"], B["p3"], B)
        self._convT(plan, 4, B["h3a"], B["h3b"])                    # dp3
        gp(2, B["h2b"], B["h3b"], B["e2"], B["h2c"])                # G3
        self._dw(plan, 3, B["h2c"], B["p2"], B)
        self._convT(plan, 3, B["h2c"], B["h2a"])                    # dp2
        gp(1, B["h1b"], B["h2a"], B["e1"], B["h1c"])                # G2
        self._dw(plan, 2, B["h1c"], B["p1"], B)
        self._convT(plan, 2, B["h1c"], B["h1a"])                    # dp1
        gp(0, B["gc"], B["h1a"], B["e0b"], B["ga"])                 # G1
        self._dw(plan, 1, B["ga"], B["e0"], B)
        self._convT(plan, 1, B["ga"], B["gb"])                      # de0
        gp(0, B["gb"], None, B["e0"], B["gc"])                      # G0
        self._dw(plan, 0, B["gc"], x0, B)
        return loss.value

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        self.ctx._check(self.L.imau_net_adam(self.ctx.h, self.p.numel(), _ptr(self.p), _ptr(self.grad), _ptr(self.m),
                                             _ptr(self.v), lr, b1, b2, eps, self.step))
        for k, (cin, cout, _, _) in enumerate(LAYERS):
            self.ctx._check(self.L.imau_net_pack(self.ctx.h, cin, cout, _ptr(self.w(k)), _ptr(self.wf[k]),
                                                 _ptr(self.wt[k])))


# ---------------------------------------------------------------------------
# problems <-> network
# ---------------------------------------------------------------------------
class ProblemIO:
    """Network plan and feature buffer of one device-resident Problem."""

    def __init__(self, net: Net, problem):
        torch = net.torch
        self.net, self.problem = net, problem
        self.plan = Plan(net.ctx, problem=problem)
        self.x0 = torch.empty(self.plan.npix[0] * NFEAT, dtype=torch.int16, device="cuda")

    def features(self, cfg):
        ctx = self.net.ctx
        ctx._check(self.net.L.imau_net_features(ctx.h, self.problem.h, ctypes.byref(cfg), self.plan.h,
                                                _ptr(self.x0)))
        return self.x0

    def targets(self, t, w):
        ctx = self.net.ctx
        ctx._check(self.net.L.imau_net_targets(ctx.h, self.problem.h, self.plan.h, _ptr(t), _ptr(w)))

    def warm_start(self, cfg):
        """Features, forward pass and the warm start of every member's velocities."""
        out = self.net.forward(self.plan, self.features(cfg))
        ctx = self.net.ctx
        ctx._check(self.net.L.imau_net_warm_start(ctx.h, self.problem.h, ctypes.byref(cfg), self.plan.h, _ptr(out)))


def forward_flops(plan):
    """Multiply-add work of one forward pass (2 flops each) over the plan."""
    return sum(2.0 * 9 * ci * co * plan.npix[l] for ci, co, l, _ in LAYERS)


class WarmStartSolver:
    """Learned warm start of a device-resident Problem, then its solve.

    The plain solve is the fallback at two levels:
    * inside the Picard loop, per member (imau_cfg.stall_window): a member
      whose FP64 residual stalls -- it does not fall by stall_drop over
      stall_window iterations, or the warm start is no better than zero
      velocities -- restarts from the cold state with the cold solve's
      relaxation, and its iterations are counted in full;
    * after the loop: if any member still ends above cfg.nonlinear_rtol, the
      whole problem is solved again from the cold state.
    If that cold solve too ends above the target, solve() raises ImauError
    with IMAU_ERR_STALLED (the result is left in the problem).

    Call problem.reset_state() before solve() when the problem holds a solved
    state (the warm start replaces the velocities, the completion passes the
    viscosity and basal stress)."""

    def __init__(self, ctx, problem, net=None, warm_iters=0, column_iters=9, relax=0.7, stall_window=4,
                 stall_drop=10.0):
        self.net = net if net is not None else Net.load(ctx)
        self.problem = problem
        self.io = ProblemIO(self.net, problem)
        self.warm_iters = warm_iters
        self.column_iters = column_iters
        # relaxation of the unmixed Picard steps (the first, and any after an
        # Anderson restart) from the warm start; the cold fallback keeps cfg's
        self.relax = relax
        self.stall_window, self.stall_drop = stall_window, stall_drop
        self.fallbacks = 0        # whole-problem cold solves after the loop
        self.restarts = 0         # per-member cold restarts inside the loop

    def warm_start(self, cfg):
        self.io.warm_start(cfg)

    def solve(self, cfg):
        fields = ("warm_visc_iters", "warm_column_iters", "DIVA_visc_it_relax", "stall_window", "stall_drop",
                  "stall_relax")
        saved = {k: getattr(cfg, k) for k in fields}
        cfg.warm_visc_iters, cfg.warm_column_iters = self.warm_iters, self.column_iters
        if cfg.stop_rule == 1 and self.stall_window > 0:
            cfg.stall_window, cfg.stall_drop = self.stall_window, self.stall_drop
            cfg.stall_relax = saved["DIVA_visc_it_relax"]   # a restarted member solves as the cold start does
        if self.relax is not None:
            cfg.DIVA_visc_it_relax = self.relax
        try:
            self.io.warm_start(cfg)
            self.problem.solve(cfg)
        finally:
            for k, v in saved.items():
                setattr(cfg, k, v)
        its, res = self.problem.stats()
        self.restarts += sum(self.problem.restarts())
        if cfg.stop_rule == 1 and any(not r <= cfg.nonlinear_rtol for r in res):
            self.fallbacks += 1
            self.problem.reset_state()
            self.problem.solve(cfg)
            its, res = self.problem.stats()
            bad = [g for g, r in enumerate(res) if not r <= cfg.nonlinear_rtol]
            if bad:
                raise solver.ImauError(
                    -5, "FP64 residual above %g after the cold fallback for member(s) %s (residuals %s)" %
                    (cfg.nonlinear_rtol, bad, ["%.3e" % res[g] for g in bad]))
        return its, res

    def close(self):
        self.net.release(self.io.plan)
        self.io.plan.close()


def training_data(ctx, cfg, n_samples, n=256, batch=64, seed0=10_000, log=print, workers=8, prepare=None):
    """Solve n_samples seeded n x n random ice caps (A and the friction scale
    log-uniform over one decade each, as the ensemble) and return device
    tensors (features [S, n*n*16] int16, targets [S, n*n*2], weights [S, n*n]).
    `prepare(states)` sets the per-call solver fields before each batch.  A
    sample whose solve did not reach cfg.nonlinear_rtol gets zero weight."""
    import multiprocessing
    import torch
    from concurrent.futures import ProcessPoolExecutor
    X = torch.empty((n_samples, n * n * NFEAT), dtype=torch.int16, device="cuda")
    T = torch.empty((n_samples, n * n * 2), dtype=torch.float32, device="cuda")
    W = torch.empty((n_samples, n * n), dtype=torch.float32, device="cuda")
    L = _bind(ctx.L)
    t0 = time.time()
    # geometry workers are fresh interpreters (spawned, not forked from a process using the GPU)
    with ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("spawn")) as pool:
        for s0 in range(0, n_samples, batch):
            seeds = list(range(s0, min(n_samples, s0 + batch)))
            members = [_unpack(x) for x in pool.map(_sample_geometry, [(seed0 + s, n) for s in seeds])]
            grids = [g for g, _ in members]
            states = [st for _, st in members]
            if prepare is not None:
                prepare(states)
            prob = solver.Problem(ctx, grids, cfg, states)
            prob.solve(cfg)
            plan = Plan(ctx, problem=prob)
            k = len(seeds)
            xb = torch.empty(plan.npix[0] * NFEAT, dtype=torch.int16, device="cuda")
            tb = torch.empty(plan.npix[0] * 2, dtype=torch.float32, device="cuda")
            wb = torch.empty(plan.npix[0], dtype=torch.float32, device="cuda")
            ctx._check(L.imau_net_features(ctx.h, prob.h, ctypes.byref(cfg), plan.h, _ptr(xb)))
            ctx._check(L.imau_net_targets(ctx.h, prob.h, plan.h, _ptr(tb), _ptr(wb)))
            ctx._check(L.imau_net_sync(ctx.h))
            out = prob.download(outputs=set())
            bad = [q for q, o in enumerate(out) if not o.stats["nonlinear_residual"] <= cfg.nonlinear_rtol]
            if bad:
                wv = wb.view(k, -1)
                for q in bad:
                    wv[q].zero_()
                log("training data: %d samples of this batch not converged, weight 0" % len(bad))
            X[s0:s0 + k] = xb.view(k, -1)
            T[s0:s0 + k] = tb.view(k, -1)
            W[s0:s0 + k] = wb.view(k, -1)
            torch.cuda.synchronize()
            plan.close()
            prob.close()
            log("training data: %d / %d solved (%.1f s)" % (s0 + k, n_samples, time.time() - t0))
    return X, T, W


def _sample_geometry(arg):
    """One seeded training geometry.  Constant fields (outputs, the cold start,
    a uniform rate factor) travel as (shape, dtype, order, value) and are
    re-created by _unpack."""
    seed, n = arg
    from . import geometry
    rng = np.random.default_rng(seed)
    A = 10.0 ** rng.uniform(-17.0, -16.0)
    bscale = 10.0 ** rng.uniform(0.0, 1.0)
    g, _, st = geometry.random_geometry(n, n, dx=4e3, seed=seed, A=A, beta_scale=bscale)
    const = {}
    for k, v in list(st.f.items()):
        if v.size > 1024 and (v.flat[0] == v).all():
            const[k] = (v.shape, v.dtype.str, "F" if v.flags.f_contiguous and not v.flags.c_contiguous else "C",
                        v.flat[0].item())
            del st.f[k]
    return g, st, const


def _unpack(sample):
    g, st, const = sample
    for k, (shape, dt, order, value) in const.items():
        st.f[k] = np.full(shape, value, dtype=np.dtype(dt), order=order)
    return g, st


def train(net: Net, X, T, W, n=256, epochs=20, batch=16, lr=2e-3, seed=0, log=print, holdout=None):
    """Adam on the weighted squared error of the asinh-scaled velocities.
    X, T, W as from training_data; the learning rate decays as a cosine."""
    import torch
    S = X.shape[0]
    plan = Plan(net.ctx, rows=[n] * batch, cols=[n] * batch)
    xb = torch.empty(batch * n * n * NFEAT, dtype=torch.int16, device="cuda")
    tb = torch.empty(batch * n * n * 2, dtype=torch.float32, device="cuda")
    wb = torch.empty(batch * n * n, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(seed)
    steps = epochs * (S // batch)
    hist, k, t0 = [], 0, time.time()
    for ep in range(epochs):
        order = rng.permutation(S)
        tot, cnt = 0.0, 0
        for b0 in range(0, S - batch + 1, batch):
            idx = torch.from_numpy(order[b0:b0 + batch].copy()).to("cuda")
            net.sync()                      # the previous step's kernels read xb / tb / wb
            xb.view(batch, -1).copy_(X.index_select(0, idx))
            tb.view(batch, -1).copy_(T.index_select(0, idx))
            wb.view(batch, -1).copy_(W.index_select(0, idx))
            wsum = float(wb.sum())
            torch.cuda.synchronize()
            net.forward(plan, xb, train=True)
            loss = net.backward(plan, xb, tb, wb, 1.0 / max(wsum, 1.0))
            rate = lr * 0.5 * (1.0 + np.cos(np.pi * k / max(steps, 1)))
            net.adam_step(rate)
            tot += loss
            cnt += 1
            k += 1
        hist.append(tot / max(cnt, 1))
        log("epoch %d: loss %.5f (%.1f s)" % (ep + 1, hist[-1], time.time() - t0))
    net.sync()
    net.release(plan)
    plan.close()
    return hist
