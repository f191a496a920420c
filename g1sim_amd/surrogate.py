"""Learned forward-dynamics surrogate: an MLP qacc = f(q, qd, tau_total).

Network: 106 inputs (qpos 36, qvel 35, total generalized force 35), four
hidden layers of 512 with SiLU, 35 outputs; inputs and outputs z-scored;
bf16 weights and activations on MFMA with fp32 accumulation; fp32 master
weights and Adam moments.  Every GEMM, activation, loss-gradient, bias-gradient
and optimizer step runs in the project's own HIP kernels
(csrc/g1_mlp.hip); torch only allocates device memory and provides the
stream.  Padded shapes: inputs to 128 columns, outputs to 128 rows, batch to
a multiple of 128.

Training data come from the ABA step kernel itself (g1_hip_step_ex with the
qacc and tau outputs) on randomized-actuation rollouts.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .model import NQ, NV
from .sim import check_soa, norm_device

NIN, NOUT, HID, NHID = NQ + NV + NV, NV, 512, 4
PIN, POUT = 128, 128
SNAP_ROWS = NIN + NV          # deferred-judgement record: network inputs, ABA qacc
EPI = dict(F32=0, F32_ATOMIC=1, BIAS_SILU=2, BIAS_SILU_H=3, BIAS_F32=4, SILU_BWD=5, F32_SPLIT=6)


def k_tiled(w16, nout):
    """bf16 [out][in] -> the fused kernels' layout [in/32][nout][32] (first
    nout rows)."""
    rows, cols = w16.shape
    assert nout <= rows and cols % 32 == 0
    return w16[:nout].reshape(nout, cols // 32, 32).permute(1, 0, 2).contiguous()


def _pad128(n):
    return (n + 127) // 128 * 128


class MLP:
    """Parameters, optimizer state and the kernel sequence of the surrogate."""

    def __init__(self, device, seed=0, hidden=HID, nhid=NHID):
        self.lib = _lib.load()
        self.dev = torch.device(device)
        self.dims = [(hidden, PIN, NIN)] + [(hidden, hidden, hidden)] * (nhid - 1) + [(POUT, hidden, hidden)]
        # (rows padded, cols padded, true fan-in); true fan-out of the last layer is NOUT
        rng = np.random.default_rng(seed)
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.W, self.b, self.W16, self.mW, self.vW, self.mb, self.vb, self.maskW, self.maskb = ([] for _ in range(9))
        self.W16T = []
        for li, (rows, cols, fan_in) in enumerate(self.dims):
            out_true = NOUT if li == len(self.dims) - 1 else rows
            bound = math.sqrt(6.0 / fan_in) if li < len(self.dims) - 1 else math.sqrt(1.0 / fan_in)
            w = np.zeros((rows, cols), np.float32)
            w[:out_true, :fan_in] = rng.uniform(-bound, bound, (out_true, fan_in))
            mask = np.zeros((rows, cols), np.uint8)
            mask[:out_true, :fan_in] = 1
            mb = np.zeros(rows, np.uint8)
            mb[:out_true] = 1
            self.W.append(torch.as_tensor(w, device=self.dev))
            self.b.append(torch.zeros(rows, **f32))
            self.W16.append(self.W[-1].to(torch.bfloat16))
            self.W16T.append(self.W16[-1].t().contiguous())      # [in][out]: the data-gradient GEMM's operand
            self.mW.append(torch.zeros_like(self.W[-1])); self.vW.append(torch.zeros_like(self.W[-1]))
            self.mb.append(torch.zeros_like(self.b[-1])); self.vb.append(torch.zeros_like(self.b[-1]))
            self.maskW.append(torch.as_tensor(mask, device=self.dev))
            self.maskb.append(torch.as_tensor(mb, device=self.dev))
        self.t = 0                # optimizer steps taken (host mirror of train_state[1])
        self._ws = {}
        # device-resident loop scalars: state int [3] = {next iteration of the
        # run, Adam step, this iteration}; sched float [3] = {lr, 1 - b1^t, 1 - b2^t}
        self.train_state = torch.zeros(4, dtype=torch.int32, device=self.dev)
        self.sched = torch.zeros(4, **f32)
        self._adam_args = None
        # z-score statistics (set by fit_normalization)
        self.xmean = torch.zeros(NIN, **f32); self.xinv = torch.ones(NIN, **f32)
        self.ymean = torch.zeros(NOUT, **f32); self.yinv = torch.ones(NOUT, **f32)

    @property
    def stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {_lib.last_error()}")

    def gemm(self, a_t, b_t, epi, A, lda, B, ldb, I, J, R, C=None, H=None, ldc=0, bias=None, Z=None, split=1,
             colpart=None):
        rc = self.lib.g1_mlp_gemm_ex(a_t, b_t, EPI[epi], A.data_ptr(), lda, B.data_ptr(), ldb, I, J, R, split,
                                     None if C is None else C.data_ptr(), None if H is None else H.data_ptr(), ldc,
                                     None if bias is None else bias.data_ptr(), None if Z is None else Z.data_ptr(),
                                     None if colpart is None else colpart.data_ptr(), self.stream)
        self._check(rc, f"gemm {epi}")

    def workspace(self, B):
        """Per-batch-size activation buffers (sample-major)."""
        if B in self._ws:
            return self._ws[B]
        f32 = dict(dtype=torch.float32, device=self.dev)
        L = len(self.dims)
        split = self._split(B)
        ws = {
            "X16": torch.zeros((B, PIN), dtype=torch.bfloat16, device=self.dev),
            "T": torch.zeros((B, POUT), **f32),
            "Z": [torch.zeros((B, self.dims[l][0]), **f32) for l in range(L - 1)],
            "H16": [torch.zeros((B, self.dims[l][0]), dtype=torch.bfloat16, device=self.dev) for l in range(L - 1)],
            "pred": torch.zeros((B, POUT), **f32),
            "dY16": torch.zeros((B, POUT), dtype=torch.bfloat16, device=self.dev),
            "dZ16": [torch.zeros((B, self.dims[l][0]), dtype=torch.bfloat16, device=self.dev) for l in range(L - 1)],
            # the gradients as the iteration leaves them for Adam: the r-slices of
            # each weight-gradient GEMM (split over the batch) and the 64-row
            # column partials of each bias gradient
            "gW_split": [torch.zeros((split, r * c), **f32) for r, c, _ in self.dims],
            "gb_part": [torch.zeros((B // 64, r), **f32) for r, _, _ in self.dims],
            # summed squared error per block of 64 rows x 32 columns
            "sse": torch.zeros((B // 64, POUT // 32), **f32),
            # the summed gradients (gradients(): for inspection; Adam adds the slices itself)
            "gW": [torch.zeros_like(w) for w in self.W],
            "gb": [torch.zeros_like(b) for b in self.b],
        }
        self._ws[B] = ws
        return ws

    # ------------------------------------------------------------ forward
    def forward(self, X16, B, train=True, out=None):
        """X16: bf16 [B][128] z-scored inputs.  Returns the fp32 [B][128]
        z-scored prediction (first 35 columns meaningful)."""
        assert B % 128 == 0
        ws = self.workspace(B)
        L = len(self.dims)
        h = X16
        for l in range(L - 1):
            rows, cols, _ = self.dims[l]
            if train:
                self.gemm(0, 0, "BIAS_SILU", h, cols, self.W16[l], cols, B, rows, cols,
                          C=ws["Z"][l], H=ws["H16"][l], ldc=rows, bias=self.b[l])
            else:
                self.gemm(0, 0, "BIAS_SILU_H", h, cols, self.W16[l], cols, B, rows, cols,
                          H=ws["H16"][l], ldc=rows, bias=self.b[l])
            h = ws["H16"][l]
        rows, cols, _ = self.dims[-1]
        pred = ws["pred"] if out is None else out
        self.gemm(0, 0, "BIAS_F32", h, cols, self.W16[-1], cols, B, rows, cols, C=pred, ldc=rows, bias=self.b[-1])
        return pred

    # ------------------------------------------------------------ backward
    @staticmethod
    def _split(B):
        split = max(1, min(8, B // 1024))        # 8 at the 8,192-sample batch: measured best of 2-16
        while B % (64 * split):
            split //= 2
        return split

    def backward(self, X16, B, dY16):
        """Gradients of the loss given dL/dpred (bf16; the output layer's bias
        partials come with it from loss_grad): the weight gradients as the
        r-slices of split GEMMs in ws['gW_split'], the bias gradients as
        64-row column partials in ws['gb_part'] (written by the data-gradient
        GEMMs' epilogue).  Adam adds the slices; gradients() sums them for
        inspection."""
        ws = self.workspace(B)
        L = len(self.dims)
        split = self._split(B)
        d16 = dY16
        for l in range(L - 1, -1, -1):
            rows, cols, _ = self.dims[l]
            hin = X16 if l == 0 else ws["H16"][l - 1]
            if l > 0:
                prow = self.dims[l - 1][0]
                # one launch: dW = d^T h (reduction over the batch in r-slices)
                # and dH_{l-1} = d W_l through SiLU'(Z_{l-1}) as bf16, with the
                # column partials of the layer l-1 bias gradient
                rc = self.lib.g1_mlp_backward_layer(d16.data_ptr(), rows, hin.data_ptr(), cols,
                                                    self.W16T[l].data_ptr(), B, split, ws["gW_split"][l].data_ptr(),
                                                    ws["dZ16"][l - 1].data_ptr(), prow, ws["Z"][l - 1].data_ptr(),
                                                    ws["gb_part"][l - 1].data_ptr(), self.stream)
                self._check(rc, "backward_layer")
                d16 = ws["dZ16"][l - 1]
            else:
                # dW = d^T h: reduction over the batch, split across gridDim.z
                # into r-slices (plain stores)
                self.gemm(1, 1, "F32_SPLIT", d16, rows, hin, cols, rows, cols, B, C=ws["gW_split"][l], ldc=cols,
                          split=split)

    def gradients(self, B):
        """Sums the slices backward() left into ws['gW'] and ws['gb'] (in the
        order Adam adds them for the weights; z ascending for the biases)."""
        ws = self.workspace(B)
        for l, (rows, cols, _) in enumerate(self.dims):
            for parts, out, n in ((ws["gW_split"][l], ws["gW"][l], rows * cols), (ws["gb_part"][l], ws["gb"][l], rows)):
                self._check(self.lib.g1_mlp_sum_splits(parts.data_ptr(), parts.shape[0], n, out.data_ptr(),
                                                       self.stream), "sum_splits")
        return ws["gW"], ws["gb"]

    def adam(self, B, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step on every weight and bias (one launch) from the
        gradient slices backward() left, with the learning rate and bias
        corrections of self.sched (g1_mlp_train_schedule)."""
        ws = self.workspace(B)
        if self._adam_args is None or self._adam_args[0] is not ws:
            L = len(self.dims)
            ps, ms, vs = self.W + self.b, self.mW + self.mb, self.vW + self.vb
            gs, p16, p16t = ws["gW_split"] + ws["gb_part"], self.W16 + [None] * L, self.W16T + [None] * L
            masks = self.maskW + self.maskb
            n = (C.c_int * (2 * L))(*[t.numel() for t in ps])
            nsum = (C.c_int * (2 * L))(*[g.shape[0] for g in gs])
            gstride = (C.c_longlong * (2 * L))(*[g.shape[1] for g in gs])
            cols = (C.c_int * (2 * L))(*([w.shape[1] for w in self.W] + [1] * L))
            wd = (C.c_float * (2 * L))(*([0.0] * (2 * L)))
            self._adam_args = (ws, [_lib.ptr_array(x) for x in (ps, ms, vs, gs, p16)], _lib.ptr_array(p16t), cols,
                               _lib.ptr_array(masks), n, nsum, gstride, wd)
        _, arrs, p16t, cols, masks, n, nsum, gstride, wd = self._adam_args
        self._check(self.lib.g1_mlp_adam_multi_sum(len(n), *arrs, p16t, cols, masks, n, nsum, gstride, wd,
                                                   self.sched.data_ptr(), b1, b2, eps, self.stream), "adam")

    def loss_grad(self, pred, target, B):
        """MSE on the 35 meaningful outputs: dY as bf16 (the GEMM operand),
        the output layer's bias-gradient partials (ws['gb_part'][-1]) and the
        summed squared error per block of 64 rows x 32 columns, which it
        returns (a device tensor; its sum is the batch's SSE)."""
        ws = self.workspace(B)
        self._check(self.lib.g1_mlp_mse_grad_part(pred.data_ptr(), target.data_ptr(), B, POUT, NOUT,
                                                  ws["dY16"].data_ptr(), ws["gb_part"][-1].data_ptr(),
                                                  ws["sse"].data_ptr(), self.stream), "mse")
        return ws["sse"]

    def iteration(self, ds, idx, B, lo, hi, seed, iters, lr, lr_end):
        """One training iteration as a fixed launch sequence (every scalar that
        changes between iterations is read on the device): schedule, batch
        sampling, gather, forward, MSE gradient (its SSE per block into
        workspace['sse']), backward, Adam."""
        ws = self.workspace(B)
        s = self.stream
        self._check(self.lib.g1_mlp_train_schedule(self.train_state.data_ptr(), self.sched.data_ptr(), iters, lr,
                                                   lr_end, 0.9, 0.999, s), "schedule")
        self._check(self.lib.g1_mlp_sample(B, lo, hi, ds.nenv, seed & 0xFFFFFFFF, self.train_state.data_ptr(),
                                           idx.data_ptr(), s), "sample")
        rc = self.lib.g1_mlp_gather_rows(ds.rows().data_ptr(), idx.data_ptr(), B, NIN, NOUT, ds.nenv,
                                         self.xmean.data_ptr(), self.xinv.data_ptr(), self.ymean.data_ptr(),
                                         self.yinv.data_ptr(), ws["X16"].data_ptr(), PIN, ws["T"].data_ptr(), POUT, s)
        self._check(rc, "gather_rows")
        pred = self.forward(ws["X16"], B, train=True)
        self.loss_grad(pred, ws["T"], B)
        self.backward(ws["X16"], B, ws["dY16"])
        self.adam(B)

    # ------------------------------------------------------ fused inference
    @property
    def fusable(self):
        return [d[:2] for d in self.dims] == [(HID, PIN), (HID, HID), (HID, HID), (HID, HID), (POUT, HID)]

    def _wb(self):
        """Weights in the fused kernel's k-tiled layout [in/32][out][32]
        (rebuilt after every optimizer step) and the bias pointers."""
        if getattr(self, "_wb_t", None) != self.t:
            outs = [HID, HID, HID, HID, 64]
            self._Wt = [k_tiled(w, o) for w, o in zip(self.W16, outs)]
            self._wb_cache = (_lib.ptr_array(self._Wt), _lib.ptr_array(self.b))
            self._wb_t = self.t
        return self._wb_cache

    def forward_state(self, nenv, qpos, qvel, tau, pred):
        """Fused forward of every env of a SoA state (features built in the
        kernel): pred fp32 [>= nenv][ldp >= 64], z-scored (35 meaningful)."""
        assert self.fusable, "the fused kernel is built for 106 -> 512 x4 -> 35"
        W, b = self._wb()
        rc = self.lib.g1_hip_surrogate_forward(nenv, qpos.data_ptr(), qvel.data_ptr(), tau.data_ptr(),
                                               self.xmean.data_ptr(), self.xinv.data_ptr(), W, b, pred.data_ptr(),
                                               pred.shape[1], self.stream)
        self._check(rc, "surrogate_forward")
        return pred

    def set_normalization(self, xmean, xstd, ymean, ystd):
        self.xmean.copy_(torch.as_tensor(xmean, dtype=torch.float32))
        self.xinv.copy_(1.0 / torch.as_tensor(xstd, dtype=torch.float32).clamp_min(1e-6))
        self.ymean.copy_(torch.as_tensor(ymean, dtype=torch.float32))
        self.yinv.copy_(1.0 / torch.as_tensor(ystd, dtype=torch.float32).clamp_min(1e-6))

    def gather(self, X, Y, idx, B, nenv, X16, Yn=None):
        """z-scored bf16 inputs (and fp32 targets) of samples idx (int32 [B][2]:
        step, env) from SoA blocks X [steps][106][nenv], Y [steps][35][nenv]."""
        rc = self.lib.g1_mlp_gather(X.data_ptr(), None if Y is None else Y.data_ptr(),
                                    None if idx is None else idx.data_ptr(), B, NIN, NOUT, nenv,
                                    self.xmean.data_ptr(), self.xinv.data_ptr(), self.ymean.data_ptr(),
                                    self.yinv.data_ptr(), X16.data_ptr(), PIN,
                                    None if Yn is None else Yn.data_ptr(), POUT, self.stream)
        self._check(rc, "gather")

    def state_dict(self):
        return {"W": [w.cpu() for w in self.W], "b": [b.cpu() for b in self.b],
                "xmean": self.xmean.cpu(), "xinv": self.xinv.cpu(), "ymean": self.ymean.cpu(),
                "yinv": self.yinv.cpu()}


# ====================================================================== data
class Dataset:
    """Transitions of the ABA step kernel: X [steps][106][nenv] = (qpos, qvel,
    tau_total) before each step, Y [steps][35][nenv] = the step's qacc."""

    def __init__(self, X, Y):
        self.X, self.Y = X, Y
        self.steps, _, self.nenv = X.shape
        self._rows = None

    @property
    def n(self):
        return self.steps * self.nenv

    def rows(self, chunk=64):
        """Sample-major copy S [steps][nenv][106 + 35] (features, then targets),
        built on first use a few steps at a time: the training minibatch reads
        each sample as one contiguous row (g1_mlp_gather_rows) instead of 141
        scattered floats."""
        if self._rows is None:
            S = torch.empty((self.steps, self.nenv, NIN + NOUT), dtype=torch.float32, device=self.X.device)
            for s0 in range(0, self.steps, chunk):
                s1 = min(self.steps, s0 + chunk)
                S[s0:s1, :, :NIN].copy_(self.X[s0:s1].transpose(1, 2))
                S[s0:s1, :, NIN:].copy_(self.Y[s0:s1].transpose(1, 2))
            self._rows = S
        return self._rows


def generate_dataset(device, nenv=16384, steps=1221, dt=2e-3, seed=7, reset_below=0.35):
    """Randomized-actuation rollouts (targets stand + U(+-0.3) resampled every
    10 steps) stepped by g1_hip_step_ex one step at a time with its qacc and
    tau outputs; envs whose pelvis drops below `reset_below` m are reset by
    g1_cuda_reset_done with fresh noise."""
    from . import workloads as W
    lib = _lib.load()
    dev = torch.device(device)
    nblocks = (steps + 9) // 10
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=seed, nblocks=nblocks)
    qpos = torch.as_tensor(qp, device=dev)
    qvel = torch.as_tensor(qv, device=dev)
    anchor = torch.as_tensor(an, device=dev)
    ctrl = torch.as_tensor(cb, device=dev)
    X = torch.empty((steps, NIN, nenv), dtype=torch.float32, device=dev)
    Y = torch.empty((steps, NOUT, nenv), dtype=torch.float32, device=dev)
    done = torch.zeros(nenv, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    for s in range(steps):
        X[s, :NQ].copy_(qpos)
        X[s, NQ:NQ + NV].copy_(qvel)
        rc = lib.g1_hip_step_ex(nenv, 1, dt, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
                                ctrl[s // 10].data_ptr(), None, None, None, Y[s].data_ptr(),
                                X[s, NQ + NV:].data_ptr(), None, None, 0, stream)
        _lib.check(rc)
        if s % 50 == 49:
            done.copy_(qpos[2] < reset_below)
            lib.g1_cuda_reset_done(nenv, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(), done.data_ptr(),
                                   (seed * 1000003 + s) & 0xFFFFFFFF, 0.3, 0.02, stream)
    return Dataset(X, Y)


def fit_normalization(mlp, ds, sample=2_000_000, seed=0):
    g = torch.Generator(device=ds.X.device).manual_seed(seed)
    k = min(sample, ds.n)
    st = torch.randint(0, ds.steps, (k,), device=ds.X.device, generator=g)
    en = torch.randint(0, ds.nenv, (k,), device=ds.X.device, generator=g)
    xs = ds.X[st, :, en].double()
    ys = ds.Y[st, :, en].double()
    mlp.set_normalization(xs.mean(0), xs.std(0), ys.mean(0), ys.std(0))


def train(mlp, ds, iters=4000, batch=8192, lr=1e-3, lr_end=1e-5, holdout_steps=None, seed=0, log=None,
          graph=True, timing=None):
    """Minibatch Adam on the steps [0, steps - holdout), cosine learning-rate
    decay from lr to lr_end; returns (iteration, mean z-scored MSE) every 100
    iterations.  With graph=True the first iteration runs eagerly and the
    rest replay it as one HIP graph (the iteration reads all its changing
    scalars on the device, so the replays are the same computation, bit for
    bit, as eager iterations).  timing (a dict, optional) receives the
    average duration of a replayed iteration, from HIP events on the stream
    around iterations 1-99 (no host synchronisation among them)."""
    dev = ds.X.device
    holdout = holdout_steps if holdout_steps is not None else max(1, ds.steps // 20)
    train_steps = ds.steps - holdout
    ws = mlp.workspace(batch)
    idx = torch.empty(2 * batch, dtype=torch.int32, device=dev)
    mlp.train_state[0].zero_()
    run = lambda: mlp.iteration(ds, idx, batch, 0, train_steps, seed, iters, lr, lr_end)  # noqa: E731
    hist = []

    def logged(it):
        if it % 100 == 0 or it == iters - 1:
            v = float(ws["sse"].sum().item()) / (batch * NOUT)
            hist.append((it, v))
            if log:
                log(f"iter {it} mse {v:.4e}")

    g = None
    ev = None
    last_timed = min(iters - 1, 99)
    for it in range(iters):
        if g is None:
            run()
            if graph and it == 0 and iters > 1:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    run()
        else:
            if timing is not None and it == 1:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record()
            g.replay()
            if ev is not None and it == last_timed:
                ev[1].record()
        mlp.t += 1
        logged(it)
    if ev is not None:
        torch.cuda.synchronize()
        timing["iterations"] = last_timed
        timing["avg_iteration_ms"] = ev[0].elapsed_time(ev[1]) / last_timed
        timing["batch"] = batch
    return hist


def heldout_indices(ds, steps_from=None, nsamp=65536, seed=1):
    """(step, env) int32 [B][2] of B = nsamp rounded up to 128 samples drawn
    from the held-out steps [steps_from, steps) (default: the last 5 %)."""
    dev = ds.X.device
    s0 = steps_from if steps_from is not None else ds.steps - max(1, ds.steps // 20)
    B = (nsamp + 127) // 128 * 128
    g = torch.Generator(device=dev).manual_seed(seed)
    idx = torch.empty((B, 2), dtype=torch.int32, device=dev)
    idx[:, 0] = torch.randint(s0, ds.steps, (B,), device=dev, generator=g, dtype=torch.int32)
    idx[:, 1] = torch.randint(0, ds.nenv, (B,), device=dev, generator=g, dtype=torch.int32)
    return idx


def rel_err_stats(err, tol):
    """Summary of per-sample rel. qacc errors (numpy float64)."""
    e = np.asarray(err, np.float64)
    return {"samples": int(e.size), "accept_frac": float((e <= tol).mean()), "median_rel_err": float(np.median(e)),
            "p90_rel_err": float(np.percentile(e, 90)), "max_rel_err": float(e.max()), "tol": tol,
            "accept_frac_at": {str(t): float((e <= t).mean()) for t in (1e-3, 1e-2, 1e-1, 1.0)}}


def evaluate(mlp, ds, steps_from=None, nsamp=65536, tol=1e-3, seed=1):
    """Held-out accuracy in physical units: rel. qacc error |a - b| / max(|b|, 1)
    (max over dofs) of the surrogate (the bf16 MFMA kernels) against ABA;
    returns a dict."""
    idx = heldout_indices(ds, steps_from, nsamp, seed)
    B = idx.shape[0]
    ws = mlp.workspace(B)
    mlp.gather(ds.X, ds.Y, idx, B, ds.nenv, ws["X16"], ws["T"])
    pred = mlp.forward(ws["X16"], B, train=False)[:, :NOUT]
    ystd = 1.0 / mlp.yinv
    qs = pred * ystd + mlp.ymean
    qa = ws["T"][:, :NOUT] * ystd + mlp.ymean
    err = ((qs - qa).abs() / qa.abs().clamp_min(1.0)).max(dim=1).values
    return rel_err_stats(err.double().cpu().numpy(), tol)


# ====================================================================== step
class SurrogateStepper:
    """qacc from the network for most envs, ABA as judge on a rotating
    1/period of the envs every step and as per-env fallback
    (g1_hip_surrogate_integrate).

    Fail-safe start: every env begins in fallback and is stepped by ABA
    until the judge has accepted the network on it; a rejected judgement puts
    it back.  Envs that the caller resets (mark_reset) start over in
    fallback.  So the network's qacc is integrated only for envs whose last
    judgement, at most period - 1 steps earlier, was within the tolerance.

    Deferred judging (defer=True, the fused path): a judged block whose envs
    are all in fallback is committed with ABA's qacc -- as a fallback env is
    whatever the judgement -- and records its network inputs and ABA's qacc;
    the records of a period are judged together by one network pass after
    the period's last step (g1_hip_surrogate_judge_batch), and the envs within
    the tolerance leave fallback then instead of the step after their
    judgement.  Blocks with an env off fallback are judged at once, so the
    guarantee above is unchanged, and a network that is never accepted costs
    one network pass per period instead of one tile latency per step.

    Fallback runs (fallback_runs=True, deferred judging, steps() only): an
    env leaves fallback only at a batch judgement, so a block whose envs are
    all in fallback at a period segment's start is ABA's until the period
    ends; steps() advances all such blocks through the segment with one
    g1_hip_fallback_run launch (each recorded at its judged step) and runs
    the per-step launches for the other blocks.  With skip_idle, while every
    env is in fallback (known at the start, then from a count after each
    batch judgement, read with one stream synchronisation per period) those
    per-step launches have nothing to do and are not issued; without it
    steps() never waits for the device.  `fallback` belongs to the stepper:
    code that writes it directly sets `_all_fallback.value = 0` before the
    next steps()."""

    def __init__(self, mlp, nenv, device, dt=2e-3, period=16, tol=1e-3, judge_flags=0, fused=True,
                 track_qacc=False, defer=True, fallback_runs=True, skip_idle=True):
        self.mlp, self.nenv, self.dt, self.period, self.tol = mlp, nenv, dt, period, tol
        self.fused = fused and mlp.fusable
        self.defer = defer and self.fused
        # steps(): all-fallback blocks stepped by one launch per period segment;
        # skip_idle: the per-step launches are left out while every env is in
        # fallback (one stream synchronisation per period reads the count)
        self.fallback_runs = fallback_runs and self.defer
        self.skip_idle = skip_idle and self.fallback_runs
        self.lib = _lib.load()
        self.dev = norm_device(device)
        self.rows = (nenv + 127) // 128 * 128
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.tau = torch.zeros((NV, nenv), **f32)
        self.qacc_aba = torch.zeros((NV, nenv), **f32)
        self.fallback = torch.ones(nenv, dtype=torch.uint8, device=self.dev)
        # [0] judged at once, [1] rejected at once, [2] max judged error (float
        # bits), [3] fallback env-steps, [4] / [5] batch-judged / rejected,
        # [6] envs off fallback after the last batch judgement of steps()
        self.stats = torch.zeros(8, dtype=torch.int32, device=self.dev)
        # known on the host: every env is in fallback (the fail-safe start);
        # steps() keeps it up to date, step() and resets leave it unknown or true
        self._all_fallback = C.c_int(1)
        if self.defer:
            self.snap = torch.zeros((SNAP_ROWS, nenv), **f32)
            self.sflag = torch.zeros(nenv, dtype=torch.uint8, device=self.dev)
            self.dropped = torch.zeros((), dtype=torch.int64, device=self.dev)     # records dropped by resets
            # blocks of 64 envs stepped by a fallback run in steps() (g1_hip_fallback_run)
            self.blk_run = torch.zeros((nenv + 63) // 64, dtype=torch.uint8, device=self.dev)
        self.X16 = torch.zeros((self.rows, PIN), dtype=torch.bfloat16, device=self.dev)
        self.pred = torch.zeros((self.rows, POUT), **f32)
        self.ystd = (1.0 / mlp.yinv).contiguous()
        self.judge_flags = judge_flags
        self.foot_pos = torch.zeros((6, nenv), **f32)
        self.foot_vel = torch.zeros((6, nenv), **f32)
        self.contact_forces = torch.zeros((8, nenv), **f32)
        # the qacc each step integrated (network or ABA), [35][nenv], when tracked
        self.qacc = torch.zeros((NV, nenv), **f32) if track_qacc else None
        self.t = 0

    def mark_reset(self, reset):
        """Envs with reset[e] != 0 (uint8 [nenv] on the device) return to
        fallback until their next accepted judgement."""
        check_soa("reset", reset.view(1, -1), 1, self.nenv, self.dev, dtype=torch.uint8)
        self.fallback.bitwise_or_(reset)
        if self.defer:
            # a pending record would judge the state before the reset
            self.dropped += (self.sflag.to(torch.int64) * (reset != 0).to(torch.int64)).sum()
            self.sflag.mul_((reset == 0).to(torch.uint8))

    def step(self, qpos, qvel, anchor, ctrl, mu=None, prediction=None):
        """One step of every env.  prediction (optional): z-scored qacc rows
        (float32 [>= nenv][128], first 35 columns) used in place of the
        network's output, through the unfused commit -- for predictors other
        than this MLP."""
        lib, s, n = self.lib, torch.cuda.current_stream(self.dev).cuda_stream, self.nenv
        for name, t, rows in (("qpos", qpos, NQ), ("qvel", qvel, NV), ("anchor", anchor, 16), ("ctrl", ctrl, NV - 6)):
            check_soa(name, t, rows, n, self.dev)
        if mu is not None:
            check_soa("mu", mu.view(1, -1), 1, n, self.dev)
        phase = self.t % self.period
        ck = _lib.check
        self._all_fallback.value = 0        # a batch judgement here may accept envs: unknown
        qacc_out = None if self.qacc is None else self.qacc.data_ptr()
        defer = self.defer and prediction is None
        if defer:
            # the same, and every block whose envs are all in fallback is
            # committed (and recorded when judged) by this kernel
            ck(lib.g1_hip_judge_forces_commit(n, self.dt, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
                                              ctrl.data_ptr(), None if mu is None else mu.data_ptr(),
                                              self.fallback.data_ptr(), phase, self.period,
                                              self.qacc_aba.data_ptr(), self.tau.data_ptr(),
                                              self.contact_forces.data_ptr(), self.foot_pos.data_ptr(),
                                              self.foot_vel.data_ptr(), qacc_out, self.snap.data_ptr(),
                                              self.sflag.data_ptr(), self.stats.data_ptr(), self.judge_flags, s),
               "judge_forces_commit")
        else:
            # ABA qacc of the judged and fallback envs, and everyone's
            # tau_total, contact update and feet, in one kernel
            ck(lib.g1_hip_judge_forces(n, self.dt, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
                                       ctrl.data_ptr(), None if mu is None else mu.data_ptr(),
                                       self.fallback.data_ptr(), phase, self.period, self.qacc_aba.data_ptr(),
                                       self.tau.data_ptr(), self.contact_forces.data_ptr(),
                                       self.foot_pos.data_ptr(), self.foot_vel.data_ptr(), self.judge_flags, s),
               "judge_forces")
        m = self.mlp
        if prediction is not None and (prediction.dtype != torch.float32 or prediction.device != self.dev
                                       or not prediction.is_contiguous() or prediction.dim() != 2
                                       or prediction.shape[0] < n or prediction.shape[1] != POUT):
            raise ValueError(f"prediction must be a contiguous float32 (>= {n}, {POUT}) tensor on {self.dev}")
        if defer:
            W, b = m._wb()
            ck(lib.g1_hip_surrogate_step_deferred(
                n, self.dt, phase, self.period, self.tol, qpos.data_ptr(), qvel.data_ptr(), self.tau.data_ptr(),
                m.xmean.data_ptr(), m.xinv.data_ptr(), W, b, m.ymean.data_ptr(), self.ystd.data_ptr(),
                self.qacc_aba.data_ptr(), self.fallback.data_ptr(), self.stats.data_ptr(), qacc_out, s),
               "surrogate_step_deferred")
            if phase == self.period - 1:
                ck(lib.g1_hip_surrogate_judge_batch(
                    n, self.tol, m.xmean.data_ptr(), m.xinv.data_ptr(), W, b, m.ymean.data_ptr(),
                    self.ystd.data_ptr(), self.snap.data_ptr(), self.sflag.data_ptr(), self.fallback.data_ptr(),
                    self.stats.data_ptr(), s), "surrogate_judge_batch")
        elif self.fused and prediction is None:
            W, b = m._wb()
            ck(lib.g1_hip_surrogate_step(n, self.dt, phase, self.period, self.tol, qpos.data_ptr(), qvel.data_ptr(),
                                         self.tau.data_ptr(), m.xmean.data_ptr(), m.xinv.data_ptr(), W, b,
                                         m.ymean.data_ptr(), self.ystd.data_ptr(), self.qacc_aba.data_ptr(),
                                         self.fallback.data_ptr(), self.stats.data_ptr(), qacc_out, s),
               "surrogate_step")
        else:
            pred = prediction
            if pred is None:
                rc = lib.g1_mlp_features(n, self.rows, qpos.data_ptr(), qvel.data_ptr(), self.tau.data_ptr(),
                                         m.xmean.data_ptr(), m.xinv.data_ptr(), self.X16.data_ptr(), PIN, s)
                ck(rc, "features")
                pred = m.forward(self.X16, self.rows, train=False, out=self.pred)
            ck(lib.g1_hip_surrogate_integrate(n, self.dt, phase, self.period, self.tol, qpos.data_ptr(),
                                              qvel.data_ptr(), pred.data_ptr(), POUT, m.ymean.data_ptr(),
                                              self.ystd.data_ptr(), self.qacc_aba.data_ptr(),
                                              self.fallback.data_ptr(), self.stats.data_ptr(), qacc_out, s),
               "surrogate_integrate")
        self.t += 1

    def steps(self, qpos, qvel, anchor, ctrl, n, mu=None):
        """n steps with ctrl (and mu) fixed: the launches of n step() calls,
        in the same order, from one host call (g1_hip_surrogate_steps) -- the
        per-step host work of step() otherwise bounds the stepper at small
        env counts.  The unfused configuration loops over step()."""
        if n < 0:
            raise ValueError("n must be >= 0")
        if not self.fused:
            for _ in range(n):
                self.step(qpos, qvel, anchor, ctrl, mu)
            return
        lib, s, ne = self.lib, torch.cuda.current_stream(self.dev).cuda_stream, self.nenv
        for name, t, rows in (("qpos", qpos, NQ), ("qvel", qvel, NV), ("anchor", anchor, 16), ("ctrl", ctrl, NV - 6)):
            check_soa(name, t, rows, ne, self.dev)
        if mu is not None:
            check_soa("mu", mu.view(1, -1), 1, ne, self.dev)
        m = self.mlp
        W, b = m._wb()
        ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        if self.defer:
            _lib.check(lib.g1_hip_surrogate_steps_deferred(
                ne, n, self.t, self.dt, self.period, self.tol, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
                ctrl.data_ptr(), ptr(mu), m.xmean.data_ptr(), m.xinv.data_ptr(), W, b, m.ymean.data_ptr(),
                self.ystd.data_ptr(), self.qacc_aba.data_ptr(), self.tau.data_ptr(), self.fallback.data_ptr(),
                self.stats.data_ptr(), ptr(self.qacc), self.contact_forces.data_ptr(), self.foot_pos.data_ptr(),
                self.foot_vel.data_ptr(), self.judge_flags, self.snap.data_ptr(), self.sflag.data_ptr(),
                self.blk_run.data_ptr() if self.fallback_runs else None,
                C.byref(self._all_fallback) if self.skip_idle else None, s), "surrogate_steps_deferred")
            self.t += n
            return
        _lib.check(lib.g1_hip_surrogate_steps(ne, n, self.t, self.dt, self.period, self.tol, qpos.data_ptr(),
                                              qvel.data_ptr(), anchor.data_ptr(), ctrl.data_ptr(), ptr(mu),
                                              m.xmean.data_ptr(), m.xinv.data_ptr(), W, b, m.ymean.data_ptr(),
                                              self.ystd.data_ptr(), self.qacc_aba.data_ptr(), self.tau.data_ptr(),
                                              self.fallback.data_ptr(), self.stats.data_ptr(), ptr(self.qacc),
                                              self.contact_forces.data_ptr(), self.foot_pos.data_ptr(),
                                              self.foot_vel.data_ptr(), self.judge_flags, s), "surrogate_steps")
        self.t += n

    def steps_schedule(self, qpos, qvel, anchor, ctrl_blocks, step0, n, period, mu=None):
        """n steps with a PD-target schedule (g1_hip_surrogate_steps_schedule):
        step s uses ctrl_blocks[(step0 + s) // period] (float32
        [nblocks][29][nenv]); the same results as one steps() call per target
        block, with fallback runs that switch blocks inside the kernel."""
        if not self.defer:
            raise ValueError("steps_schedule needs deferred judging")
        if n < 0 or step0 < 0 or period <= 0:
            raise ValueError("n >= 0, step0 >= 0 and period > 0 required")
        ne = self.nenv
        for name, t, rows in (("qpos", qpos, NQ), ("qvel", qvel, NV), ("anchor", anchor, 16)):
            check_soa(name, t, rows, ne, self.dev)
        if (ctrl_blocks.dtype != torch.float32 or ctrl_blocks.device != self.dev or ctrl_blocks.dim() != 3
                or tuple(ctrl_blocks.shape[1:]) != (NV - 6, ne) or not ctrl_blocks.is_contiguous()
                or (step0 + max(n, 1) - 1) // period >= ctrl_blocks.shape[0]):
            raise ValueError(f"ctrl_blocks must be a contiguous float32 [nblocks][29][{ne}] tensor on {self.dev} "
                             f"covering steps {step0} .. {step0 + n - 1}")
        if mu is not None:
            check_soa("mu", mu.view(1, -1), 1, ne, self.dev)
        lib, s = self.lib, torch.cuda.current_stream(self.dev).cuda_stream
        m = self.mlp
        W, b = m._wb()
        ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        _lib.check(lib.g1_hip_surrogate_steps_schedule(
            ne, n, self.t, self.dt, self.period, self.tol, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
            ctrl_blocks.data_ptr(), ctrl_blocks.shape[1] * ctrl_blocks.shape[2], period, step0, ptr(mu),
            m.xmean.data_ptr(), m.xinv.data_ptr(), W, b, m.ymean.data_ptr(), self.ystd.data_ptr(),
            self.qacc_aba.data_ptr(), self.tau.data_ptr(), self.fallback.data_ptr(), self.stats.data_ptr(),
            ptr(self.qacc), self.contact_forces.data_ptr(), self.foot_pos.data_ptr(), self.foot_vel.data_ptr(),
            self.judge_flags, self.snap.data_ptr(), self.sflag.data_ptr(),
            self.blk_run.data_ptr() if self.fallback_runs else None,
            C.byref(self._all_fallback) if self.skip_idle else None, s), "surrogate_steps_schedule")
        self.t += n

    def summary(self):
        """judged / rejected count every judgement (at once and in batches);
        aba_frac is the share of env-steps that integrated ABA's qacc (judged
        at once, or in fallback -- a batch-judged env-step was in fallback)."""
        st = self.stats.cpu().numpy().astype(np.uint32)
        now, now_rej, fb, batch, batch_rej = int(st[0]), int(st[1]), int(st[3]), int(st[4]), int(st[5])
        judged, rejected = now + batch, now_rej + batch_rej
        maxerr = float(np.array([st[2]], np.uint32).view(np.float32)[0])
        total = self.t * self.nenv
        return {"env_steps": total, "judged": judged, "rejected": rejected,
                "reject_frac": rejected / max(judged, 1), "fallback_env_steps": fb,
                "aba_frac": (now + fb) / max(total, 1), "max_judged_rel_err": maxerr,
                "fallback_envs_now": int(self.fallback.sum().item()), "batch_judged": batch,
                "pending_records": int(self.sflag.sum().item()) if self.defer else 0,
                "dropped_records": int(self.dropped.item()) if self.defer else 0}
