"""BF16 MFMA kernels of the surrogate MLP against CPU references.

GEMMs: every supported (a_trans, b_trans, epilogue) against a float64 CPU
product of the same bf16-rounded operands (tolerance: fp32 accumulation over
R terms, ~1e-5 relative).  Training: the full forward/backward of a small
batch against a float64 CPU reference that applies the same bf16 roundings
(activations and gradients are rounded to bf16 where the kernels store them
as GEMM operands), and one Adam step against its closed form.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from g1sim_amd import _lib
from g1sim_amd.surrogate import EPI, MLP, NIN, NOUT, PIN, POUT

pytestmark = pytest.mark.gpu


def bf16(a):
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def run_gemm(device, a_t, b_t, epi, A, B, I, J, R, split=1, bias=None, Z=None, C0=None):
    lib = _lib.load()
    At = torch.as_tensor(A.astype(np.float32), device=device).to(torch.bfloat16).contiguous()
    Bt = torch.as_tensor(B.astype(np.float32), device=device).to(torch.bfloat16).contiguous()
    C = torch.zeros((I, J), dtype=torch.float32, device=device) if C0 is None else \
        torch.as_tensor(C0.astype(np.float32), device=device)
    H = torch.zeros((I, J), dtype=torch.bfloat16, device=device)
    bt = None if bias is None else torch.as_tensor(bias.astype(np.float32), device=device)
    Zt = None if Z is None else torch.as_tensor(Z.astype(np.float32), device=device)
    lda = A.shape[1]
    ldb = B.shape[1]
    rc = lib.g1_mlp_gemm(a_t, b_t, EPI[epi], At.data_ptr(), lda, Bt.data_ptr(), ldb, I, J, R, split, C.data_ptr(),
                         H.data_ptr(), J, None if bt is None else bt.data_ptr(), None if Zt is None else Zt.data_ptr(),
                         torch.cuda.current_stream(device).cuda_stream)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return C.cpu().numpy().astype(np.float64), H.to(torch.float32).cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("a_t,b_t,epi,split", [
    (0, 0, "F32", 1), (0, 0, "BIAS_F32", 1), (0, 0, "BIAS_SILU", 1), (0, 0, "BIAS_SILU_H", 1),
    (0, 0, "SILU_BWD", 1), (0, 1, "F32", 1), (0, 1, "SILU_BWD", 1), (1, 1, "F32", 1), (1, 1, "F32_ATOMIC", 4)])
def test_gemm_variants(device, a_t, b_t, epi, split):
    rng = np.random.default_rng(100 * a_t + 10 * b_t + EPI[epi])
    I, J, R = 256, 384, 256
    Aop = rng.normal(size=(I, R))
    Bop = rng.normal(size=(J, R))
    A = Aop.T.copy() if a_t else Aop          # stored layouts
    B = Bop.T.copy() if b_t else Bop
    bias = rng.normal(size=J)
    Z = rng.normal(size=(I, J))
    C0 = rng.normal(size=(I, J)) if epi == "F32_ATOMIC" else None
    C, H = run_gemm(device, a_t, b_t, epi, A, B, I, J, R, split=split, bias=bias, Z=Z, C0=C0)
    acc = bf16(Aop) @ bf16(Bop).T
    tol = 1e-5 * np.abs(bf16(Aop)) @ np.abs(bf16(Bop)).T + 1e-6
    if epi == "F32":
        assert (np.abs(C - acc) <= tol).all()
    elif epi == "F32_ATOMIC":
        assert (np.abs(C - (C0.astype(np.float32) + acc)) <= tol + 1e-5).all()
    elif epi == "BIAS_F32":
        assert (np.abs(C - (acc + bias.astype(np.float32))) <= tol + 1e-6).all()
    elif epi in ("BIAS_SILU", "BIAS_SILU_H"):
        z = acc + bias.astype(np.float32)
        if epi == "BIAS_SILU":
            assert (np.abs(C - z) <= tol + 1e-6).all()
        assert np.allclose(H, bf16(silu(z)), rtol=1e-2, atol=1e-2)
        assert (np.abs(H - silu(z)) <= 8e-3 * np.abs(silu(z)) + tol + 1e-3).all()
    elif epi == "SILU_BWD":
        d = acc * dsilu(Z.astype(np.float32))
        assert (np.abs(C - d) <= tol * 1.2 + 1e-6).all()
        assert (np.abs(H - d) <= 8e-3 * np.abs(d) + tol + 1e-6).all()


def test_gemm_rejects_bad_shapes(device):
    lib = _lib.load()
    d = torch.zeros(16, device=device)
    rc = lib.g1_mlp_gemm(0, 0, EPI["F32"], d.data_ptr(), 32, d.data_ptr(), 32, 100, 128, 32, 1, d.data_ptr(),
                         None, 128, None, None, None)
    assert rc == -1
    rc = lib.g1_mlp_gemm(0, 0, EPI["F32"], d.data_ptr(), 32, d.data_ptr(), 32, 128, 128, 32, 2, d.data_ptr(),
                         None, 128, None, None, None)
    assert rc == -1          # split only with atomic accumulation


def _cpu_reference(mlp, X, T):
    """float64 forward/backward with the kernels' bf16 roundings."""
    W = [bf16(w.cpu().numpy()) for w in mlp.W]
    b = [bb.cpu().numpy().astype(np.float64) for bb in mlp.b]
    B = X.shape[0]
    h = [bf16(X)]
    zs = []
    for l in range(len(W) - 1):
        z = h[-1] @ W[l].T + b[l]
        zs.append(z)
        h.append(bf16(silu(z)))
    pred = h[-1] @ W[-1].T + b[-1]
    d = np.zeros_like(pred)
    d[:, :NOUT] = 2.0 * (pred[:, :NOUT] - T[:, :NOUT]) / (B * NOUT)
    gW, gb = [None] * len(W), [None] * len(W)
    d16 = bf16(d)
    gb[-1] = d.sum(0)
    for l in range(len(W) - 1, -1, -1):
        gW[l] = d16.T @ h[l]
        if l > 0:
            dz = (d16 @ W[l]) * dsilu(zs[l - 1])
            gb[l - 1] = dz.sum(0)
            d16 = bf16(dz)
    return pred, gW, gb


def test_forward_backward_against_cpu(device):
    mlp = MLP(device, seed=3)
    B = 256
    rng = np.random.default_rng(0)
    X = np.zeros((B, PIN)); X[:, :NIN] = rng.normal(size=(B, NIN))
    T = np.zeros((B, POUT)); T[:, :NOUT] = rng.normal(size=(B, NOUT))
    ws = mlp.workspace(B)
    ws["X16"].copy_(torch.as_tensor(X.astype(np.float32), device=device).to(torch.bfloat16))
    ws["T"].copy_(torch.as_tensor(T.astype(np.float32), device=device))
    pred = mlp.forward(ws["X16"], B, train=True)
    mlp.loss_grad(pred, ws["T"], B)
    mlp.backward(ws["X16"], B, ws["dY16"])
    mlp.gradients(B)
    torch.cuda.synchronize()
    rpred, rgW, rgb = _cpu_reference(mlp, X, T)
    assert np.abs(pred.cpu().numpy()[:, :NOUT] - rpred[:, :NOUT]).max() < 2e-3 * max(1, np.abs(rpred).max())
    for l in range(len(mlp.W)):
        g = ws["gW"][l].cpu().numpy()
        scale = np.abs(rgW[l]).max()
        assert np.abs(g - rgW[l]).max() < 2e-2 * scale, l
        gb = ws["gb"][l].cpu().numpy()
        assert np.abs(gb - rgb[l]).max() < 2e-2 * max(np.abs(rgb[l]).max(), 1e-8), l
        # padding stays exactly zero
        if l == 0:
            assert (g[:, NIN:] == 0).all()
        if l == len(mlp.W) - 1:
            assert (g[NOUT:] == 0).all() and (gb[NOUT:] == 0).all()


def test_adam_step_closed_form(device):
    lib = _lib.load()
    n = 1000
    rng = np.random.default_rng(1)
    p = rng.normal(size=n).astype(np.float32)
    g = rng.normal(size=n).astype(np.float32)
    mask = (rng.random(n) < 0.9).astype(np.uint8)
    p[mask == 0] = 0
    t = {k: torch.as_tensor(v, device=device) for k, v in (("p", p), ("g", g), ("mask", mask))}
    m = torch.zeros(n, device=device); v = torch.zeros(n, device=device)
    p16 = torch.zeros(n, dtype=torch.bfloat16, device=device)
    for step in (1, 2):
        rc = lib.g1_mlp_adam(t["p"].data_ptr(), m.data_ptr(), v.data_ptr(), t["g"].data_ptr(), p16.data_ptr(),
                             t["mask"].data_ptr(), n, 1e-3, 0.9, 0.999, 1e-8, step, 0.0,
                             torch.cuda.current_stream(device).cuda_stream)
        assert rc == 0
    torch.cuda.synchronize()
    # closed form after two identical gradients
    m1 = 0.1 * g; v1 = 0.001 * g * g
    p1 = p - 1e-3 * (m1 / 0.1) / (np.sqrt(v1 / 0.001) + 1e-8)
    m2 = 0.9 * m1 + 0.1 * g; v2 = 0.999 * v1 + 0.001 * g * g
    p2 = p1 - 1e-3 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    p2[mask == 0] = 0
    assert np.allclose(t["p"].cpu().numpy(), p2, rtol=1e-5, atol=1e-7)
    assert (p16.to(torch.float32).cpu().numpy()[mask == 0] == 0).all()


def test_gather_normalises(device):
    lib = _lib.load()
    rng = np.random.default_rng(2)
    steps, nenv, B = 3, 200, 128
    X = rng.normal(size=(steps, NIN, nenv)).astype(np.float32)
    Y = rng.normal(size=(steps, NOUT, nenv)).astype(np.float32)
    idx = np.stack([rng.integers(0, steps, B), rng.integers(0, nenv, B)], 1).astype(np.int32)
    xm, xs = rng.normal(size=NIN).astype(np.float32), rng.uniform(0.5, 2, NIN).astype(np.float32)
    ym, ys = rng.normal(size=NOUT).astype(np.float32), rng.uniform(0.5, 2, NOUT).astype(np.float32)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device)  # noqa: E731
    Xn = torch.zeros((B, PIN), dtype=torch.bfloat16, device=device)
    Yn = torch.zeros((B, POUT), device=device)
    tX, tY, ti = d(X), d(Y), d(idx)
    txm, txi, tym, tyi = d(xm), d(1 / xs), d(ym), d(1 / ys)
    rc = lib.g1_mlp_gather(tX.data_ptr(), tY.data_ptr(), ti.data_ptr(), B, NIN, NOUT, nenv, txm.data_ptr(),
                           txi.data_ptr(), tym.data_ptr(), tyi.data_ptr(), Xn.data_ptr(), PIN, Yn.data_ptr(), POUT,
                           torch.cuda.current_stream(device).cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    ex = (X[idx[:, 0], :, idx[:, 1]] - xm) / xs
    ey = (Y[idx[:, 0], :, idx[:, 1]] - ym) / ys
    got = Xn.to(torch.float32).cpu().numpy()
    assert np.allclose(got[:, :NIN], bf16(ex), rtol=0, atol=1e-6) or np.abs(got[:, :NIN] - ex).max() < 2e-2
    assert (got[:, NIN:] == 0).all()
    assert np.allclose(Yn.cpu().numpy()[:, :NOUT], ey, rtol=1e-5, atol=1e-5)
    # the sample-major rows give the same bits
    S = d(np.concatenate([X, Y], axis=1).transpose(0, 2, 1))
    Xn2 = torch.full((B, PIN), 3.0, dtype=torch.bfloat16, device=device)
    Yn2 = torch.full((B, POUT), 3.0, device=device)
    rc = lib.g1_mlp_gather_rows(S.data_ptr(), ti.data_ptr(), B, NIN, NOUT, nenv, txm.data_ptr(), txi.data_ptr(),
                                tym.data_ptr(), tyi.data_ptr(), Xn2.data_ptr(), PIN, Yn2.data_ptr(), POUT,
                                torch.cuda.current_stream(device).cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(Xn2, Xn) and torch.equal(Yn2, Yn)


def test_gemm_split_slices_sum_to_product(device):
    """F32_SPLIT writes the r-slices of the weight-gradient GEMM; summing them
    (g1_mlp_sum_splits) gives the full product."""
    lib = _lib.load()
    rng = np.random.default_rng(7)
    I, J, R, split = 256, 128, 1024, 4
    Aop, Bop = rng.normal(size=(I, R)), rng.normal(size=(J, R))
    At = torch.as_tensor(Aop.T.copy().astype(np.float32), device=device).to(torch.bfloat16).contiguous()
    Bt = torch.as_tensor(Bop.T.copy().astype(np.float32), device=device).to(torch.bfloat16).contiguous()
    parts = torch.full((split, I, J), 7.0, device=device)
    s = torch.cuda.current_stream(device).cuda_stream
    rc = lib.g1_mlp_gemm(1, 1, EPI["F32_SPLIT"], At.data_ptr(), I, Bt.data_ptr(), J, I, J, R, split, parts.data_ptr(),
                         None, J, None, None, s)
    assert rc == 0, _lib.last_error()
    out = torch.full((I, J), -1.0, device=device)
    assert lib.g1_mlp_sum_splits(parts.data_ptr(), split, I * J, out.data_ptr(), s) == 0
    torch.cuda.synchronize()
    acc = bf16(Aop) @ bf16(Bop).T
    tol = 1e-5 * np.abs(bf16(Aop)) @ np.abs(bf16(Bop)).T + 1e-6
    assert (np.abs(out.cpu().numpy() - acc) <= tol).all()
    # each slice is the product over its quarter of r
    q = R // split
    p1 = bf16(Aop[:, q:2 * q]) @ bf16(Bop[:, q:2 * q]).T
    assert np.allclose(parts[1].cpu().numpy(), p1, rtol=1e-5, atol=1e-4)


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def test_colsum_det_is_exact_and_repeatable(device):
    lib = _lib.load()
    rng = np.random.default_rng(5)
    B, ld = 1000, 200
    dZ = rng.normal(size=(B, ld)).astype(np.float32)
    tz = torch.as_tensor(dZ, device=device)
    part = torch.zeros(((B + 255) // 256) * ld, device=device)
    outs = []
    for _ in range(2):
        db = torch.full((ld,), 7.0, device=device)           # overwritten, not accumulated
        assert lib.g1_mlp_colsum_det(tz.data_ptr(), B, ld, part.data_ptr(), db.data_ptr(), _stream(device)) == 0
        outs.append(db.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    np.testing.assert_allclose(outs[0], dZ.astype(np.float64).sum(axis=0), rtol=1e-5, atol=1e-4)


def test_train_schedule_and_sampler(device):
    lib = _lib.load()
    st = torch.zeros(4, dtype=torch.int32, device=device)
    st[1] = 10                                                # Adam steps taken before this run
    sched = torch.zeros(4, device=device)
    B, lo, hi, nenv, iters = 4096, 3, 40, 1000, 7
    idx = torch.zeros(2 * B, dtype=torch.int32, device=device)
    draws = []
    for k in range(iters):
        assert lib.g1_mlp_train_schedule(st.data_ptr(), sched.data_ptr(), iters, 1e-3, 1e-5, 0.9, 0.999,
                                         _stream(device)) == 0
        assert lib.g1_mlp_sample(B, lo, hi, nenv, 99, st.data_ptr(), idx.data_ptr(), _stream(device)) == 0
        s = st.cpu().numpy()
        assert list(s[:3]) == [k + 1, 11 + k, k]
        lr, bc1, bc2 = sched.cpu().numpy()[:3]
        t = 11 + k
        np.testing.assert_allclose(lr, 1e-5 + 0.5 * (1e-3 - 1e-5) * (1 + np.cos(np.pi * k / (iters - 1))), rtol=1e-5)
        # b1, b2 as the float32 arguments they are (1 - b^t cancels: float32(0.999) != 0.999)
        b1, b2 = (float(np.float32(b)) for b in (0.9, 0.999))
        np.testing.assert_allclose([bc1, bc2], [1 - b1 ** t, 1 - b2 ** t], rtol=1e-5)
        d = idx.cpu().numpy().reshape(B, 2)
        assert d[:, 0].min() >= lo and d[:, 0].max() < hi and d[:, 1].min() >= 0 and d[:, 1].max() < nenv
        draws.append(d)
    # uniform over the step range (loose chi-square), different every iteration
    counts = np.bincount(draws[0][:, 0] - lo, minlength=hi - lo)
    expect = B / (hi - lo)
    assert ((counts - expect) ** 2 / expect).sum() < 3 * (hi - lo)
    assert not np.array_equal(draws[0], draws[1])
    # the same (seed, iteration) draws the same batch
    st2 = torch.tensor([0, 0, 0, 0], dtype=torch.int32, device=device)
    idx2 = torch.zeros_like(idx)
    assert lib.g1_mlp_sample(B, lo, hi, nenv, 99, st2.data_ptr(), idx2.data_ptr(), _stream(device)) == 0
    assert np.array_equal(idx2.cpu().numpy().reshape(B, 2), draws[0])


def test_adam_multi_matches_single_tensor_adam(device):
    """The multi-tensor step with sched = (lr, 1 - b1^t, 1 - b2^t) gives the
    per-tensor kernel's bits, masks and bf16 copies included."""
    lib = _lib.load()
    rng = np.random.default_rng(6)
    sizes = [1000, 37, 4096]
    mk = lambda a: torch.as_tensor(a, device=device)  # noqa: E731
    ps = [rng.normal(size=n).astype(np.float32) for n in sizes]
    gs = [mk(rng.normal(size=n).astype(np.float32)) for n in sizes]
    masks = [mk((rng.random(n) < 0.8).astype(np.uint8)) for n in sizes]
    A = [mk(p.copy()) for p in ps]; Am = [torch.zeros(n, device=device) for n in sizes]
    Av = [torch.zeros(n, device=device) for n in sizes]; A16 = [torch.zeros(n, dtype=torch.bfloat16, device=device)
                                                              for n in sizes]
    Bp = [mk(p.copy()) for p in ps]; Bm = [torch.zeros(n, device=device) for n in sizes]
    Bv = [torch.zeros(n, device=device) for n in sizes]; B16 = [torch.zeros(n, dtype=torch.bfloat16, device=device)
                                                              for n in sizes]
    sched = torch.zeros(4, device=device)
    n_arr = (C.c_int * 3)(*sizes)
    # transposed copies: the 4096 tensor as 64 x 64, the 1000 one as 8 x 125; none for the 37
    cols = (C.c_int * 3)(125, 1, 64)
    B16T = [torch.zeros(n, dtype=torch.bfloat16, device=device) for n in sizes]
    B16T[1] = None
    for step in (1, 2, 3):
        for k in range(3):
            assert lib.g1_mlp_adam(A[k].data_ptr(), Am[k].data_ptr(), Av[k].data_ptr(), gs[k].data_ptr(),
                                   A16[k].data_ptr(), masks[k].data_ptr(), sizes[k], 1e-3, 0.9, 0.999, 1e-8, step,
                                   0.0, _stream(device)) == 0
        # the bias corrections g1_mlp_adam computes on the host (to float rounding)
        bc = torch.tensor([1e-3, 1 - np.float32(0.9) ** step, 1 - np.float32(0.999) ** step, 0],
                          dtype=torch.float32)
        sched.copy_(bc)
        rc = lib.g1_mlp_adam_multi(3, _lib.ptr_array(Bp), _lib.ptr_array(Bm), _lib.ptr_array(Bv), _lib.ptr_array(gs),
                                   _lib.ptr_array(B16), _lib.ptr_array(B16T), cols, _lib.ptr_array(masks), n_arr,
                                   None, sched.data_ptr(), 0.9, 0.999, 1e-8, _stream(device))
        assert rc == 0
    torch.cuda.synchronize()
    for k in range(3):
        np.testing.assert_allclose(Bp[k].cpu().numpy(), A[k].cpu().numpy(), rtol=2e-6, atol=1e-9)
        assert np.array_equal(Bp[k].cpu().numpy()[masks[k].cpu().numpy() == 0], ps[k][masks[k].cpu().numpy() == 0])
        assert (B16[k].float().cpu().numpy()[masks[k].cpu().numpy() == 0] == 0).all()
        assert np.array_equal(B16[k].float().cpu().numpy(), Bp[k].to(torch.bfloat16).float().cpu().numpy()
                              * (masks[k].cpu().numpy() != 0))
        if B16T[k] is not None:
            c = cols[k]
            assert np.array_equal(B16T[k].float().cpu().numpy().reshape(c, -1),
                                  B16[k].float().cpu().numpy().reshape(-1, c).T)


@pytest.mark.parametrize("B", [512, 2048])
def test_graph_replays_equal_eager_iterations(device, B):
    """One training iteration captured as a HIP graph and replayed, with eager
    work and synchronisations between the replays, gives after every replay
    the same weights, bit for bit, as eager iterations from the same start
    (2,048 samples: 32 bias partials per layer, Adam's one-wave-per-element
    sum)."""
    from g1sim_amd.surrogate import fit_normalization, generate_dataset
    ds = generate_dataset(device, nenv=1024, steps=12, seed=41)
    iters = 6
    nets = [MLP(device, seed=3) for _ in range(2)]
    for m in nets:
        fit_normalization(m, ds)
    idx = [torch.zeros(2 * B, dtype=torch.int32, device=device) for _ in range(2)]
    run = lambda m, ix: m.iteration(ds, ix, B, 0, 10, 5, iters, 1e-3, 1e-5)  # noqa: E731
    eager = []
    for _ in range(iters):
        run(nets[0], idx[0])
        torch.cuda.synchronize()
        eager.append([w.cpu().numpy().copy() for w in nets[0].W])
    run(nets[1], idx[1])                            # iteration 0 eagerly, then capture
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run(nets[1], idx[1])
    assert all(np.array_equal(a, w.cpu().numpy()) for a, w in zip(eager[0], nets[1].W))
    for k in range(1, iters):
        g.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device=device).sum()   # eager work between replays
        assert np.isfinite(junk.item())
        got = [w.cpu().numpy() for w in nets[1].W]
        for l, (a, b) in enumerate(zip(eager[k], got)):
            assert np.array_equal(a, b), (k, l)
    assert int(nets[1].train_state[1].item()) == iters


@pytest.mark.parametrize("a_t,b_t,epi", [(0, 0, "BIAS_SILU"), (0, 1, "SILU_BWD"), (1, 1, "F32")])
def test_gemm_128_tiles(device, a_t, b_t, epi):
    """A grid large enough for the 128 x 128 tiles (the small shapes above
    run the 64 x 64 tiles)."""
    rng = np.random.default_rng(7 + EPI[epi])
    I, J, R = 8192, 1024, 64                    # (I / 128) (J / 128) = 512 workgroups
    Aop = rng.normal(size=(I, R))
    Bop = rng.normal(size=(J, R))
    A = Aop.T.copy() if a_t else Aop
    B = Bop.T.copy() if b_t else Bop
    bias = rng.normal(size=J)
    Z = rng.normal(size=(I, J))
    C, H = run_gemm(device, a_t, b_t, epi, A, B, I, J, R, bias=bias, Z=Z)
    acc = bf16(Aop) @ bf16(Bop).T
    tol = 1e-5 * np.abs(bf16(Aop)) @ np.abs(bf16(Bop)).T + 1e-6
    if epi == "F32":
        assert (np.abs(C - acc) <= tol).all()
    elif epi == "BIAS_SILU":
        assert (np.abs(C - (acc + bias.astype(np.float32))) <= tol + 1e-6).all()
    else:
        d = acc * dsilu(Z.astype(np.float32))
        assert (np.abs(C - d) <= tol * 1.2 + 1e-6).all()


@pytest.mark.parametrize("B,rows,cols,split", [(8192, 512, 512, 8), (1024, 128, 128, 1), (2048, 128, 512, 2),
                                                (16384, 512, 512, 32), (32768, 512, 512, 8)])
def test_backward_layer_matches_two_gemms(device, B, rows, cols, split):
    """g1_mlp_backward_layer gives the bits of its two g1_mlp_gemm_ex calls:
    one grid of 64 x 64 tiles (the training shape, 8,192 samples, and two
    small ones), one of 128 x 128 tiles (16,384 samples in 32 r-slices), and
    two launches where the tile sizes differ (32,768 samples)."""
    lib = _lib.load()
    rng = np.random.default_rng(B + rows + cols + split)
    s = _stream(device)
    bf = lambda *shape: torch.as_tensor(rng.normal(size=shape).astype(np.float32), device=device).to(torch.bfloat16)  # noqa: E731
    d16, hin, WT16 = bf(B, rows), bf(B, cols), bf(cols, rows)
    Z = torch.as_tensor(rng.normal(size=(B, cols)).astype(np.float32), device=device)
    outs = []
    for fused, fill in ((False, 0.0), (True, 7.0)):
        gW = torch.full((split, rows * cols), fill, device=device)
        dZ = torch.full((B, cols), fill, device=device).to(torch.bfloat16)
        cp = torch.full((B // 64, cols), fill, device=device)
        if fused:
            rc = lib.g1_mlp_backward_layer(d16.data_ptr(), rows, hin.data_ptr(), cols, WT16.data_ptr(), B, split,
                                           gW.data_ptr(), dZ.data_ptr(), cols, Z.data_ptr(), cp.data_ptr(), s)
            assert rc == 0, _lib.last_error()
        else:
            assert lib.g1_mlp_gemm_ex(1, 1, EPI["F32_SPLIT"], d16.data_ptr(), rows, hin.data_ptr(), cols, rows, cols,
                                      B, split, gW.data_ptr(), None, cols, None, None, None, s) == 0
            assert lib.g1_mlp_gemm_ex(0, 0, EPI["SILU_BWD"], d16.data_ptr(), rows, WT16.data_ptr(), rows, B, cols,
                                      rows, 1, None, dZ.data_ptr(), cols, None, Z.data_ptr(), cp.data_ptr(), s) == 0
        torch.cuda.synchronize()
        outs.append((gW.cpu(), dZ.float().cpu(), cp.cpu()))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    # the data gradient's bias partials are required (the fused form stores no fp32 gradient)
    assert lib.g1_mlp_backward_layer(d16.data_ptr(), rows, hin.data_ptr(), cols, WT16.data_ptr(), B, split,
                                     gW.data_ptr(), dZ.data_ptr(), cols, Z.data_ptr(), None, s) == -1


@pytest.mark.parametrize("I,J,R", [(256, 384, 256), (8192, 1024, 64)])
def test_silu_bwd_column_partials(device, I, J, R):
    """The SILU_BWD epilogue's bias-gradient partials: column sums of its fp32
    output over each 64-row block (both tile sizes: 256 x 384 runs 64 x 64
    tiles, 8192 x 1024 the 128 x 128 ones), and the same bf16 output with
    the fp32 one left out."""
    lib = _lib.load()
    rng = np.random.default_rng(11)
    Aop, Bop, Z = rng.normal(size=(I, R)), rng.normal(size=(J, R)), rng.normal(size=(I, J))
    s = _stream(device)
    At = torch.as_tensor(Aop.astype(np.float32), device=device).to(torch.bfloat16).contiguous()
    Bt = torch.as_tensor(Bop.T.copy().astype(np.float32), device=device).to(torch.bfloat16).contiguous()
    Zt = torch.as_tensor(Z.astype(np.float32), device=device)
    Cf = torch.zeros((I, J), device=device)
    H1 = torch.zeros((I, J), dtype=torch.bfloat16, device=device)
    H2 = torch.zeros_like(H1)
    part = torch.full((I // 64, J), 7.0, device=device)
    assert lib.g1_mlp_gemm(0, 1, EPI["SILU_BWD"], At.data_ptr(), R, Bt.data_ptr(), J, I, J, R, 1, Cf.data_ptr(),
                           H1.data_ptr(), J, None, Zt.data_ptr(), s) == 0
    assert lib.g1_mlp_gemm_ex(0, 1, EPI["SILU_BWD"], At.data_ptr(), R, Bt.data_ptr(), J, I, J, R, 1, None,
                              H2.data_ptr(), J, None, Zt.data_ptr(), part.data_ptr(), s) == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert torch.equal(H1, H2)
    c = Cf.cpu().numpy().astype(np.float64).reshape(I // 64, 64, J)
    got = part.cpu().numpy()
    np.testing.assert_allclose(got, c.sum(axis=1), rtol=1e-5, atol=1e-5 * np.abs(c).sum(axis=1).max())
    # colpart belongs to the SILU_BWD epilogue only
    assert lib.g1_mlp_gemm_ex(0, 1, EPI["F32"], At.data_ptr(), R, Bt.data_ptr(), J, I, J, R, 1, Cf.data_ptr(),
                              None, J, None, None, part.data_ptr(), s) == -1


def test_mse_grad_partials_match_atomic_version(device):
    """g1_mlp_mse_grad_part: the same bf16 dY as g1_mlp_mse_grad, 64-row column
    sums of the fp32 dY and per-block squared errors that add up to its SSE."""
    lib = _lib.load()
    rng = np.random.default_rng(12)
    B, ld, n_out = 1024, POUT, NOUT
    pred = rng.normal(size=(B, ld)).astype(np.float32)
    tgt = rng.normal(size=(B, ld)).astype(np.float32)
    tp, tt = torch.as_tensor(pred, device=device), torch.as_tensor(tgt, device=device)
    d16a = torch.zeros((B, ld), dtype=torch.bfloat16, device=device)
    d16b = torch.full((B, ld), 5.0, dtype=torch.bfloat16, device=device)
    d32 = torch.zeros((B, ld), device=device)
    sse = torch.zeros(1, device=device)
    part = torch.full((B // 64, ld), 7.0, device=device)
    sp = torch.full((B // 64, ld // 32), 7.0, device=device)
    s = _stream(device)
    assert lib.g1_mlp_mse_grad(tp.data_ptr(), tt.data_ptr(), B, ld, n_out, d16a.data_ptr(), d32.data_ptr(),
                               sse.data_ptr(), s) == 0
    assert lib.g1_mlp_mse_grad_part(tp.data_ptr(), tt.data_ptr(), B, ld, n_out, d16b.data_ptr(), part.data_ptr(),
                                    sp.data_ptr(), s) == 0
    torch.cuda.synchronize()
    assert torch.equal(d16a, d16b)
    d = d32.cpu().numpy().astype(np.float64)
    assert (d[:, n_out:] == 0).all()
    np.testing.assert_allclose(part.cpu().numpy(), d.reshape(B // 64, 64, ld).sum(axis=1), rtol=1e-5, atol=1e-9)
    e = np.zeros((B, ld))
    e[:, :n_out] = (pred[:, :n_out].astype(np.float64) - tgt[:, :n_out]) ** 2
    np.testing.assert_allclose(sp.cpu().numpy(), e.reshape(B // 64, 64, ld // 32, 32).sum(axis=(1, 3)), rtol=1e-5,
                               atol=1e-9)
    np.testing.assert_allclose(sp.cpu().numpy().sum(), sse.item(), rtol=1e-5)
    assert lib.g1_mlp_mse_grad_part(tp.data_ptr(), tt.data_ptr(), B + 1, ld, n_out, d16b.data_ptr(),
                                    part.data_ptr(), sp.data_ptr(), s) == -1        # B not a multiple of 64


def test_adam_over_gradient_slices(device):
    """g1_mlp_adam_multi_sum: a gradient given as up to 16 slices gives the
    bits of g1_mlp_sum_splits followed by the one-slice step; 128 or 200
    slices (one wave per element) match the step on the float64 sum of the
    slices."""
    lib = _lib.load()
    rng = np.random.default_rng(13)
    s = _stream(device)
    mk = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device)  # noqa: E731
    # 8 and 12 slices (all loads in flight, then 4 more one at a time), 1
    # slice; 128 and 200 slices (one wave per element; 200 leaves lanes with
    # an odd count)
    sizes, nslices = [3000, 512, 35, 700, 257, 77], [8, 128, 128, 12, 1, 200]
    nt = len(sizes)
    slices = [rng.normal(size=(k, n)).astype(np.float32) for n, k in zip(sizes, nslices)]
    p0 = [rng.normal(size=n).astype(np.float32) for n in sizes]
    masks = [mk((rng.random(n) < 0.9).astype(np.uint8)) for n in sizes]
    sched = mk(np.array([1e-3, 1 - 0.9, 1 - 0.999, 0], np.float32))
    n_arr = (C.c_int * nt)(*sizes)
    # reference: summed gradients (sum_splits for the 8-slice tensor, float64 for the wide ones), one-slice step
    summed = []
    for k in range(nt):
        if nslices[k] <= 16:
            out = torch.zeros(sizes[k], device=device)
            assert lib.g1_mlp_sum_splits(mk(slices[k]).data_ptr(), nslices[k], sizes[k], out.data_ptr(), s) == 0
            summed.append(out)
        else:
            summed.append(mk(slices[k].astype(np.float64).sum(axis=0).astype(np.float32)))
    runs = []
    for use_slices in (False, True):
        P = [mk(p.copy()) for p in p0]
        M = [torch.zeros(n, device=device) for n in sizes]
        V = [torch.zeros(n, device=device) for n in sizes]
        P16 = [torch.zeros(n, dtype=torch.bfloat16, device=device) for n in sizes]
        if use_slices:
            G = [mk(x) for x in slices]
            nsum = (C.c_int * nt)(*nslices)
            gstride = (C.c_longlong * nt)(*sizes)
            rc = lib.g1_mlp_adam_multi_sum(nt, _lib.ptr_array(P), _lib.ptr_array(M), _lib.ptr_array(V),
                                           _lib.ptr_array(G), _lib.ptr_array(P16), None, None,
                                           _lib.ptr_array(masks), n_arr, nsum, gstride, None, sched.data_ptr(), 0.9,
                                           0.999, 1e-8, s)
        else:
            rc = lib.g1_mlp_adam_multi(nt, _lib.ptr_array(P), _lib.ptr_array(M), _lib.ptr_array(V),
                                       _lib.ptr_array(summed), _lib.ptr_array(P16), None, None,
                                       _lib.ptr_array(masks), n_arr, None, sched.data_ptr(), 0.9, 0.999, 1e-8, s)
        assert rc == 0, _lib.last_error()
        torch.cuda.synchronize()
        runs.append(([x.cpu().numpy() for x in P], [x.cpu().numpy() for x in M]))
    (pa, ma), (pb, mb) = runs
    for k in range(nt):
        if nslices[k] <= 16:                                                      # up to 16 slices: same bits
            assert np.array_equal(pa[k], pb[k]) and np.array_equal(ma[k], mb[k]), k
    for k in (k for k in range(nt) if nslices[k] > 16):
        np.testing.assert_allclose(mb[k], ma[k], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(pb[k], pa[k], rtol=1e-6, atol=1e-7)
        mk_ = masks[k].cpu().numpy() == 0
        assert np.array_equal(pb[k][mk_], p0[k][mk_])
    # slices closer than the tensor length are refused
    bad = (C.c_longlong * 3)(sizes[0] - 1, sizes[1], sizes[2])
    assert lib.g1_mlp_adam_multi_sum(3, _lib.ptr_array(P), _lib.ptr_array(M), _lib.ptr_array(V),
                                     _lib.ptr_array([mk(x) for x in slices]), None, None, None, None, n_arr,
                                     (C.c_int * 3)(*nslices[:3]), bad, None, sched.data_ptr(), 0.9, 0.999, 1e-8, s) == -1
