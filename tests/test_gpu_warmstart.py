"""GPU tests of the warm-start network kernels (libwarmstart.so) against the
FP64 oracle (oracle/mlp.py) and exact integer products."""

import numpy as np
import pytest
import torch

from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws(device):
    from lpe import warmstart
    warmstart.load()
    return warmstart


def _int_bf16(rng, shape, lo=-8, hi=9):
    """Small integers: exact in BF16, products and sums exact in FP32."""
    return rng.integers(lo, hi, size=shape).astype(np.float32)


@pytest.mark.parametrize("trans_a,trans_b", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 45, 29), (128, 128, 32), (200, 130, 300), (257, 63, 96)])
def test_gemm_exact_on_integers(ws, device, trans_a, trans_b, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    a = _int_bf16(rng, (k, m) if trans_a else (m, k))
    b = _int_bf16(rng, (n, k) if trans_b else (k, n))
    want = (a.T if trans_a else a).astype(np.float64) @ (b.T if trans_b else b).astype(np.float64)
    ta = torch.tensor(a, device=device).to(torch.bfloat16).contiguous()
    tb = torch.tensor(b, device=device).to(torch.bfloat16).contiguous()
    got = ws.gemm(ta, tb, trans_a=trans_a, trans_b=trans_b)
    np.testing.assert_array_equal(got.cpu().numpy(), want.astype(np.float32))


@pytest.mark.parametrize("splits", [1, 2, 3, 8])
@pytest.mark.parametrize("m,n,k", [(96, 200, 4136), (130, 257, 1000), (1, 9, 72)])
def test_gemm_split_k_exact_on_integers(ws, device, splits, m, n, k):
    """ws_gemm_bf16_splitk: k ranges multiplied by separate blocks and summed
    in range order; exact on small integers, with and without accumulate."""
    import ctypes
    rng = np.random.default_rng(splits * 7 + m + n + k)
    a = _int_bf16(rng, (m, k))
    b = _int_bf16(rng, (n, k))
    want = a.astype(np.float64) @ b.astype(np.float64).T
    ta = torch.tensor(a, device=device).to(torch.bfloat16).contiguous()
    tb = torch.tensor(b, device=device).to(torch.bfloat16).contiguous()
    out = torch.full((m, n), 2.0, device=device)
    work = torch.empty(splits * m * n, dtype=torch.float32, device=device)
    lib = ws.load()
    p = lambda t: ctypes.c_void_p(t.data_ptr())   # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.ws_gemm_bf16_splitk(m, n, k, p(ta), k, p(tb), k, p(out), n, 1, p(work), splits, stream) == 0
    np.testing.assert_array_equal(out.cpu().numpy(), (want + 2.0).astype(np.float32))
    # the wrapper picks the split itself for few output tiles and a long k
    got = ws.gemm(ta, tb, trans_b=True)
    np.testing.assert_array_equal(got.cpu().numpy(), want.astype(np.float32))
    # misaligned rows are rejected
    assert lib.ws_gemm_bf16_splitk(m, n, k, p(ta), k + 1, p(tb), k, p(out), n, 0, p(work), splits, stream) == -1


def test_gemm_accumulate_and_lane_layout(ws, device):
    """Distinct entries per (row, col) expose any lane/row/column mix-up."""
    m, n, k = 64, 96, 8
    a = np.zeros((m, k), np.float32)
    b = np.zeros((k, n), np.float32)
    a[np.arange(m), np.arange(m) % k] = 1.0
    b[np.arange(k)[:, None], np.arange(n)[None, :]] = (np.arange(k)[:, None] * 7 + np.arange(n)[None, :] % 11)
    want = a.astype(np.float64) @ b
    base = torch.full((m, n), 3.0, device=device)
    out = ws.gemm(torch.tensor(a, device=device).to(torch.bfloat16), torch.tensor(b, device=device).to(torch.bfloat16),
                  out=base, accumulate=True)
    np.testing.assert_array_equal(out.cpu().numpy(), (want + 3.0).astype(np.float32))


def test_gemm_random_bf16_within_fp32_accumulation_bound(ws, device):
    rng = np.random.default_rng(5)
    m, n, k = 300, 500, 1024
    a = mlp.bf16_round(rng.normal(size=(m, k)).astype(np.float32))
    b = mlp.bf16_round(rng.normal(size=(k, n)).astype(np.float32))
    got = ws.gemm(torch.tensor(a, dtype=torch.float32, device=device).to(torch.bfloat16),
                  torch.tensor(b, dtype=torch.float32, device=device).to(torch.bfloat16)).cpu().numpy()
    want = a @ b
    bound = 2 * k * np.finfo(np.float32).eps * (np.abs(a) @ np.abs(b))     # gamma_k |A||B|
    assert np.all(np.abs(got - want) <= bound)


@pytest.mark.parametrize("r,c", [(1, 1), (63, 65), (128, 64), (1000, 37)])
def test_transpose_exact(ws, device, r, c):
    x = torch.randn(r, c, device=device).to(torch.bfloat16)
    assert torch.equal(ws.transpose(x), x.t().contiguous())


def test_to_bf16_matches_torch_rounding(ws, device):
    x = torch.tensor(np.random.default_rng(0).normal(size=4097).astype(np.float32) * 1e3, device=device)
    x[0] = float("nan")
    got = ws.to_bf16(x)
    want = x.to(torch.bfloat16)
    assert torch.isnan(got[0].float())
    assert torch.equal(got[1:].float(), want[1:].float())


def _small_net(ws, device, dims, nx, ny, na, seed=0):
    net = ws.WarmStartNet(dims[0], nx, ny, na, hidden=dims[1], layers=len(dims) - 2, device=device, seed=seed)
    rng = np.random.default_rng(seed)
    wts = [mlp.bf16_round(rng.normal(scale=1 / np.sqrt(k), size=(k, n)).astype(np.float32))
           for k, n in zip(dims[:-1], dims[1:])]
    bs = [rng.normal(scale=0.1, size=n).astype(np.float32).astype(np.float64) for n in dims[1:]]
    net.set_params(wts, bs)
    return net, wts, bs


def test_forward_and_gradients_match_oracle(ws, device):
    """BF16 operands, FP32 accumulation vs the FP64 oracle on the same BF16
    weights and inputs; tolerance 3e-2 relative (norm-wise), set by BF16
    rounding of the stored activations and gradients (2^-8 each)."""
    dims, nx, ny, na = [40, 96, 96, 72], 30, 20, 22
    net, wts, bs = _small_net(ws, device, dims, nx, ny, na)
    rng = np.random.default_rng(1)
    mb = 64
    x = mlp.bf16_round(rng.normal(size=(mb, dims[0])).astype(np.float32))
    t = np.concatenate([rng.normal(size=(mb, nx + ny)), (rng.random((mb, na)) < 0.4)], axis=1)
    xb = torch.tensor(x, dtype=torch.float32, device=device).to(torch.bfloat16)
    tb = torch.tensor(t, dtype=torch.float32, device=device)
    out, _ = net._forward(xb, keep=False)
    o_ref, _, _ = mlp.forward(wts, bs, x)
    rel = np.linalg.norm(out.cpu().numpy() - o_ref) / np.linalg.norm(o_ref)
    assert rel < 3e-2, rel
    loss, grads = net.loss_and_grad(xb, tb)
    val, dws, dbs = mlp.loss_and_grad(wts, bs, x, t, nx, ny, na)
    assert abs(loss.item() - val) <= 3e-2 * abs(val)
    for l in range(len(wts)):
        g = grads[l].cpu().numpy()
        assert np.linalg.norm(g - dws[l]) <= 3e-2 * np.linalg.norm(dws[l]), l
        gb = grads[len(wts) + l].cpu().numpy()
        assert np.linalg.norm(gb - dbs[l]) <= 3e-2 * np.linalg.norm(dbs[l]), l


def test_adam_step_matches_oracle(ws, device):
    rng = np.random.default_rng(2)
    n = 1000
    w = rng.normal(size=n).astype(np.float32)
    tw = torch.tensor(w, device=device)
    tm = torch.zeros_like(tw)
    tv = torch.zeros_like(tw)
    wb = torch.empty(n, dtype=torch.bfloat16, device=device)
    wr, mr, vr = w.astype(np.float64), np.zeros(n), np.zeros(n)
    lib = ws.load()
    for step in range(1, 4):
        g = rng.normal(size=n).astype(np.float32)
        tg = torch.tensor(g, device=device)
        assert lib.ws_adam(ws._ptr(tw), ws._ptr(tg), ws._ptr(tm), ws._ptr(tv), ws._ptr(wb), n, 1e-2, 0.9, 0.999,
                           1e-8, step, ws._stream()) == 0
        wr, mr, vr = mlp.adam(wr, g.astype(np.float64), mr, vr, step, 1e-2)
    np.testing.assert_allclose(tw.cpu().numpy(), wr, rtol=1e-5, atol=1e-6)
    assert torch.equal(wb.float(), tw.to(torch.bfloat16).float())


def test_training_reduces_loss(ws, device):
    """A learnable target (a fixed random network) is fitted by Adam."""
    dims, nx, ny, na = [16, 64, 64, 24], 10, 8, 6
    rng = np.random.default_rng(3)
    teacher = [rng.normal(scale=1 / np.sqrt(k), size=(k, n)) for k, n in zip(dims[:-1], dims[1:])]
    tb = [np.zeros(n) for n in dims[1:]]
    feats = rng.normal(size=(2048, dims[0]))
    o, _, _ = mlp.forward(teacher, tb, feats)
    targets = np.concatenate([o[:, :nx + ny], (o[:, nx + ny:] > 0).astype(float)], axis=1)
    net = ws.WarmStartNet(dims[0], nx, ny, na, hidden=64, layers=2, device=device, seed=4)
    hist = net.fit(torch.tensor(feats, dtype=torch.float32, device=device),
                   torch.tensor(targets, dtype=torch.float32, device=device), epochs=30, batch=256, lr=3e-3)
    assert hist[-1] < 0.35 * hist[0], hist


def _snap_reference(x, logits, lower, upper):
    """Projection onto [lower, upper], then predicted at-bound variables moved
    to the nearer finite bound (the rule of the fused output epilogue)."""
    x = torch.minimum(torch.maximum(x, lower), upper)
    flo, fhi = torch.isfinite(lower), torch.isfinite(upper)
    to_lo = flo & (~fhi | ((x - lower) <= (upper - x)))
    at = logits > 0
    return torch.where(at & to_lo, lower, torch.where(at & fhi, upper, x))


def test_fused_inference_equals_the_unfused_forward(ws, device):
    """infer_start (standardisation, bias + GELU epilogues, the interleaved
    output layer with de-standardisation, projection and snapping) equals the
    unfused forward followed by the same host-side steps, bit for bit - with
    free, one-sided and boxed columns."""
    nx, ny, na, neq = 50, 30, 50, 22
    net = ws.WarmStartNet(37, nx, ny, na, hidden=128, layers=2, device=device, seed=5)
    g = torch.Generator(device="cpu").manual_seed(7)
    feats = torch.randn(300, 37, generator=g).to(device)
    targ = torch.randn(300, nx + ny + na, generator=g).to(device) * 3.0 + 1.0
    net.fit_normalization(feats, targ)
    for l in range(len(net.w)):                   # a trained-looking network: larger random weights
        net.w[l].copy_(torch.randn(net.w[l].shape, generator=g).to(device) / np.sqrt(net.w[l].shape[0]))
        net.b[l].copy_(torch.randn(net.b[l].shape, generator=g).to(device) * 0.1)
        ws.to_bf16(net.w[l], net.wb[l])
        ws.transpose(net.wb[l], net.wbt[l])
    net.trained = True
    lower = torch.randn(300, nx, generator=g, dtype=torch.float64).to(device) - 2.0
    upper = lower + torch.rand(300, nx, generator=g, dtype=torch.float64).to(device) * 6.0
    lower[:, :5] = -float("inf"); upper[:, :5] = float("inf")      # free columns
    lower[:, 5:10] = -float("inf")                                 # upper bound only
    upper[:, 10:15] = float("inf")                                 # lower bound only
    x0, y0 = net.infer_start(feats, lower, upper, neq)
    xr, yr, logits = net.predict(feats)
    x_ref = _snap_reference(xr.to(torch.float64), logits, lower, upper)
    y_ref = yr.to(torch.float64)
    y_ref[:, neq:] = torch.clamp(y_ref[:, neq:], max=0.0)
    assert torch.isfinite(x0).all() and torch.isfinite(y0).all()
    assert torch.equal(x0, x_ref)
    assert torch.equal(y0, y_ref)
    # predicted at-bound variables sit exactly on a finite bound; free ones are never snapped
    at = logits > 0
    on_bound = (x0 == lower) | (x0 == upper)
    assert bool(on_bound[:, 5:][at[:, 5:]].all())
    assert bool((x0[:, :5] == xr[:, :5].to(torch.float64)).all())


def test_checkpoint_round_trip_restores_predictions_and_adam(ws, device, tmp_path):
    """save / load: the restored network predicts the same starts bit for
    bit, and one more Adam step from the restored state equals the step the
    original takes (moments and step count restored)."""
    rng = np.random.default_rng(11)
    feats = torch.tensor(rng.normal(size=(512, 20)), dtype=torch.float32, device=device)
    targ = torch.tensor(np.concatenate([rng.normal(size=(512, 14)), (rng.random((512, 8)) > 0.5)], axis=1),
                        dtype=torch.float32, device=device)
    a = ws.WarmStartNet(20, 8, 6, 8, hidden=64, layers=2, device=device, seed=1)
    a.fit(feats, targ, epochs=3, batch=128, lr=1e-3)
    path = str(tmp_path / "net.safetensors")
    a.save(path)
    b = ws.WarmStartNet(20, 8, 6, 8, hidden=64, layers=2, device=device, seed=99)
    b.load(path)
    lo = torch.full((512, 8), -5.0, dtype=torch.float64, device=device)
    hi = torch.full((512, 8), 5.0, dtype=torch.float64, device=device)
    xa, ya = a.infer_start(feats, lo, hi, 3)
    xb, yb = b.infer_start(feats, lo, hi, 3)
    assert torch.equal(xa, xb) and torch.equal(ya, yb)
    xs, ts = a._inputs(feats[:128]), a._targets(targ[:128])
    _, ga = a.loss_and_grad(xs, ts)
    a.adam_step(ga, 1e-3)
    _, gb = b.loss_and_grad(xs, ts)
    b.adam_step(gb, 1e-3)
    for pa, pb in zip(a.params, b.params):
        assert torch.equal(pa, pb)
    c = ws.WarmStartNet(20, 8, 6, 9, hidden=64, layers=2, device=device)
    with pytest.raises(ValueError):
        c.load(path)
