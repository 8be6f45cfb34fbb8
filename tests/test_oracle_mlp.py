"""Pins of the FP64 warm-start network oracle (oracle/mlp.py): central finite
differences of the loss, torch CPU autograd of the same network, torch's Adam,
and BF16 rounding known answers."""

import numpy as np
import pytest
import torch

from oracle import mlp


def _net(rng, dims):
    ws = [rng.normal(scale=1.0 / np.sqrt(k), size=(k, n)) for k, n in zip(dims[:-1], dims[1:])]
    bs = [rng.normal(scale=0.1, size=n) for n in dims[1:]]
    return ws, bs


def _data(rng, m, dims, nx, ny, na):
    x = rng.normal(size=(m, dims[0]))
    t = np.concatenate([rng.normal(size=(m, nx + ny)), (rng.random((m, na)) < 0.3).astype(float)], axis=1)
    return x, t


def test_gelu_known_values():
    assert mlp.gelu(np.array(0.0)) == 0.0
    np.testing.assert_allclose(mlp.gelu(np.array([10.0, -10.0])), [10.0, 0.0], atol=1e-12)
    # gelu(1) = 0.5 (1 + tanh(sqrt(2/pi) 1.044715))
    np.testing.assert_allclose(mlp.gelu(np.array(1.0)), 0.8411919906082768, rtol=1e-15)
    z = np.linspace(-4, 4, 41)
    fd = (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6
    np.testing.assert_allclose(mlp.gelu_grad(z), fd, atol=1e-8)


def test_backward_matches_central_finite_differences():
    rng = np.random.default_rng(0)
    dims, nx, ny, na = [5, 7, 6, 9], 3, 2, 4
    ws, bs = _net(rng, dims)
    x, t = _data(rng, 4, dims, nx, ny, na)
    _, dws, dbs = mlp.loss_and_grad(ws, bs, x, t, nx, ny, na, 1.0, 0.5, 2.0)
    h = 1e-6
    for l in range(len(ws)):
        for (i, j) in [(0, 0), (ws[l].shape[0] - 1, ws[l].shape[1] - 1), (1, 2)]:
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[l][i, j] += h
            wm[l][i, j] -= h
            fd = (mlp.loss_and_grad(wp, bs, x, t, nx, ny, na, 1.0, 0.5, 2.0)[0]
                  - mlp.loss_and_grad(wm, bs, x, t, nx, ny, na, 1.0, 0.5, 2.0)[0]) / (2 * h)
            assert abs(fd - dws[l][i, j]) <= 1e-7 * max(1.0, abs(fd)), (l, i, j)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1] += h
        bm[l][1] -= h
        fd = (mlp.loss_and_grad(ws, bp, x, t, nx, ny, na, 1.0, 0.5, 2.0)[0]
              - mlp.loss_and_grad(ws, bm, x, t, nx, ny, na, 1.0, 0.5, 2.0)[0]) / (2 * h)
        assert abs(fd - dbs[l][1]) <= 1e-7 * max(1.0, abs(fd))


def test_matches_torch_autograd():
    rng = np.random.default_rng(1)
    dims, nx, ny, na = [6, 8, 8, 11], 4, 3, 4
    ws, bs = _net(rng, dims)
    x, t = _data(rng, 5, dims, nx, ny, na)
    val, dws, dbs = mlp.loss_and_grad(ws, bs, x, t, nx, ny, na, 1.0, 2.0, 0.5)
    tw = [torch.tensor(w, requires_grad=True) for w in ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    h = torch.tensor(x)
    for l in range(len(tw)):
        h = h @ tw[l] + tb[l]
        if l < len(tw) - 1:
            h = torch.nn.functional.gelu(h, approximate="tanh")
    tt = torch.tensor(t)
    m = x.shape[0]
    loss = (((h[:, :nx] - tt[:, :nx]) ** 2).sum() / (m * nx)
            + 2.0 * ((h[:, nx:nx + ny] - tt[:, nx:nx + ny]) ** 2).sum() / (m * ny)
            + 0.5 * torch.nn.functional.binary_cross_entropy_with_logits(h[:, nx + ny:], tt[:, nx + ny:],
                                                                         reduction="sum") / (m * na))
    loss.backward()
    # BCE-with-logits = softplus(o) - t o
    np.testing.assert_allclose(val, loss.item(), rtol=1e-12)
    for l in range(len(ws)):
        np.testing.assert_allclose(dws[l], tw[l].grad.numpy(), rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(dbs[l], tb[l].grad.numpy(), rtol=1e-10, atol=1e-13)


def test_adam_matches_torch():
    rng = np.random.default_rng(2)
    w = rng.normal(size=10)
    tw = torch.tensor(w.copy(), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for step in range(1, 6):
        g = rng.normal(size=10)
        w, m, v = mlp.adam(w, g, m, v, step, 1e-2)
        opt.zero_grad()
        tw.grad = torch.tensor(g)
        opt.step()
    np.testing.assert_allclose(w, tw.detach().numpy(), rtol=1e-12)


def test_bf16_round_known_answers():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9 * 1.5, -2.5, 3.0e38, 1e-40], np.float32)
    got = mlp.bf16_round(vals)
    want = torch.tensor(vals).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(got, want)
    assert got[1] == 1.0                       # tie to even
    assert got[2] == 1.0 + 4 * 2 ** -8         # tie to even (upwards)
    assert np.isnan(mlp.bf16_round(np.array([np.nan], np.float32)))[0]


@pytest.mark.parametrize("mistake", ["transpose", "sign"])
def test_planted_backward_mistakes_fail_the_finite_difference_pin(mistake):
    """A plausible mistake in the backward is caught by the finite-difference check."""
    rng = np.random.default_rng(3)
    dims, nx, ny, na = [4, 5, 3], 1, 1, 1
    ws, bs = _net(rng, dims)
    x, t = _data(rng, 3, dims, nx, ny, na)
    val, dws, _ = mlp.loss_and_grad(ws, bs, x, t, nx, ny, na)
    if mistake == "transpose":
        bad = dws[0][::-1]                      # rows in the wrong order
    else:
        bad = -dws[0]
    h = 1e-6
    wp = [w.copy() for w in ws]
    wm = [w.copy() for w in ws]
    wp[0][0, 0] += h
    wm[0][0, 0] -= h
    fd = (mlp.loss_and_grad(wp, bs, x, t, nx, ny, na)[0] - mlp.loss_and_grad(wm, bs, x, t, nx, ny, na)[0]) / (2 * h)
    assert abs(fd - dws[0][0, 0]) <= 1e-7 * max(1, abs(fd))
    assert abs(fd - bad[0, 0]) > 1e-6 * max(1, abs(fd))
