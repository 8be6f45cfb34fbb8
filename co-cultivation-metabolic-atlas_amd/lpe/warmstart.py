"""Learned warm start for the dFBA step: a multilayer perceptron on MI355X.

The network maps an environment's LP parameters (log medium concentrations,
species biomass, Michaelis-Menten uptake bounds) to an initial primal x0, an
initial dual y0 and one logit per column for "this variable sits at a bound".
It trains on solutions the engine itself certified.  Every GEMM, activation,
loss gradient and Adam update runs in libwarmstart.so (hand-written HIP: BF16
MFMA GEMMs with FP32 accumulation, FP32 master weights); torch only allocates
device memory and does the input/output standardisation.

The reference has no MLP warm start; its learned warm starts are GRU/GNN
research modules (src/temporal_lp_model.py:58-129,
src/supervised_graph_lp.py:13-91).  Cold start stays the fallback: a predicted
start is only a starting point, the engine's certificate decides acceptance,
and ``LearnedWarmStart`` returns no start before the network is trained.
"""

from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional, Sequence

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwarmstart.so")
EXPORTS = ("ws_gemm_bf16", "ws_gemm_bf16_splitk", "ws_bias_act", "ws_act_backward", "ws_loss", "ws_adam", "ws_transpose_bf16",
           "ws_to_bf16", "ws_gemm_bf16_act", "ws_gemm_bf16_start", "ws_standardize", "ws_version")

_lib = None


def load():
    """Load libwarmstart.so (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P, I, F = C.c_void_p, C.c_int, C.c_float
    lib.ws_gemm_bf16.argtypes = [I, I, I, P, I, I, P, I, I, P, I, I, P]
    lib.ws_gemm_bf16_splitk.argtypes = [I, I, I, P, I, P, I, P, I, I, P, I, P]
    lib.ws_bias_act.argtypes = [P, P, P, I, I, I, P]
    lib.ws_act_backward.argtypes = [P, P, P, P, I, I, I, P]
    lib.ws_loss.argtypes = [P, P, P, P, I, I, I, I, F, F, F, P]
    lib.ws_adam.argtypes = [P, P, P, P, P, C.c_int64, F, F, F, F, I, P]
    lib.ws_transpose_bf16.argtypes = [P, P, I, I, P]
    lib.ws_to_bf16.argtypes = [P, P, C.c_int64, P]
    lib.ws_gemm_bf16_act.argtypes = [I, I, I, P, I, P, I, P, P, I, P]
    lib.ws_gemm_bf16_start.argtypes = [I, I, I, I, P, I, P, I, P, P, P, P, P, I, I, P, P, P]
    lib.ws_standardize.argtypes = [P, P, P, P, I, I, I, P]
    lib.ws_version.restype = C.c_char_p
    for name in EXPORTS[:-1]:
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def gemm(a: torch.Tensor, b: torch.Tensor, *, trans_a: bool = False, trans_b: bool = False,
         out: Optional[torch.Tensor] = None, accumulate: bool = False) -> torch.Tensor:
    """out = op(a) @ op(b) with BF16 operands and FP32 result (ws_gemm_bf16)."""
    lib = load()
    if a.dtype != torch.bfloat16 or b.dtype != torch.bfloat16:
        raise ValueError("gemm operands must be bfloat16")
    if not (a.is_contiguous() and b.is_contiguous()):
        raise ValueError("gemm operands must be contiguous")
    m, k = (a.shape[1], a.shape[0]) if trans_a else (a.shape[0], a.shape[1])
    kb, n = (b.shape[1], b.shape[0]) if trans_b else (b.shape[0], b.shape[1])
    if k != kb:
        raise ValueError(f"inner dimensions differ: {k} vs {kb}")
    if out is None:
        out = torch.empty((m, n), dtype=torch.float32, device=a.device)
    elif out.shape != (m, n) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 [M, N] tensor")
    splits = _splits(m, n, k) if (not trans_a and trans_b) else 1
    if splits > 1:
        # few output tiles and a long k: the k range is split over more blocks
        # (partial products summed in a fixed order)
        work = _workspace(splits * m * n, a.device)
        _ok(lib.ws_gemm_bf16_splitk(m, n, k, _ptr(a), a.shape[1], _ptr(b), b.shape[1], _ptr(out), n,
                                    int(accumulate), _ptr(work), splits, _stream()), "ws_gemm_bf16_splitk")
        return out
    _ok(lib.ws_gemm_bf16(m, n, k, _ptr(a), a.shape[1], int(trans_a), _ptr(b), b.shape[1], int(trans_b),
                         _ptr(out), n, int(accumulate), _stream()), "ws_gemm_bf16")
    return out


TILE = 128          # output tile of the GEMM kernels
TARGET_BLOCKS = 256  # one block per compute unit


def _splits(m: int, n: int, k: int) -> int:
    """K splits for a product whose output tiles alone cannot fill the chip."""
    tiles = (-(-m // TILE)) * (-(-n // TILE))
    if tiles >= TARGET_BLOCKS or k < 2048:
        return 1
    return int(max(1, min(8, TARGET_BLOCKS // tiles, k // 1024)))


_work = {}


def _workspace(numel: int, device) -> torch.Tensor:
    """FP32 scratch of the split-K products, kept per device and grown on demand."""
    w = _work.get(device)
    if w is None or w.numel() < numel:
        w = torch.empty(numel, dtype=torch.float32, device=device)
        _work[device] = w
    return w


def transpose(x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """BF16 [R, C] -> [C, R] (ws_transpose_bf16)."""
    lib = load()
    if x.dtype != torch.bfloat16 or x.dim() != 2 or not x.is_contiguous():
        raise ValueError("transpose takes a contiguous 2-D bfloat16 tensor")
    r, c = x.shape
    out = torch.empty((c, r), dtype=torch.bfloat16, device=x.device) if out is None else out
    _ok(lib.ws_transpose_bf16(_ptr(x), _ptr(out), r, c, _stream()), "ws_transpose_bf16")
    return out


def to_bf16(x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    lib = load()
    x = x.contiguous()
    if x.dtype != torch.float32:
        raise ValueError("to_bf16 takes float32")
    out = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device) if out is None else out
    _ok(lib.ws_to_bf16(_ptr(x), _ptr(out), x.numel(), _stream()), "ws_to_bf16")
    return out


class WarmStartNet:
    """MLP n_in -> hidden x layers (GELU) -> nx + ny + na, trained with Adam.

    Weights are stored [in][out] (FP32 master, BF16 copy for the GEMMs).
    Inputs and regression targets are standardised per column with statistics
    fixed at the first call of ``fit_normalization``.
    """

    def __init__(self, n_in: int, nx: int, ny: int, na: int, *, hidden: int = 1024, layers: int = 3,
                 device=None, seed: int = 0, loss_weights=(1.0, 1.0, 1.0)):
        load()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.n_in, self.nx, self.ny, self.na = int(n_in), int(nx), int(ny), int(na)
        # widths padded to multiples of 8 so every GEMM operand row is 16-byte aligned:
        # zero input columns, and extra bound logits trained towards label 0
        self.n_in_p = -(-self.n_in // 8) * 8
        self.na_p = self.na + (-(self.nx + self.ny + self.na)) % 8
        self.n_out = self.nx + self.ny + self.na_p
        self.dims = [self.n_in_p] + [int(hidden)] * int(layers) + [self.n_out]
        self.loss_weights = tuple(float(w) for w in loss_weights)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w, self.b = [], []
        for k, n in zip(self.dims[:-1], self.dims[1:]):
            bound = 1.0 / math.sqrt(k)
            self.w.append(((torch.rand((k, n), generator=g, dtype=torch.float32) * 2 - 1) * bound)
                          .to(self.device))
            self.b.append(torch.zeros(n, dtype=torch.float32, device=self.device))
        # BF16 copies in both layouts: every GEMM then reads two k-contiguous operands
        self.wb = [to_bf16(w) for w in self.w]              # [in][out]
        self.wbt = [transpose(w) for w in self.wb]          # [out][in]
        self.params = self.w + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.step_count = 0
        self.in_mean = self.in_std = self.out_mean = self.out_std = None
        self.trained = False
        self._inf = None

    @property
    def n_params(self) -> int:
        return sum(p.numel() for p in self.params)

    def flops_per_sample(self) -> int:
        return 2 * sum(k * n for k, n in zip(self.dims[:-1], self.dims[1:]))

    def set_params(self, weights: Sequence, biases: Sequence) -> None:
        """Load weights [in][out] and biases given at the unpadded widths."""
        for l, (w, b) in enumerate(zip(weights, biases)):
            w = torch.as_tensor(np.asarray(w), dtype=torch.float32)
            b = torch.as_tensor(np.asarray(b), dtype=torch.float32)
            self.w[l].zero_()
            self.b[l].zero_()
            self.w[l][:w.shape[0], :w.shape[1]].copy_(w)
            self.b[l][:b.shape[0]].copy_(b)
            to_bf16(self.w[l], self.wb[l])
            transpose(self.wb[l], self.wbt[l])

    # ------------------------------------------------------------ statistics
    def fit_normalization(self, features: torch.Tensor, targets: torch.Tensor) -> None:
        f = features.to(torch.float64)
        self.in_mean = f.mean(0).to(torch.float32)
        self.in_std = f.std(0).clamp_min(1e-6).to(torch.float32)
        t = targets[:, :self.nx + self.ny].to(torch.float64)
        self.out_mean = t.mean(0).to(torch.float32)
        self.out_std = t.std(0).clamp_min(1e-6).to(torch.float32)

    def _inputs(self, features):
        x = torch.zeros((features.shape[0], self.n_in_p), dtype=torch.float32, device=self.device)
        x[:, :self.n_in] = (features.to(torch.float32) - self.in_mean) / self.in_std
        return to_bf16(x)

    def _targets(self, targets):
        t = torch.zeros((targets.shape[0], self.n_out), dtype=torch.float32, device=self.device)
        k = self.nx + self.ny
        t[:, :k] = (targets[:, :k].to(torch.float32) - self.out_mean) / self.out_std
        t[:, k:k + self.na] = targets[:, k:k + self.na].to(torch.float32)
        return t

    # ------------------------------------------------------------- forward
    def _forward(self, xb: torch.Tensor, keep: bool):
        lib = load()
        s = _stream()
        mb = xb.shape[0]
        h = xb
        zs, hs = [], [xb]
        last = len(self.w) - 1
        for l in range(len(self.w)):
            n = self.dims[l + 1]
            z = gemm(h, self.wbt[l], trans_b=True)
            hn = None if l == last else torch.empty((mb, n), dtype=torch.bfloat16, device=self.device)
            _ok(lib.ws_bias_act(_ptr(z), _ptr(self.b[l]), _ptr(hn), mb, n, 0 if l == last else 1, s),
                "ws_bias_act")
            zs.append(z)
            if hn is not None:
                hs.append(hn)
                h = hn
        return zs[-1], (zs, hs) if keep else None

    def predict_raw(self, features: torch.Tensor) -> torch.Tensor:
        """Network output [M, nx+ny+na] in standardised units (logits last)."""
        o, _ = self._forward(self._inputs(features), keep=False)
        return o

    def predict(self, features: torch.Tensor):
        """(x [M,nx], y [M,ny], bound logits [M,na]) in original units."""
        o = self.predict_raw(features)
        k = self.nx + self.ny
        reg = o[:, :k] * self.out_std + self.out_mean
        return reg[:, :self.nx], reg[:, self.nx:k], o[:, k:k + self.na]

    # ------------------------------------------------------ checkpoints
    def save(self, path: str) -> None:
        """Weights, biases, Adam moments and step, and the input / output
        standardisation in one safetensors file (nothing executable)."""
        from safetensors.torch import save_file
        if self.in_mean is None:
            raise RuntimeError("the network has no normalisation statistics yet (fit or fit_normalization first)")
        t = {}
        for i, (p, m, v) in enumerate(zip(self.params, self.m, self.v)):
            t[f"param.{i}"], t[f"adam_m.{i}"], t[f"adam_v.{i}"] = p, m, v
        t.update(in_mean=self.in_mean, in_std=self.in_std, out_mean=self.out_mean, out_std=self.out_std)
        meta = {"dims": ",".join(str(d) for d in self.dims), "nx": str(self.nx), "ny": str(self.ny),
                "na": str(self.na), "n_in": str(self.n_in), "step_count": str(self.step_count),
                "trained": str(int(self.trained))}
        save_file({k: v.detach().contiguous().cpu() for k, v in t.items()}, path, metadata=meta)

    def load(self, path: str) -> None:
        """Restore a checkpoint written by ``save`` for a network of the same
        widths (the BF16 copies are rebuilt from the FP32 master weights)."""
        from safetensors import safe_open
        with safe_open(path, framework="pt", device="cpu") as f:
            meta = f.metadata()
            if meta.get("dims") != ",".join(str(d) for d in self.dims) or int(meta["nx"]) != self.nx \
                    or int(meta["ny"]) != self.ny or int(meta["na"]) != self.na or int(meta["n_in"]) != self.n_in:
                raise ValueError("the checkpoint's network widths differ from this network's")
            for i in range(len(self.params)):
                self.params[i].copy_(f.get_tensor(f"param.{i}"))
                self.m[i].copy_(f.get_tensor(f"adam_m.{i}"))
                self.v[i].copy_(f.get_tensor(f"adam_v.{i}"))
            self.in_mean, self.in_std, self.out_mean, self.out_std = (
                f.get_tensor(k).to(self.device) for k in ("in_mean", "in_std", "out_mean", "out_std"))
            self.step_count = int(meta["step_count"])
            self.trained = bool(int(meta["trained"]))
        for l in range(len(self.w)):
            to_bf16(self.w[l], self.wb[l])
            transpose(self.wb[l], self.wbt[l])
        self._inf = None

    # ------------------------------------------------- fused inference
    def _inference_layout(self):
        """Output layer for ``infer_start``: rows of the [out][in] BF16 weights
        interleaved as (x_j, logit_j) for j < nx, then the y rows, with the
        matching bias and the de-standardisation (scale, shift) per row."""
        if getattr(self, "_inf_step", None) == self.step_count and self._inf is not None:
            return self._inf
        nx, ny = self.nx, self.ny
        if self.na < nx:
            raise ValueError("the fused output layer needs one bound logit per x column")
        dev = self.device
        jx = torch.arange(nx, device=dev)
        order = torch.empty(2 * nx + ny, dtype=torch.long, device=dev)
        order[0:2 * nx:2] = jx
        order[1:2 * nx:2] = nx + ny + jx
        order[2 * nx:] = nx + torch.arange(ny, device=dev)
        w_out = self.wbt[-1][order].contiguous()
        bias = self.b[-1][order].contiguous()
        scale = torch.ones(2 * nx + ny, dtype=torch.float32, device=dev)
        shift = torch.zeros(2 * nx + ny, dtype=torch.float32, device=dev)
        scale[0:2 * nx:2] = self.out_std[:nx]
        shift[0:2 * nx:2] = self.out_mean[:nx]
        scale[2 * nx:] = self.out_std[nx:nx + ny]
        shift[2 * nx:] = self.out_mean[nx:nx + ny]
        self._inf = (w_out, bias, scale, shift)
        self._inf_step = self.step_count
        return self._inf

    def infer_start(self, features: torch.Tensor, lower: torch.Tensor, upper: torch.Tensor, neq: int,
                    snap: bool = True):
        """Warm start (x0 [M, nx], y0 [M, ny], FP64) straight from the network:
        input standardisation, three GEMMs with fused bias + GELU writing BF16,
        and the output GEMM whose epilogue de-standardises, projects x onto
        [lower, upper], snaps predicted at-bound variables and clips the
        inequality duals (ws_standardize, ws_gemm_bf16_act, ws_gemm_bf16_start)."""
        lib = load()
        s = _stream()
        mb = features.shape[0]
        w_out, bias, scale, shift = self._inference_layout()
        f = features.to(torch.float32).contiguous()
        xb = torch.empty((mb, self.n_in_p), dtype=torch.bfloat16, device=self.device)
        _ok(lib.ws_standardize(_ptr(f), _ptr(self.in_mean), _ptr(self.in_std), _ptr(xb), mb, self.n_in,
                               self.n_in_p, s), "ws_standardize")
        h = xb
        for l in range(len(self.w) - 1):
            n, k = self.dims[l + 1], self.dims[l]
            hn = torch.empty((mb, n), dtype=torch.bfloat16, device=self.device)
            _ok(lib.ws_gemm_bf16_act(mb, n, k, _ptr(h), k, _ptr(self.wbt[l]), k, _ptr(self.b[l]), _ptr(hn), n, s),
                "ws_gemm_bf16_act")
            h = hn
        lo = lower.to(torch.float64).contiguous()
        up = upper.to(torch.float64).contiguous()
        x0 = torch.empty((mb, self.nx), dtype=torch.float64, device=self.device)
        y0 = torch.empty((mb, self.ny), dtype=torch.float64, device=self.device)
        k = self.dims[-2]
        _ok(lib.ws_gemm_bf16_start(mb, self.nx, self.ny, k, _ptr(h), k, _ptr(w_out), k, _ptr(bias), _ptr(scale),
                                   _ptr(shift), _ptr(lo), _ptr(up), int(neq), 1 if snap else 0, _ptr(x0), _ptr(y0),
                                   s), "ws_gemm_bf16_start")
        return x0, y0

    # ------------------------------------------------------------ training
    def loss_and_grad(self, xb: torch.Tensor, tb: torch.Tensor):
        """Forward + backward on a standardised BF16 batch; returns (loss,
        gradients) with gradients listed like ``params``."""
        lib = load()
        s = _stream()
        mb = xb.shape[0]
        o, (zs, hs) = self._forward(xb, keep=True)
        d_out = torch.empty_like(o)
        row_loss = torch.empty(mb, dtype=torch.float32, device=self.device)
        wx, wy, wa = self.loss_weights
        _ok(lib.ws_loss(_ptr(o), _ptr(tb), _ptr(d_out), _ptr(row_loss), mb, self.nx, self.ny, self.na_p,
                        wx, wy, wa, s), "ws_loss")
        gw = [None] * len(self.w)
        gb = [torch.zeros_like(b) for b in self.b]
        dh = d_out
        for l in range(len(self.w) - 1, -1, -1):
            n = self.dims[l + 1]
            dz = torch.empty((mb, n), dtype=torch.bfloat16, device=self.device)
            _ok(lib.ws_act_backward(_ptr(dh), _ptr(zs[l]), _ptr(dz), _ptr(gb[l]), mb, n,
                                    0 if l == len(self.w) - 1 else 1, s), "ws_act_backward")
            # dW = H^T dZ from the two transposed activations (k = batch contiguous)
            gw[l] = gemm(transpose(hs[l]), transpose(dz), trans_b=True)
            if l > 0:
                dh = gemm(dz, self.wb[l], trans_b=True)
        return row_loss.sum(), gw + gb

    def adam_step(self, grads, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        lib = load()
        s = _stream()
        self.step_count += 1
        nw = len(self.w)
        for i, (p, g, m, v) in enumerate(zip(self.params, grads, self.m, self.v)):
            wb = self.wb[i] if i < nw else torch.empty(p.shape, dtype=torch.bfloat16, device=self.device)
            _ok(lib.ws_adam(_ptr(p), _ptr(g), _ptr(m), _ptr(v), _ptr(wb), p.numel(), lr, beta1, beta2, eps,
                            self.step_count, s), "ws_adam")
            if i < nw:
                transpose(wb, self.wbt[i])

    def fit(self, features: torch.Tensor, targets: torch.Tensor, *, epochs: int = 20, batch: int = 1024,
            lr: float = 1e-3, lr_final: float = 1e-4, seed: int = 0, log=None) -> list:
        """Mini-batch Adam over device-resident data; cosine learning-rate decay."""
        if self.in_mean is None:
            self.fit_normalization(features, targets)
        x_all = self._inputs(features)
        t_all = self._targets(targets)
        n = x_all.shape[0]
        batch = min(batch, n)
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        gen = torch.Generator(device=self.device).manual_seed(seed)
        history = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device=self.device)
            acc = torch.zeros((), dtype=torch.float32, device=self.device)
            for i in range(steps_per_epoch):
                idx = perm[i * batch:(i + 1) * batch]
                loss, grads = self.loss_and_grad(x_all[idx].contiguous(), t_all[idx].contiguous())
                rate = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * k / max(1, total - 1)))
                self.adam_step(grads, rate)
                acc += loss
                k += 1
            history.append(float(acc) / steps_per_epoch)
            if log:
                log(f"epoch {ep}: loss {history[-1]:.5f}")
        self.trained = True
        return history


# --------------------------------------------------------------- dFBA step
def consortium_features(model, conc: torch.Tensor, lower: torch.Tensor, biomass: torch.Tensor) -> torch.Tensor:
    """[B, K + S*K] network input: log medium concentrations and the exchange
    bounds of every (species, metabolite) term scaled by the species' biomass,
    X_s * lb_(s,i) - the bound of that term in the medium row, in mmol/L/h.
    Biomass enters the LP only through those rows' coefficients and is
    recoverable from these inputs (lb depends on C alone), so 60 + 180 = 240
    values determine the consortium LP."""
    s_idx, _ = model.terms()
    mc = torch.as_tensor(model.term_cols, device=lower.device, dtype=torch.long)
    si = torch.as_tensor(s_idx, device=lower.device, dtype=torch.long)
    return torch.cat([torch.log(conc + 1e-6), biomass[:, si] * lower[:, mc]], dim=1).to(torch.float32)


def bound_labels(x: torch.Tensor, lower: torch.Tensor, upper: torch.Tensor, tol: float = 1e-6) -> torch.Tensor:
    """1 where x sits at a finite bound (within tol * max(1, width)), else 0."""
    width = torch.clamp(torch.nan_to_num(upper - lower, posinf=1.0), min=1.0)
    at = (torch.isfinite(lower) & ((x - lower) <= tol * width)) | (torch.isfinite(upper) & ((upper - x) <= tol * width))
    return at.to(torch.float32)


def start_score(metrics: torch.Tensor) -> torch.Tensor:
    """Distance of a candidate start from optimality, per LP, from its FP64
    certificate [B, 8]: the largest of the primal residual, the dual violation
    and the relative complementarity gap, each over its gate (1e-5, 1e-7,
    1e-7); non-finite candidates score +inf."""
    primal, dual, gap, finite = metrics[:, 0] / 1e-5, metrics[:, 1] / 1e-7, metrics[:, 2] / 1e-7, metrics[:, 5]
    sc = torch.maximum(primal, torch.maximum(dual, gap))
    return torch.where((finite > 0.5) & torch.isfinite(sc), sc, torch.full_like(sc, float("inf")))


class LearnedWarmStart:
    """Warm-start callable for ``ConsortiumRollout``.

    mode "hybrid" (default): an LP starts from the previous step's certified
    solution when that solution, scored by ``start_score`` of its FP64
    certificate on the new LP (``engine.certify``), is within ``keep_score``
    times the certificate gates, and from the network's prediction otherwise
    (or when it has no previous solution).  The network is only evaluated
    when some LP needs it.  The rule and the default of 100 are measured
    (1024 consortium environments): the previous solution scores ~5e3
    (median) one step after the start of a rollout, when biomass has grown by
    tens of percent, and needs 352 PDHG iterations and 3 polish rounds where
    the network's start needs 64 and 1; from the second step on it scores
    below ~10 and needs 32 iterations (0 once the rollout settles) where the
    network's start needs 64.  The network's own score (1e4-3e4, a larger
    residual but the right active set) does not rank its starts, so it is
    not used for the choice.
    mode "predict": the network's prediction for every LP.
    Before training it passes the previous solution through (None on the
    first step: a cold start, the method without the network).  The last
    call's choice is in ``last_choice`` (LPs started from the network)."""

    wants_engine = True

    def __init__(self, model, *, hidden: int = 1024, layers: int = 3, device=None, seed: int = 0,
                 snap: bool = True, mode: str = "hybrid", keep_score: float = 100.0):
        if mode not in ("hybrid", "predict"):
            raise ValueError("mode must be 'hybrid' or 'predict'")
        self.model = model
        m, n = model.shape
        self.n_features = model.n_medium + len(model.term_cols)
        self.net = WarmStartNet(self.n_features, n, m, n, hidden=hidden, layers=layers, device=device, seed=seed)
        self.snap = snap
        self.mode = mode
        self.keep_score = float(keep_score)
        self.features, self.targets = [], []
        self.last_choice = None

    def add_solutions(self, conc, lower, upper, biomass, x, y, certified) -> int:
        """Keep the certified LPs of one solved step as training samples."""
        ok = torch.as_tensor(np.asarray(certified), device=x.device)
        if not bool(ok.any()):
            return 0
        f = consortium_features(self.model, conc, lower, biomass)[ok]
        t = torch.cat([x[ok].to(torch.float32), y[ok].to(torch.float32),
                       bound_labels(x[ok], lower[ok], upper[ok])], dim=1)
        self.features.append(f)
        self.targets.append(t)
        return int(ok.sum())

    @property
    def n_samples(self) -> int:
        return sum(int(f.shape[0]) for f in self.features)

    def train(self, **kwargs) -> list:
        f = torch.cat(self.features)
        t = torch.cat(self.targets)
        self.features, self.targets = [f], [t]
        hist = self.net.fit(f, t, **kwargs)
        # load the scoring kernels now: their first use costs ~0.1 s, which
        # would otherwise land on the first step that scores previous solutions
        start_score(torch.zeros((1, 8), dtype=torch.float64, device=self.net.device))
        torch.cuda.synchronize(self.net.device)
        return hist

    def predict(self, conc, lower, biomass):
        """(x0, y0) of the network for every LP (fused inference)."""
        md = self.model
        upper = torch.as_tensor(md.upper, device=lower.device).expand_as(lower)
        return self.net.infer_start(consortium_features(md, conc, lower, biomass), lower, upper, md.neq,
                                    snap=self.snap)

    def __call__(self, conc, lower, biomass, x_prev=None, y_prev=None, engine=None):
        if not self.net.trained:
            self.last_choice = None
            return x_prev, y_prev
        B = conc.shape[0]
        if self.mode == "predict" or x_prev is None or engine is None or not hasattr(engine, "certify"):
            self.last_choice = B
            return self.predict(conc, lower, biomass)
        use = start_score(engine.certify(x_prev, y_prev)) > self.keep_score
        self.last_choice = int(use.sum())
        if self.last_choice == 0:
            return x_prev, y_prev
        x_m, y_m = self.predict(conc, lower, biomass)
        return (torch.where(use[:, None], x_m, x_prev).contiguous(),
                torch.where(use[:, None], y_m, y_prev).contiguous())
