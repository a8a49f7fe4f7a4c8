"""Learned nonlinear column step on gfx950 (kernels: csrc/learned.h, C ABI: ash_net_*).

The reference has no learned component.  A 4-layer 1-D conv net over t
(kernel 9, 32 channels, bf16 weights and activations on the matrix cores,
float32 accumulation) stands in for the KT update over one dZ (ASS:833-839)
on smooth columns; plain KT stays the default, the fallback wherever a
column's max|dp/dt| exceeds a threshold, and the yardstick for the net's error.
Training pairs (column before, column after one KT step) come from the plain
gfx950 KT kernel itself, and the network is
trained on the GPU by the library's own forward, backward and Adam kernels.

Scaling: with A = max|u| of a column the net sees x = u / A (features x, x^2)
and predicts y = du / (c A^2), c = dZ N / dT, N = beta / (2 c0^3 rho0): the KT
flux is homogeneous of degree two in u, so y is close to amplitude-free.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, current_stream

C, KW, NLAY = 32, 9, 4
W_SIZE = C * C * KW
LAYER_SIZE = W_SIZE + C
PARAM_COUNT = NLAY * LAYER_SIZE


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial parameters; held-at-zero entries are zero (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAM_COUNT, dtype=np.float32)
    fan_in = [2 * KW, C * KW, C * KW, C * KW]
    for j in range(NLAY):
        W = p[j * LAYER_SIZE:j * LAYER_SIZE + W_SIZE].reshape(C, C, KW)
        W[...] = rng.normal(0.0, math.sqrt(2.0 / fan_in[j]), size=W.shape).astype(np.float32)
        if j == 0:
            W[:, 2:, :] = 0.0
        if j == NLAY - 1:
            W[1:, :, :] = 0.0
            W *= 0.1
    return p


class LearnedStep:
    """Owns one ``ash_net``: parameters, Adam state and training workspace."""

    def __init__(self, nT: int, max_cols: int = 256, params: Optional[np.ndarray] = None, seed: int = 0):
        self.lib = _lib.load()
        if int(self.lib.ash_net_param_count()) != PARAM_COUNT:
            raise RuntimeError("parameter layout of libaship.so differs from this module")
        self.nT, self.max_cols = int(nT), int(max_cols)
        h = ctypes.c_void_p()
        check(self.lib.ash_net_create(self.nT, self.max_cols, ctypes.byref(h)), "ash_net_create")
        self.h = h
        self.dev = torch.device("cuda")
        self.grads = torch.zeros(PARAM_COUNT, dtype=torch.float32, device=self.dev)
        self._loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self.step_count = 0
        self.set_params(init_params(seed) if params is None else params)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.ash_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters --------------------------------------------------------
    def set_params(self, params) -> None:
        t = torch.as_tensor(np.asarray(params, dtype=np.float32)).to(self.dev).contiguous()
        if t.numel() != PARAM_COUNT:
            raise ValueError(f"expected {PARAM_COUNT} parameters")
        check(self.lib.ash_net_set_params(self.h, t.data_ptr(), current_stream()), "ash_net_set_params")
        torch.cuda.current_stream().synchronize()

    def get_params(self) -> np.ndarray:
        t = torch.empty(PARAM_COUNT, dtype=torch.float32, device=self.dev)
        check(self.lib.ash_net_get_params(self.h, t.data_ptr(), current_stream()), "ash_net_get_params")
        return t.cpu().numpy()

    # -- training path -----------------------------------------------------
    def _cols(self, u: torch.Tensor) -> torch.Tensor:
        if u.device.type != "cuda" or u.dtype != torch.float32 or u.dim() != 2 or u.shape[1] != self.nT:
            raise ValueError(f"expected a CUDA float32 tensor of shape (ncols, {self.nT})")
        if u.shape[0] > self.max_cols:
            raise ValueError(f"at most {self.max_cols} columns per call")
        return u.contiguous()

    def forward(self, u: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """Normalised network output y (ncols, nT) and column amplitudes A."""
        u = self._cols(u)
        y = torch.empty_like(u)
        amp = torch.empty(u.shape[0], dtype=torch.float32, device=self.dev)
        check(self.lib.ash_net_forward(self.h, u.data_ptr(), u.shape[0], y.data_ptr(), amp.data_ptr(),
                                       current_stream()), "ash_net_forward")
        return y, amp

    def loss_grad(self, u: torch.Tensor, u_next: torch.Tensor, c: float) -> Tuple[float, torch.Tensor]:
        """Mean-square loss against y* = (u_next - u) / (c A^2) and its parameter gradient."""
        u = self._cols(u)
        un = self._cols(u_next)
        self._loss.zero_()
        check(self.lib.ash_net_loss_grad(self.h, u.data_ptr(), un.data_ptr(), u.shape[0], float(c),
                                         self._loss.data_ptr(), self.grads.data_ptr(), current_stream()),
              "ash_net_loss_grad")
        return float(self._loss.item()), self.grads

    def adam(self, grads: torch.Tensor, lr: float, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8) -> None:
        self.step_count += 1
        check(self.lib.ash_net_adam(self.h, grads.data_ptr(), lr, beta1, beta2, eps, self.step_count,
                                    current_stream()), "ash_net_adam")

    # -- inference ---------------------------------------------------------
    def apply(self, u: torch.Tensor, c: float, dpdt_max: float, dT: float) -> torch.Tensor:
        """In place u += c A^2 y on columns with max|du/dt| <= dpdt_max (Pa/s); returns the fallback flags.

        Columns with c max|u[t+1] - u[t]| < 1e-5 (a Burgers increment below 1e-5 of their
        amplitude) and quiet 128-sample segments (below 1e-4 A) are left unchanged."""
        if u.device.type != "cuda" or u.dtype != torch.float32 or u.shape[-1] != self.nT or not u.is_contiguous():
            raise ValueError(f"expected a contiguous CUDA float32 tensor with last dimension {self.nT}")
        ncols = u.numel() // self.nT
        flags = torch.empty(ncols, dtype=torch.int32, device=self.dev)
        check(self.lib.ash_net_apply(self.h, u.data_ptr(), ncols, float(c), float(dpdt_max), float(dT),
                                     flags.data_ptr(), current_stream()), "ash_net_apply")
        return flags


# ---------------------------------------------------------------------------
# Training data from the plain KT kernel
# ---------------------------------------------------------------------------
def burgers_coefficient(params) -> float:
    """N = beta / (2 c0^3 rho0) (ASS:1297)."""
    return params.beta / (2.0 * params.c0 ** 3 * params.rho0)


def tone_bursts(n: int, nT: int, dT: float, f0: float, rng: np.random.Generator,
                p0_range=(1e5, 1.5e6)) -> np.ndarray:
    """Seeded Gaussian-enveloped bursts exp(-(1.05 t w0 / (n_c pi))^4) sin(w0 t + phi) p0
    (the pulse of validate_solver.py:41-57) with random amplitude, cycles, delay, phase and a
    +-15 % frequency jitter; float32 (n, nT)."""
    t = (np.arange(nT) - (nT - 1) / 2) * dT
    p0 = rng.uniform(*p0_range, size=(n, 1))
    ncyc = rng.uniform(3.0, 6.0, size=(n, 1))
    delay = rng.uniform(-0.25, 0.25, size=(n, 1)) * nT * dT
    phase = rng.uniform(0.0, 2 * np.pi, size=(n, 1))
    w0 = 2 * np.pi * f0 * rng.uniform(0.85, 1.15, size=(n, 1))
    tt = t[None, :] - delay
    env = np.exp(-(1.05 * tt * w0 / (ncyc * np.pi)) ** 4)
    return (env * np.sin(w0 * tt + phase) * p0).astype(np.float32)


def generate_pairs(params, nT: int, dz: float, n_pairs: int = 200_000, stride: int = 8,
                   pairs_per_column: int = 50, seed: int = 0,
                   progress: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
    """(before, after) column pairs of one plain-KT step over dz, on the GPU.

    Bursts with amplitudes 0.1-1.5 MPa are marched by the KT kernel; every
    ``stride``-th step contributes its (before, after) pair, so the set spans
    the waveforms from the source up to ``stride * pairs_per_column`` steps of
    nonlinear distortion.  Returns two float32 (n_pairs, nT) CUDA tensors.
    """
    from .plan import MarchPlan
    nx = 64
    cols = nx * nx
    rng = np.random.default_rng(seed)
    plan = MarchPlan((nx, nx, nT), params)
    before = torch.empty((n_pairs, nT), dtype=torch.float32, device="cuda")
    after = torch.empty_like(before)
    got = 0
    while got < n_pairs:
        u = torch.from_numpy(tone_bursts(cols, nT, params.dT, params.f0, rng)).cuda().view(nx, nx, nT)
        nxt = torch.empty_like(u)
        for k in range(stride * pairs_per_column):
            plan.nonlinear_step(u, _lib.NL_KT, dz, 1.0, 0, out=nxt)
            if k % stride == stride - 1:
                take = min(cols, n_pairs - got)
                before[got:got + take] = u.view(cols, -1)[:take]
                after[got:got + take] = nxt.view(cols, -1)[:take]
                got += take
                if got >= n_pairs:
                    break
            u, nxt = nxt, u
        if progress:
            print(f"  pairs {got}/{n_pairs}", flush=True)
    plan.close()
    return before, after


def train(net: LearnedStep, before: torch.Tensor, after: torch.Tensor, c: float, epochs: int = 4,
          batch: int = 256, lr: float = 2e-3, lr_final: float = 1e-4, seed: int = 0,
          progress: bool = False) -> list:
    """Adam on mini-batches of (before, after) pairs with a cosine learning-rate decay.

    Returns the mean loss of every epoch."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = before.shape[0]
    nb = n // batch
    total = epochs * nb
    hist = []
    it = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).cuda()
        acc = 0.0
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            loss, grads = net.loss_grad(before[idx], after[idx], c)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / max(total - 1, 1)))
            net.adam(grads, lr_t)
            acc += loss
            it += 1
        hist.append(acc / nb)
        if progress:
            print(f"  epoch {ep}: loss {hist[-1]:.3e}", flush=True)
    return hist
