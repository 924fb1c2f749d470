"""Learned WENO weights: a 5 -> 32 -> 32 -> 3 perceptron in place of the
WENO-JS nonlinear weights of ``weno.py:247-256``.

The network maps the normalised 5-cell stencil of a reconstruction to the
three candidate weights (softmax over logits + ln d, so an all-zero network
gives the ideal weights d); cells whose smoothness indicators flag a shock,
``max beta > tau (min beta + eps)``, keep the WENO-JS weights. Inference runs
inside the fused stage kernels (BF16 matrix-core products, FP32
accumulation); training runs on the device through the library's own
forward/backward kernel (``psa_mlp_grad``) and Adam update
(``psa_mlp_adam``). The reference has no counterpart: this is the learned
form of the reconstruction the solvers take through
:class:`LearnedWENOJS53`.
"""

from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import torch

from pyshocks_amd import _lib

NPARAM = _lib.MLP_NPARAM
RECORD = _lib.MLP_RECORD
NIN, NHID, NOUT = 5, 32, 3
IDEAL = (0.1, 0.6, 0.3)
TRAINED_PATH = Path(__file__).with_name("wenonet.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden layers, zero biases and a zero output layer (so that
    the untrained network gives the ideal weights)."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((NHID, NIN)) * math.sqrt(2.0 / NIN)
    w2 = rng.standard_normal((NHID, NHID)) * math.sqrt(2.0 / NHID)
    w3 = np.zeros((NOUT, NHID))
    p = np.concatenate([w1.ravel(), np.zeros(NHID), w2.ravel(), np.zeros(NHID),
                        w3.ravel(), np.zeros(NOUT)])
    return p.astype(np.float32)


@dataclass
class MLPWeights:
    """Device parameters of the network and the shock threshold."""

    params: torch.Tensor
    tau: float = 20.0

    def __post_init__(self) -> None:
        if self.params.dtype != torch.float32 or self.params.numel() != NPARAM:
            raise ValueError(f"expected {NPARAM} FP32 parameters")
        if not self.params.is_contiguous():
            raise ValueError("parameters must be contiguous")
        if not self.tau >= 0.0:
            raise ValueError("tau must be >= 0")

    @classmethod
    def from_numpy(cls, p: np.ndarray, tau: float = 20.0,
                   device: torch.device | str = "cuda") -> "MLPWeights":
        t = torch.as_tensor(np.asarray(p, dtype=np.float32)).to(device).contiguous()
        return cls(params=t, tau=float(tau))

    @classmethod
    def load(cls, path: str | os.PathLike[str] | None = None,
             device: torch.device | str = "cuda") -> "MLPWeights":
        """Weights saved by :meth:`save` (default: the trained network shipped
        with the package)."""
        with np.load(TRAINED_PATH if path is None else path, allow_pickle=False) as f:
            return cls.from_numpy(f["params"], float(f["tau"]), device)

    def save(self, path: str | os.PathLike[str], **meta: Any) -> None:
        np.savez(path, params=self.params.detach().cpu().numpy(),
                 tau=np.float64(self.tau), **meta)

    @property
    def ptr(self) -> int:
        return self.params.data_ptr()


def features(v: torch.Tensor) -> torch.Tensor:
    """Inputs of stencils ``v`` (..., 5), computed in ``v``'s dtype and
    returned as FP32 (the same formula the kernels use)."""
    mean = (((v[..., 0] + v[..., 4]) + (v[..., 1] + v[..., 3])) + v[..., 2]) * 0.2
    d = v - mean[..., None]
    s = d.abs().amax(dim=-1, keepdim=True)
    x = torch.where(s > 0, d / torch.where(s > 0, s, torch.ones_like(s)),
                    torch.zeros_like(d))
    return x.to(torch.float32)


def weights(w: MLPWeights, x: torch.Tensor, *, logits: bool = False
            ) -> torch.Tensor | tuple[torch.Tensor, torch.Tensor]:
    """Candidate weights (n, 3) of inputs ``x`` (n, 5) on the device."""
    x = x.to(torch.float32).contiguous()
    if x.ndim != 2 or x.shape[1] != NIN:
        raise ValueError("inputs must be (n, 5)")
    _lib.require_device(x, w.params)
    n = x.shape[0]
    omega = torch.empty((n, NOUT), dtype=torch.float32, device=x.device)
    z = torch.empty_like(omega) if logits else None
    _lib.call("psa_mlp_weights", w.ptr, x.data_ptr(), n, omega.data_ptr(),
              _lib.ptr(z), _lib.stream_of(x))
    return (omega, z) if logits else omega


def learned_reconstruct(w: MLPWeights, u: torch.Tensor, eps: float
                        ) -> tuple[torch.Tensor, torch.Tensor]:
    """``(ul, ur)`` of every cell of ``u`` with the learned weights (WENO-JS
    in flagged cells), zero padding outside the array as in
    ``reconstruction.py:358-377``."""
    u = u.contiguous()
    _lib.require_device(u, w.params)
    ul, ur = torch.empty_like(u), torch.empty_like(u)
    _lib.call("psa_learned_reconstruct", _lib.dtype_code(u.dtype), w.ptr, float(w.tau),
              u.data_ptr(), ul.data_ptr(), ur.data_ptr(), u.numel(), float(eps),
              _lib.stream_of(u))
    return ul, ur


# {{{ training


@dataclass
class Trainer:
    """Adam on the mean sample loss of ``psa_mlp_grad``:
    ``c (r^2 / (mean e^2 + delta) + lambda_js |omega - omega_JS|^2)`` with
    ``r = sum_k omega_k e_k`` the normalised reconstruction error."""

    weights: MLPWeights
    lr: float = 2.0e-3
    lambda_js: float = 1.0e-3
    delta: float = 1.0e-12
    beta1: float = 0.9
    beta2: float = 0.999
    adam_eps: float = 1.0e-8
    step_count: int = 0
    history: list[float] = field(default_factory=list)

    def __post_init__(self) -> None:
        self.m = torch.zeros_like(self.weights.params)
        self.v = torch.zeros_like(self.weights.params)
        self.grad = torch.zeros(NPARAM + 1, dtype=torch.float32,
                                device=self.weights.params.device)
        self._work: torch.Tensor | None = None

    def _workspace(self, n: int) -> torch.Tensor:
        size = ctypes.c_size_t(0)
        _lib.call("psa_mlp_grad_workspace_bytes", n, ctypes.byref(size))
        if self._work is None or self._work.numel() < size.value:
            self._work = torch.empty(size.value, dtype=torch.uint8,
                                     device=self.weights.params.device)
        return self._work

    def loss_and_grad(self, samples: torch.Tensor) -> torch.Tensor:
        """Enqueue the loss and gradient of a batch of sample records (n, 12);
        returns the device vector of ``NPARAM`` gradients followed by the loss."""
        if samples.dtype != torch.float32 or samples.ndim != 2 or samples.shape[1] != RECORD:
            raise ValueError(f"samples must be FP32 (n, {RECORD})")
        samples = samples.contiguous()
        _lib.require_device(samples, self.weights.params)
        n = samples.shape[0]
        work = self._workspace(n)
        _lib.call("psa_mlp_grad", self.weights.ptr, samples.data_ptr(), n,
                  float(self.lambda_js), float(self.delta), self.grad.data_ptr(),
                  work.data_ptr(), work.numel(), _lib.stream_of(samples))
        return self.grad

    def step(self, samples: torch.Tensor, lr: float | None = None) -> torch.Tensor:
        """One Adam step on a batch; returns the batch loss (device scalar)."""
        self.loss_and_grad(samples)
        self.step_count += 1
        _lib.call("psa_mlp_adam", self.weights.ptr, self.grad.data_ptr(),
                  self.m.data_ptr(), self.v.data_ptr(), self.step_count,
                  float(self.lr if lr is None else lr), self.beta1, self.beta2,
                  self.adam_eps, _lib.stream_of(samples))
        return self.grad[NPARAM]

    def fit(self, samples: torch.Tensor, *, epochs: int = 4, batch: int = 65536,
            seed: int = 0, lr_final: float | None = None) -> list[float]:
        """Epochs of shuffled mini-batches (cosine learning-rate decay to
        ``lr_final``); records the mean loss of each epoch."""
        n = samples.shape[0]
        gen = torch.Generator(device=samples.device)
        gen.manual_seed(seed)
        nb = max(1, n // batch)
        total = epochs * nb
        lr0 = self.lr
        lr1 = lr0 * 0.05 if lr_final is None else lr_final
        k = 0
        for _ in range(epochs):
            perm = torch.randperm(n, device=samples.device, generator=gen)
            acc = torch.zeros((), dtype=torch.float64, device=samples.device)
            for b in range(nb):
                lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + math.cos(math.pi * k / total))
                acc += self.step(samples[perm[b * batch:(b + 1) * batch]], lr).double()
                k += 1
            self.history.append(float(acc) / nb)
        return self.history


# }}}
