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
    """Shock regime: ``max beta > tau (min beta + eps)`` keeps WENO-JS."""
    sigma: float = 0.0
    """Smooth regime: ``max beta <= sigma min beta`` takes the ideal weights
    without evaluating the network (0: only constant stencils)."""

    def __post_init__(self) -> None:
        if self.params.dtype != torch.float32 or self.params.numel() != NPARAM:
            raise ValueError(f"expected {NPARAM} FP32 parameters")
        if not self.params.is_contiguous():
            raise ValueError("parameters must be contiguous")
        if not self.tau >= 0.0 or not self.sigma >= 0.0:
            raise ValueError("tau and sigma must be >= 0")

    @classmethod
    def from_numpy(cls, p: np.ndarray, tau: float = 20.0, sigma: float = 0.0,
                   device: torch.device | str = "cuda") -> "MLPWeights":
        t = torch.as_tensor(np.asarray(p, dtype=np.float32)).to(device).contiguous()
        return cls(params=t, tau=float(tau), sigma=float(sigma))

    @classmethod
    def load(cls, path: str | os.PathLike[str] | None = None,
             device: torch.device | str = "cuda") -> "MLPWeights":
        """Weights saved by :meth:`save` (default: the trained network shipped
        with the package)."""
        with np.load(TRAINED_PATH if path is None else path, allow_pickle=False) as f:
            sigma = float(f["sigma"]) if "sigma" in f.files else 0.0
            return cls.from_numpy(f["params"], float(f["tau"]), sigma, device)

    def save(self, path: str | os.PathLike[str], **meta: Any) -> None:
        np.savez(path, params=self.params.detach().cpu().numpy(),
                 tau=np.float64(self.tau), sigma=np.float64(self.sigma), **meta)

    @property
    def ptr(self) -> int:
        return self.params.data_ptr()


def features(v: torch.Tensor) -> torch.Tensor:
    """Inputs of stencils ``v`` (..., 5), computed in ``v``'s dtype and
    returned as FP32 (the same formula the kernels use)."""
    mean = (((v[..., 0] + v[..., 4]) + (v[..., 1] + v[..., 3])) + v[..., 2]) * 0.2
    d = v - mean[..., None]
    s = d.abs().amax(dim=-1, keepdim=True)
    inv = 1.0 / torch.where(s > 0, s, torch.ones_like(s))
    x = torch.where(s > 0, d * inv, torch.zeros_like(d))
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
              float(w.sigma), u.data_ptr(), ul.data_ptr(), ur.data_ptr(), u.numel(),
              float(eps), _lib.stream_of(u))
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
        ``lr_final``); records the mean loss of each epoch. Raises
        ``RuntimeError`` after an epoch whose loss or parameters are not
        finite (a diverged run, whose weights would turn solutions into NaN)."""
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
            if not (math.isfinite(self.history[-1])
                    and bool(torch.isfinite(self.weights.params).all())):
                raise RuntimeError(f"training diverged in epoch {len(self.history)} "
                                   f"(loss {self.history[-1]}); lower the learning rate")
        return self.history


# }}}


# {{{ training data


def training_ics(count: int, seed: int = 0) -> list:
    """Seeded sine-plus-step initial conditions ``c + sum_{k=1..8} A_k
    sin(2 pi k x + phi_k) + H 1[x >= xs]`` with ``c ~ U[0, 1]``,
    ``A_k ~ U[-1, 1] / k``, ``H ~ U[-1, 1]``, ``xs ~ U(0.05, 0.95)``."""
    from pyshocks_amd.funcs import SineStep

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        c = rng.uniform(0.0, 1.0)
        amps = rng.uniform(-1.0, 1.0, 8) / np.arange(1, 9)
        phis = rng.uniform(0.0, 2.0 * np.pi, 8)
        out.append(SineStep(c=float(c), amplitudes=tuple(float(a) for a in amps),
                            wavenumbers=tuple(range(1, 9)),
                            phases=tuple(float(p) for p in phis),
                            step_height=float(rng.uniform(-1.0, 1.0)),
                            step_location=float(rng.uniform(0.05, 0.95))))
    return out


def _js_weights(v: torch.Tensor, eps: float) -> tuple[torch.Tensor, torch.Tensor]:
    """WENO-JS weights (``weno.py:247-256``) and indicators of right-side
    stencils ``v`` (..., 5), FP64."""
    um2, um1, u0, up1, up2 = v.unbind(-1)
    b0 = 13 / 12 * (um2 - 2 * um1 + u0) ** 2 + 1 / 4 * (um2 - 4 * um1 + 3 * u0) ** 2
    b1 = 13 / 12 * (um1 - 2 * u0 + up1) ** 2 + 1 / 4 * (um1 - up1) ** 2
    b2 = 13 / 12 * (u0 - 2 * up1 + up2) ** 2 + 1 / 4 * (3 * u0 - 4 * up1 + up2) ** 2
    beta = torch.stack([b0, b1, b2], dim=-1)
    d = torch.tensor(IDEAL, dtype=v.dtype, device=v.device)
    alpha = d / (eps + beta) ** 2
    return alpha / alpha.sum(dim=-1, keepdim=True), beta


def _candidates(v: torch.Tensor) -> torch.Tensor:
    um2, um1, u0, up1, up2 = v.unbind(-1)
    return torch.stack([11 / 6 * u0 - 7 / 6 * um1 + 2 / 6 * um2,
                        2 / 6 * up1 + 5 / 6 * u0 - 1 / 6 * um1,
                        -1 / 6 * up2 + 5 / 6 * up1 + 2 / 6 * u0], dim=-1)


def stencil_records(u: torch.Tensor, uface: torch.Tensor, *, eps: float, tau: float,
                    sigma: float = 0.0, smin: float = 1.0e-9) -> torch.Tensor:
    """Training records of every interior cell of the periodic states ``u``
    (B, n + 6) against the exact face values ``uface`` (B, n + 1): the right
    reconstruction of cell i with face i + 1/2 and the left one (mirrored
    stencil) with face i - 1/2. Cells in the shock regime (they keep
    WENO-JS) or the smooth regime (ideal weights, ``sigma``) and
    near-constant stencils get weight 0."""
    n = u.shape[1] - 6
    idx = torch.arange(3, n + 3, device=u.device)
    v = torch.stack([u[:, idx + k] for k in (-2, -1, 0, 1, 2)], dim=-1)  # (B, n, 5)
    recs = []
    for side in ("right", "left"):
        vs = v if side == "right" else v.flip(-1)
        ustar = uface[:, 1:] if side == "right" else uface[:, :-1]
        mean = (((vs[..., 0] + vs[..., 4]) + (vs[..., 1] + vs[..., 3])) + vs[..., 2]) * 0.2
        s = (vs - mean[..., None]).abs().amax(dim=-1)
        x = features(vs)
        wjs, beta = _js_weights(vs, eps)
        e = (_candidates(vs) - ustar[..., None]) / torch.where(s > 0, s, torch.ones_like(s))[..., None]
        bmax, bmin = beta.amax(dim=-1), beta.amin(dim=-1)
        learned = (bmax <= tau * (bmin + eps)) & (bmax > sigma * bmin)
        c = ((s > smin) & learned).to(torch.float64)
        rec = torch.cat([x.double(), e, wjs, c[..., None]], dim=-1)
        recs.append(rec.reshape(-1, RECORD))
    return torch.cat(recs).to(torch.float32)


def make_training_samples(ics: list, *, n: int = 4096, times: tuple[float, ...] = (
        0.01, 0.02, 0.04, 0.06, 0.08, 0.1, 0.13, 0.16, 0.2, 0.25, 0.3, 0.35, 0.4, 0.45, 0.5),
        eps: float = 1.0e-6, tau: float = 20.0, sigma: float = 0.0,
        device: str = "cuda") -> torch.Tensor:
    """Records from plain WENO-JS5 / Rusanov / SSP-RK3 runs (FP64, the
    batched solver) of ``ics`` on ``n`` cells, snapshotted at ``times``,
    against the exact Hopf-Lax face values at the snapshot times."""
    from pyshocks_amd import exact
    from pyshocks_amd.batched import BatchedSolver

    B = len(ics)
    u0 = exact.burgers_cell_averages(ics, [n] * B, 0.0, device=device)
    solver = BatchedSolver([n] * B, dtype=torch.float64, eps=eps, device=device)
    out = []
    for t in times:
        solver.set_states(u0)
        res = solver.solve(t)
        if np.any(res.status != 0):
            raise RuntimeError("a training run did not finish")
        u = torch.stack([solver.state(i) for i in range(B)])
        uf = exact.burgers_face_values(ics, n, res.t, device=device)
        out.append(stencil_records(u, uf, eps=eps, tau=tau, sigma=sigma))
    return torch.cat(out)


# }}}


# {{{ evaluation


def l1_errors(ics: list, ns: np.ndarray, t: float, *, mlp: MLPWeights | None = None,
              dtype: torch.dtype = torch.float64, eps: float = 1.0e-6,
              device: str = "cuda", tiles: tuple[int, int] | None = None,
              team_cells: int | None = None,
              ) -> tuple[np.ndarray, dict[str, Any]]:
    """Interior L1 errors ``sum_i |u_i - u_exact_i| h`` at time ``t`` of the
    batched solver (WENO-JS, or the learned weights ``mlp``) on instances
    ``ics`` with ``ns`` cells, against exact Hopf-Lax cell averages.
    ``team_cells`` splits larger instances among teams of workgroups."""
    from pyshocks_amd import exact
    from pyshocks_amd.batched import BatchedSolver

    ns = np.asarray(ns, dtype=np.int64)
    u0 = exact.burgers_cell_averages(ics, ns, 0.0, device=device)
    ue = exact.burgers_cell_averages(ics, ns, t, device=device)
    solver = BatchedSolver(ns, dtype=dtype, eps=eps, device=device, mlp=mlp, tiles=tiles,
                           team_cells=team_cells)
    solver.set_states(u0)
    torch.cuda.synchronize()
    import time

    t0 = time.perf_counter()
    res = solver.solve(t)
    elapsed = time.perf_counter() - t0
    if np.any(res.status != 0):
        raise RuntimeError("an evaluation run did not finish")
    err = np.array([float((solver.state(i)[3:-3].double() - ue[i][3:-3]).abs().sum()) / n
                    for i, n in enumerate(ns)])
    return err, {"seconds": elapsed, "cell_updates": float(np.sum(ns * res.steps)),
                 "steps": res.steps}


# }}}


def main(argv: list[str] | None = None) -> int:
    """Train the network on the device and save it:
    ``python -m pyshocks_amd.learned --out wenonet.npz``."""
    import argparse
    import time

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", default=str(TRAINED_PATH))
    ap.add_argument("--load", default=None,
                    help="validate these weights instead of training new ones")
    ap.add_argument("--ics", type=int, default=84)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--lambda-js", type=float, default=3e-3)
    ap.add_argument("--delta", type=float, default=1e-12)
    ap.add_argument("--tau", type=float, default=20.0)
    ap.add_argument("--sigma", type=float, default=1.5,
                    help="smooth-regime threshold stored with the weights (training "
                         "always uses every record outside the shock regime)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sigma-sweep", type=float, nargs="*", default=[])
    args = ap.parse_args(argv)

    _lib.load()
    if args.load:
        w = MLPWeights.load(args.load)
        w = MLPWeights(w.params, args.tau, args.sigma)
    else:
        t0 = time.perf_counter()
        # every record outside the shock regime: training on the learned
        # regime alone (a few 1e4 records next to shocks) generalises badly
        samples = make_training_samples(training_ics(args.ics, seed=1000 + args.seed),
                                        n=args.n, tau=args.tau, sigma=0.0)
        total = samples.shape[0]
        samples = samples[samples[:, 11] > 0].contiguous()
        print(f"[train] {total} records ({samples.shape[0]} used) in "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        w = MLPWeights.from_numpy(init_params(args.seed), args.tau, args.sigma)
        tr = Trainer(w, lr=args.lr, lambda_js=args.lambda_js, delta=args.delta)
        t0 = time.perf_counter()
        tr.fit(samples, epochs=args.epochs, batch=args.batch, seed=args.seed)
        torch.cuda.synchronize()
        print(f"[train] {args.epochs} epochs in {time.perf_counter() - t0:.1f} s; loss per "
              f"epoch {' '.join(f'{x:.4e}' for x in tr.history)}", flush=True)
        w.save(args.out, loss=np.asarray(tr.history), ics=args.ics, n=args.n,
               lambda_js=args.lambda_js, delta=args.delta, seed=args.seed)

    # validation on unseen instances, FP64: plain WENO-JS, the untrained
    # network (ideal weights with the same fallback) and the trained one
    ics = training_ics(32, seed=2000 + args.seed)
    ns = np.round(np.exp(np.random.default_rng(args.seed).uniform(
        np.log(1024), np.log(8192), 32))).astype(np.int64)
    ideal = MLPWeights.from_numpy(np.zeros(NPARAM, np.float32), args.tau)
    for tv in (0.05, 0.2, 0.5):
        e0, _ = l1_errors(ics, ns, tv)
        ei, _ = l1_errors(ics, ns, tv, mlp=ideal)
        e1, _ = l1_errors(ics, ns, tv, mlp=w)
        print(f"[train] validation t={tv}: mean L1 plain {e0.mean():.4e} ideal+fallback "
              f"{ei.mean():.4e} learned {e1.mean():.4e} (learned/plain {e1.mean() / e0.mean():.4f},"
              f" learned/ideal {e1.mean() / ei.mean():.4f})", flush=True)
    for sg in args.sigma_sweep:
        ws = MLPWeights(w.params, w.tau, sg)
        ratios = []
        for tv in (0.05, 0.5):
            e0, i0 = l1_errors(ics, ns, tv)
            e1, i1 = l1_errors(ics, ns, tv, mlp=ws)
            ratios.append(f"t={tv}: learned/plain {e1.mean() / e0.mean():.4f} "
                          f"time learned/plain {i1['seconds'] / i0['seconds']:.2f}")
        print(f"[train] sigma {sg}: " + "; ".join(ratios), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
