"""Discrete adjoint of the fused SSP-RK3 step.

The reference computes adjoints with a dense ``jax.jacfwd`` Jacobian of one
step (``timestepping.py:155-215``, feasible only for small grids, as noted at
``:167-172``). Here the transposed Jacobian is applied matrix-free by the
library's reverse stage kernels:

* :func:`ssprk33_step_vjp` -- ``J(u)^T p`` for one step (used by
  :func:`~pyshocks_amd.timestepping.adjoint_step`);
* :func:`checkpointed_solve` / :class:`SSPRK33Solve` -- a fixed-dt solve of
  ``nsteps`` steps whose backward pass recomputes windows of ``every`` steps
  from checkpoints and sweeps them in reverse, exposed to torch autograd;
* :class:`BatchedCheckpointedSolve` -- the same for many instances of one
  grid, one launch per step for all of them.

Cotangents use the grid layout (ghost entries are ignored on input and are
zero on output).

Rusanov, Godunov and Engquist-Osher schemes with
:class:`~pyshocks_amd.reconstruction.LearnedWENOJS53` are supported too (the learned weights have no counterpart in the reference).
Their transposed Jacobian is that of the forward solver's smooth surrogate,
with each cell's regime fixed at the forward state:
WENO-JS in shock cells, the ideal weights in smooth cells, and the network in
learned cells. The network's weights are rounded to BF16 as in the forward
pass; its activations stay unrounded, since the forward pass's rounding of
them is a step function whose derivative is zero almost everywhere.
"""

from __future__ import annotations

import ctypes
from typing import Any

import torch

from pyshocks_amd import _lib


def _plan_of(source: Any) -> _lib.PsaSolverDesc:
    from pyshocks_amd.timestepping import _scheme_plan

    sp = _scheme_plan(source)
    if sp is None:
        raise NotImplementedError(
            "the adjoint needs a BoundOperator of a hot-path scheme")
    return sp[0]


def _dt_tensor(dt: Any, like: torch.Tensor) -> torch.Tensor:
    return torch.as_tensor(dt, dtype=like.dtype, device=like.device).reshape(()).contiguous()


def _array_bytes(m: int, dtype: torch.dtype) -> int:
    item = 8 if dtype == torch.float64 else 4
    return ((m * item + 255) // 256) * 256


def operator_vjp(source: Any, u: torch.Tensor, mu: torch.Tensor) -> torch.Tensor:
    """``J_L(u)^T mu`` for the operator ``L = apply_operator(scheme, grid,
    bc, t, .)`` of a bound source."""
    desc = _plan_of(source)
    _lib.require_device(u, mu)
    out = torch.empty_like(u)
    _lib.call("psa_ssprk33_stage_vjp", ctypes.byref(desc), 0, _lib.ptr(u),
              _lib.ptr(mu), None, None, _lib.ptr(out), None, _lib.stream_of(u))
    return out


def ssprk33_step_vjp(source: Any, u: torch.Tensor, p: torch.Tensor,
                     dt: Any) -> torch.Tensor:
    """``J(u)^T p`` for one SSP-RK3 step of size ``dt`` from ``u``."""
    desc = _plan_of(source)
    _lib.require_device(u, p)
    dt_dev = _dt_tensor(dt, u)
    ab = _array_bytes(u.numel(), u.dtype)
    work = torch.empty(4 * ab, dtype=torch.uint8, device=u.device)
    out = torch.empty_like(u)
    _lib.call("psa_ssprk33_step_vjp", ctypes.byref(desc), _lib.ptr(u), _lib.ptr(p),
              _lib.ptr(out), _lib.ptr(work), work.numel(), _lib.ptr(dt_dev),
              _lib.stream_of(u))
    return out


class CheckpointedSolve:
    """A fixed-dt forward solve that keeps checkpoints for its adjoint."""

    def __init__(self, source: Any, u0: torch.Tensor, nsteps: int, dt: Any,
                 every: int = 50, schedule: int | None = None) -> None:
        self.desc = _plan_of(source)
        if schedule is not None:
            # PSA_SCHEDULE_STEP (default): whole-step kernels for the forward and
            # recompute passes where they apply; PSA_SCHEDULE_TILE: stage kernels
            self.desc.schedule = int(schedule)
        _lib.require_device(u0)
        self.nsteps = int(nsteps)
        self.every = int(every)
        self.dt = _dt_tensor(dt, u0)
        nbytes = ctypes.c_size_t(0)
        _lib.call("psa_adjoint_workspace_bytes", ctypes.byref(self.desc),
                  self.nsteps, self.every, ctypes.byref(nbytes))
        self.workspace = torch.empty(nbytes.value, dtype=torch.uint8, device=u0.device)
        self.u_final = torch.empty_like(u0)
        _lib.call("psa_ssprk33_forward_checkpointed", ctypes.byref(self.desc),
                  _lib.ptr(u0), _lib.ptr(self.u_final), self.nsteps,
                  _lib.ptr(self.dt), self.every, _lib.ptr(self.workspace),
                  self.workspace.numel(), _lib.stream_of(u0))

    def gradient(self, lam_final: torch.Tensor) -> torch.Tensor:
        """``(d u_final / d u0)^T lam_final``."""
        lam_final = lam_final.contiguous()
        _lib.require_device(lam_final)
        grad = torch.empty_like(lam_final)
        _lib.call("psa_ssprk33_adjoint", ctypes.byref(self.desc), self.nsteps,
                  _lib.ptr(self.dt), self.every, _lib.ptr(lam_final),
                  _lib.ptr(grad), _lib.ptr(self.workspace), self.workspace.numel(),
                  _lib.stream_of(lam_final))
        return grad


def checkpointed_solve(source: Any, u0: torch.Tensor, nsteps: int, dt: Any,
                       every: int = 50) -> CheckpointedSolve:
    return CheckpointedSolve(source, u0, nsteps, dt, every)


class BatchedCheckpointedSolve:
    """:class:`CheckpointedSolve` for ``B`` instances of one grid at once:
    ``u0`` of shape ``(B, m)``, ``dt`` a scalar or one step size per
    instance. Each step of the forward, recompute and reverse passes is one
    launch for the whole batch, so the gradients share its latency."""

    def __init__(self, source: Any, u0: torch.Tensor, nsteps: int, dt: Any,
                 every: int = 50) -> None:
        self.desc = _plan_of(source)
        _lib.require_device(u0)
        if u0.ndim != 2:
            raise ValueError("u0 must have shape (batch, m)")
        self.batch, self.m = int(u0.shape[0]), int(u0.shape[1])
        # rows padded to a multiple of 32 cells (16-byte aligned in either dtype)
        self.ld = ((self.m + 31) // 32) * 32
        self.nsteps = int(nsteps)
        self.every = int(every)
        self.dt = torch.as_tensor(dt, dtype=u0.dtype, device=u0.device).expand(
            self.batch).contiguous()
        nbytes = ctypes.c_size_t(0)
        _lib.call("psa_adjoint_workspace_bytes_batch", ctypes.byref(self.desc),
                  self.batch, self.ld, self.nsteps, self.every, ctypes.byref(nbytes))
        self.workspace = torch.empty(nbytes.value, dtype=torch.uint8, device=u0.device)
        u0p = self._padded(u0)
        self._u_final = torch.empty_like(u0p)
        _lib.call("psa_ssprk33_forward_checkpointed_batch", ctypes.byref(self.desc),
                  self.batch, self.ld, _lib.ptr(u0p), _lib.ptr(self._u_final),
                  self.nsteps, _lib.ptr(self.dt), self.every, _lib.ptr(self.workspace),
                  self.workspace.numel(), _lib.stream_of(u0))

    def _padded(self, a: torch.Tensor) -> torch.Tensor:
        if tuple(a.shape) != (self.batch, self.m):
            raise ValueError(f"expected shape ({self.batch}, {self.m})")
        out = torch.zeros((self.batch, self.ld), dtype=a.dtype, device=a.device)
        out[:, : self.m] = a
        return out

    @property
    def u_final(self) -> torch.Tensor:
        """The final states, shape ``(B, m)`` (a view)."""
        return self._u_final[:, : self.m]

    def gradient(self, lam_final: torch.Tensor) -> torch.Tensor:
        """``(d u_final_b / d u0_b)^T lam_final_b`` for every instance b,
        shape ``(B, m)``."""
        if not lam_final.is_cuda:
            raise ValueError("lam_final must be on the GPU")
        lam = self._padded(lam_final)
        grad = torch.empty_like(lam)
        _lib.call("psa_ssprk33_adjoint_batch", ctypes.byref(self.desc), self.batch,
                  self.ld, self.nsteps, _lib.ptr(self.dt), self.every, _lib.ptr(lam),
                  _lib.ptr(grad), _lib.ptr(self.workspace), self.workspace.numel(),
                  _lib.stream_of(lam))
        return grad[:, : self.m]


class SSPRK33Solve(torch.autograd.Function):
    """``u_final = S^nsteps(u0)`` for the fused fixed-dt SSP-RK3 step, with a
    backward pass by the reverse kernels and checkpoint/recompute every
    ``every`` steps."""

    @staticmethod
    def forward(ctx: Any, u0: torch.Tensor, source: Any, nsteps: int, dt: Any,
                every: int = 50) -> torch.Tensor:
        solve = CheckpointedSolve(source, u0.detach().contiguous(), nsteps, dt, every)
        ctx.solve = solve
        return solve.u_final.clone()

    @staticmethod
    def backward(ctx: Any, grad_out: torch.Tensor) -> tuple[Any, ...]:
        grad = ctx.solve.gradient(grad_out.contiguous())
        return grad, None, None, None, None


def ssprk33_solve(u0: torch.Tensor, source: Any, nsteps: int, dt: Any,
                  every: int = 50) -> torch.Tensor:
    """Differentiable fixed-dt solve (see :class:`SSPRK33Solve`)."""
    return SSPRK33Solve.apply(u0, source, nsteps, dt, every)
