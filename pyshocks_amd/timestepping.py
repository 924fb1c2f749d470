"""Time integration: steppers, the ``step`` generator and ``advance``.

Mirrors ``timestepping.py`` of pyshocks: ``StepCompleted`` (``:47-65``),
``AdjointStepCompleted`` (``:69-71``), ``Stepper`` (``:75-86``), ``step``
(``:89-152``), ``adjoint_step`` (``:155-215``), ``advance`` (``:219-233``),
the fixed-step helpers (``:242-261``), ``ForwardEuler`` (``:290-298``) and
``SSPRK33`` (``:308-320``).

Two execution paths sit behind the same API:

* **fused** -- when ``stepper.source`` is a :class:`BoundOperator` of a
  hot-path scheme (Burgers Rusanov/Godunov or advection Godunov with
  WENO-JS5 and periodic boundaries) and ``stepper.predict_timestep`` is a
  :class:`CourantTimestep` or :class:`FixedTimestep`, every step is three
  fused HIP stage kernels and the step control (termination test, CFL time
  step, the ``min(dt, tfinal - t) + 1e-15`` rule, the finiteness check) runs
  on the device;
* **generic** -- any other callables: the reference's host loop, calling
  ``source`` three times per step.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from functools import singledispatch
from typing import TYPE_CHECKING, Any, Callable, Iterator

import numpy as np
import torch

from pyshocks_amd import _lib
from pyshocks_amd.checkpointing import Checkpoint, load, save

if TYPE_CHECKING:
    from pyshocks_amd.grid import Grid
    from pyshocks_amd.schemes import Boundary, SchemeBase

# {{{ interface


@dataclass(frozen=True)
class StepCompleted:
    t: Any
    """Current time."""
    tfinal: Any
    """Final time of the simulation."""
    dt: Any
    """Time step used to get to :attr:`t`."""
    iteration: int
    """Current iteration."""
    u: torch.Tensor
    """Solution at the time :attr:`t`."""

    def __str__(self) -> str:
        return (f"[{self.iteration:5d}] "
                f"t = {float(self.t):.5e} / {float(self.tfinal):.5e} "
                f"dt {float(self.dt):.5e}")


@dataclass(frozen=True)
class AdjointStepCompleted(StepCompleted):
    p: torch.Tensor
    """Adjoint solution at the time :attr:`~StepCompleted.t`."""


@dataclass(frozen=True)
class Stepper:
    """Generic time stepping method for first-order ODEs."""

    predict_timestep: Callable[[Any, torch.Tensor], Any]
    source: Callable[[Any, torch.Tensor], torch.Tensor]
    checkpoint: Checkpoint | None


@dataclass(frozen=True)
class ForwardEuler(Stepper):
    """Forward Euler (generic path only)."""


@dataclass(frozen=True)
class SSPRK33(Stepper):
    """The optimal third order SSP Runge-Kutta method with 3 stages."""


# }}}


# {{{ bound callables recognised by the fused path


@dataclass(frozen=True)
class BoundOperator:
    """``source(t, u) = apply_operator(scheme, grid, bc, t, u)``."""

    scheme: SchemeBase
    grid: Grid
    bc: Boundary

    def __call__(self, t: Any, u: torch.Tensor) -> torch.Tensor:
        from pyshocks_amd.schemes import apply_operator

        return apply_operator(self.scheme, self.grid, self.bc, t, u)


@dataclass(frozen=True)
class CourantTimestep:
    """``predict_timestep(t, u) = theta * predict_timestep(scheme, grid, bc, t, u)``
    (the harness form of ``examples/burgers.py:161-162``)."""

    scheme: SchemeBase
    grid: Grid
    bc: Boundary
    theta: float = 1.0

    def __call__(self, t: Any, u: torch.Tensor) -> torch.Tensor:
        from pyshocks_amd.schemes import predict_timestep

        return self.theta * predict_timestep(self.scheme, self.grid, self.bc, t, u)


@dataclass(frozen=True)
class FixedTimestep:
    """``predict_timestep(t, u) = dt`` (``tests/test_convergence.py:151-157``)."""

    dt: float

    def __call__(self, t: Any, u: torch.Tensor) -> float:
        return self.dt


def make_ssprk33(scheme: SchemeBase, grid: Grid, bc: Boundary, *,
                 theta: float = 1.0, dt: float | None = None,
                 checkpoint: Checkpoint | None = None) -> SSPRK33:
    """The stepper the reference harness builds (``examples/burgers.py:164-179``),
    with callables that the fused path recognises."""
    predict: Any = CourantTimestep(scheme, grid, bc, theta) if dt is None \
        else FixedTimestep(float(dt))
    return SSPRK33(predict_timestep=predict, source=BoundOperator(scheme, grid, bc),
                   checkpoint=checkpoint)


# }}}


# {{{ fused plan


def _np_dtype(dtype: torch.dtype) -> type:
    return np.float64 if dtype == torch.float64 else np.float32


@dataclass
class FusedPlan:
    """Solver descriptor and step-control parameters of a fused stepper."""

    desc: _lib.PsaSolverDesc
    dtype: torch.dtype
    device: torch.device
    m: int
    dt_mode: int
    theta: float
    cfl_num: float
    fixed_dt: float
    keepalive: tuple[Any, ...]

    @property
    def np_dtype(self) -> type:
        return _np_dtype(self.dtype)

    def new_control(self, u0: torch.Tensor, t: Any, tfinal: Any,
                    maxit: int | None) -> torch.Tensor:
        ctrl = torch.empty(ctypes.sizeof(_lib.PsaControl), dtype=torch.uint8,
                           device=self.device)
        _lib.call("psa_control_init", ctypes.byref(self.desc), _lib.ptr(u0),
                  _lib.ptr(ctrl), float(t), float(tfinal),
                  -1 if maxit is None else int(maxit), self.dt_mode,
                  self.theta, self.cfl_num, self.fixed_dt, _lib.stream_of(u0))
        return ctrl

    def steps(self, u_in: torch.Tensor, u_out: torch.Tensor, k1: torch.Tensor,
              k2: torch.Tensor, ctrl: torch.Tensor, nsteps: int) -> None:
        _lib.call("psa_ssprk33_steps", ctypes.byref(self.desc), _lib.ptr(u_in),
                  _lib.ptr(u_out), _lib.ptr(k1), _lib.ptr(k2), _lib.ptr(ctrl),
                  int(nsteps), _lib.stream_of(u_in))

    def stage(self, stage: int, un: torch.Tensor | None, uk: torch.Tensor,
              out: torch.Tensor, ctrl: torch.Tensor | None = None,
              dt: torch.Tensor | None = None) -> None:
        """One SSP-RK3 stage (1, 2 or 3) with the control block or a device
        ``dt``."""
        _lib.call("psa_ssprk33_stage", ctypes.byref(self.desc), int(stage),
                  _lib.ptr(un), _lib.ptr(uk), _lib.ptr(out), _lib.ptr(ctrl),
                  _lib.ptr(dt), _lib.stream_of(uk))

    def advance(self, u: torch.Tensor, out: torch.Tensor, k1: torch.Tensor,
                k2: torch.Tensor, dt: torch.Tensor) -> None:
        _lib.call("psa_ssprk33_advance", ctypes.byref(self.desc), _lib.ptr(u),
                  _lib.ptr(out), _lib.ptr(k1), _lib.ptr(k2), _lib.ptr(dt),
                  _lib.stream_of(u))

    def uses_step_kernel(self) -> bool:
        """Whether :meth:`steps` runs one whole-step kernel per step."""
        return not _uses_ring_kernel(self.desc) and _uses_step_kernel(self.desc)

    def uses_ring_kernel(self) -> bool:
        """Whether :meth:`steps` runs all its steps in one launch of one
        workgroup that holds the state in LDS."""
        return _uses_ring_kernel(self.desc)

    @staticmethod
    def read_control(ctrl: torch.Tensor) -> _lib.PsaControl:
        host = ctrl.cpu().numpy().tobytes()
        return _lib.PsaControl.from_buffer_copy(host)


def _uses_step_kernel(d: _lib.PsaSolverDesc) -> bool:
    """The library's rule for the whole-step kernel: schedule default or
    step, a constant or no velocity, no Rusanov ``nu``, not Lax-Friedrichs,
    tiles (256, 4) or (256, 2), or (256, 1) for WENO-JS5 (learned weights
    or not); by default 1 cell per thread below 2^17 cells where this kernel
    takes it (then it also takes 2, the answer here), 2 below 2^20, 4
    above."""
    if d.schedule not in (_lib.PSA_SCHEDULE_DEFAULT, _lib.PSA_SCHEDULE_STEP):
        return False
    if d.nu or (d.flux == _lib.PSA_FLUX_ADVECTION_UPWIND and not d.velocity_is_const):
        return False
    if d.flux == _lib.PSA_FLUX_LAX_FRIEDRICHS:
        return False
    block = d.block if d.block > 0 else 256
    cpt = d.cells_per_thread if d.cells_per_thread > 0 else (2 if d.n < (1 << 20) else 4)
    if block == 256 and cpt == 1:
        return d.rec == _lib.PSA_REC_WENOJS53 and d.flux != _lib.PSA_FLUX_ESWENO
    return block == 256 and cpt in (2, 4)


def _uses_ring_kernel(d: _lib.PsaSolverDesc) -> bool:
    """The library's rule for the ring schedule: schedule default or ring,
    periodic, 8 to ``PSA_RING_MAX_CELLS`` cells, no learned weights, a
    constant or no velocity, no Rusanov ``nu``, not Lax-Friedrichs."""
    if d.schedule not in (_lib.PSA_SCHEDULE_DEFAULT, _lib.PSA_SCHEDULE_RING):
        return False
    if d.bc != _lib.PSA_BC_PERIODIC or d.mlp or d.nu:
        return False
    if d.flux == _lib.PSA_FLUX_ADVECTION_UPWIND and not d.velocity_is_const:
        return False
    if d.flux == _lib.PSA_FLUX_LAX_FRIEDRICHS:
        return False
    return 8 <= d.n <= _lib.PSA_RING_MAX_CELLS and 2 <= d.g <= d.n


def _scheme_plan(source: Any, *, fills: bool = False
                 ) -> tuple[_lib.PsaSolverDesc, tuple[Any, ...]] | None:
    """The descriptor of a bound operator the fused kernels run, or ``None``:
    periodic, or with ``fills`` (the whole-step kernel of
    ``psa_ssprk33_steps``, which fills those ghost cells itself) Neumann or
    Dirichlet on both sides with values that do not depend on time."""
    from pyshocks_amd.scalar import PeriodicBoundary

    if not isinstance(source, BoundOperator):
        return None
    if isinstance(source.bc, PeriodicBoundary):
        return _scheme_desc(source.scheme, source.grid)
    if not fills:
        return None
    fill = _constant_fill(source.bc, source.grid)
    if fill is None:
        return None
    sd = _scheme_desc(source.scheme, source.grid)
    if sd is None:
        return None
    desc, keep = sd
    code, left, right, values = fill
    desc.bc = code
    desc.bc_left, desc.bc_right = left, right
    if values is not None:
        desc.bc_values = values.data_ptr()
        keep = keep + (values,)
    return desc, keep


def _constant_fill(bc: Any, grid: Any) -> tuple[int, float, float, Any] | None:
    """``(PSA_BC_*, left gradient, right gradient, ghost values)`` of a
    two-sided Neumann or Dirichlet boundary whose ``g`` are constants
    (``scalar.py:452-511``), or ``None``."""
    from pyshocks_amd.scalar import (DirichletBoundary, NeumannBoundary, TwoSidedBoundary,
                                     dirichlet_values, neumann_gradient)

    if not isinstance(bc, TwoSidedBoundary):
        return None
    sides = (bc.left, bc.right)
    if any(callable(s.g) for s in sides):
        return None
    if all(isinstance(s, NeumannBoundary) for s in sides):
        return (_lib.PSA_BC_NEUMANN, neumann_gradient(bc.left, 0.0),
                neumann_gradient(bc.right, 0.0), None)
    if all(isinstance(s, DirichletBoundary) for s in sides):
        g = grid.nghosts
        values = torch.cat([dirichlet_values(bc.left, 0.0, grid.x[:g], grid.dtype),
                            dirichlet_values(bc.right, 0.0, grid.x[grid.n - g:],
                                             grid.dtype)]).contiguous()
        return _lib.PSA_BC_DIRICHLET, 0.0, 0.0, values
    return None


def _scheme_desc(scheme: Any, grid: Any) -> tuple[_lib.PsaSolverDesc, tuple[Any, ...]] | None:
    """The library descriptor of a scheme on a grid (any boundary), with the
    tensors it points to, or ``None`` if the library does not run it."""
    from pyshocks_amd import advection, burgers
    from pyshocks_amd.reconstruction import ESWENO32, MUSCL, WENOJS32, WENOJS53
    from pyshocks_amd.scalar import rusanov_nu

    if not isinstance(scheme.rec, (WENOJS53, WENOJS32, ESWENO32, MUSCL)):
        return None

    desc = _lib.PsaSolverDesc()
    desc.dtype = _lib.dtype_code(grid.dtype)
    desc.n = grid.ncells
    desc.g = grid.nghosts
    desc.dx = grid.h
    desc.eps = float(getattr(scheme.rec, "eps", 0.0))
    keep: tuple[Any, ...] = ()
    from pyshocks_amd.reconstruction import LearnedWENOJS53, rec_code

    if not isinstance(scheme.rec, LearnedWENOJS53):
        desc.rec = rec_code(scheme.rec)
    if isinstance(scheme.rec, MUSCL):
        from pyshocks_amd.limiters import limiter_code

        desc.limiter, desc.limiter_theta = limiter_code(scheme.rec.lm)
        desc.muscl_atol = float(scheme.rec.atol)

    if isinstance(scheme.rec, LearnedWENOJS53):
        if not isinstance(scheme, (burgers.Rusanov, burgers.Godunov,
                                   burgers.EngquistOsher)):
            return None
        desc.mlp = scheme.rec.weights.ptr
        desc.mlp_tau = float(scheme.rec.weights.tau)
        desc.mlp_sigma = float(scheme.rec.weights.sigma)
        keep = (scheme.rec.weights,)
    if isinstance(scheme, burgers.Rusanov):
        # and its subclass LaxFriedrichs
        desc.flux = burgers.flux_kind(scheme)
        nu = rusanov_nu(grid, scheme.alpha)
        if nu is not None:
            desc.nu = nu.data_ptr()
            keep = keep + (nu,)
    elif isinstance(scheme, burgers.Godunov):
        desc.flux = _lib.PSA_FLUX_BURGERS_UPWIND
    elif isinstance(scheme, burgers.ESWENO32):
        desc.flux = _lib.PSA_FLUX_ESWENO
        desc.es_delta = float(scheme.rec.delta)
    elif isinstance(scheme, burgers.EngquistOsher):
        desc.flux = _lib.PSA_FLUX_ENGQUIST_OSHER
    elif isinstance(scheme, advection.Godunov):
        vel = scheme.velocity
        if vel is None:
            return None
        _lib.require_device(vel)
        desc.flux = _lib.PSA_FLUX_ADVECTION_UPWIND
        desc.velocity = vel.data_ptr()
        const = bool(torch.all(vel == vel[0]).item())
        desc.velocity_is_const = 1 if const else 0
        desc.velocity_const = float(vel[0].item()) if const else 0.0
        keep = (vel,)
    else:
        return None
    return desc, keep


def fused_plan(stepper: Stepper, *,
               tiles: tuple[int, int] | None = None) -> FusedPlan | None:
    """The fused plan of ``stepper``, or ``None`` for the generic path.

    ``tiles = (threads per workgroup, cells per thread)`` overrides the
    library's tile choice for the stage kernels."""
    if not isinstance(stepper, SSPRK33):
        return None
    sp = _scheme_plan(stepper.source, fills=True)
    if sp is None:
        return None
    desc, keep = sp
    if tiles is None:
        from pyshocks_amd.autotune import lookup_stage

        tiles = lookup_stage(desc)
    if tiles is not None:
        desc.block, desc.cells_per_thread = int(tiles[0]), int(tiles[1])
    if desc.bc != _lib.PSA_BC_PERIODIC and not _uses_step_kernel(desc):
        # the Neumann and Dirichlet fills live in the whole-step kernel only
        return None
    source: BoundOperator = stepper.source  # type: ignore[assignment]
    grid = source.grid
    npt = _np_dtype(grid.dtype)

    predict = stepper.predict_timestep
    from pyshocks_amd import advection, burgers

    if isinstance(predict, CourantTimestep):
        if (predict.scheme is not source.scheme or predict.grid is not grid
                or predict.bc is not source.bc):
            return None
        dx_min = npt(grid.h)
        scheme = source.scheme
        if isinstance(scheme, burgers.Rusanov):
            # burgers/schemes.py:121-127, inherited by LaxFriedrichs
            cfl = npt(0.5 * dx_min ** npt(2 - scheme.alpha))
        elif isinstance(scheme, (burgers.Godunov, burgers.EngquistOsher,
                                 burgers.ESWENO32)):
            # burgers/schemes.py:42-49
            cfl = npt(0.5 * dx_min)
        elif isinstance(scheme, advection.Godunov):
            cfl = dx_min
        else:
            return None
        return FusedPlan(desc=desc, dtype=grid.dtype, device=grid.device,
                         m=grid.n, dt_mode=_lib.PSA_DT_CFL,
                         theta=float(predict.theta), cfl_num=float(cfl),
                         fixed_dt=0.0, keepalive=keep)
    if isinstance(predict, FixedTimestep):
        return FusedPlan(desc=desc, dtype=grid.dtype, device=grid.device,
                         m=grid.n, dt_mode=_lib.PSA_DT_FIXED, theta=1.0,
                         cfl_num=0.0, fixed_dt=float(npt(predict.dt)),
                         keepalive=keep)
    return None


# }}}


# {{{ step


def step(stepper: Stepper, u0: torch.Tensor, *, maxit: int | None = None,
         tstart: float = 0.0, tfinal: float | None = None) -> Iterator[StepCompleted]:
    """Advance to ``tfinal`` (or ``maxit`` steps), yielding a
    :class:`StepCompleted` after every step and once for the initial state."""
    plan = fused_plan(stepper)
    if plan is not None:
        yield from _step_fused(plan, stepper, u0, maxit=maxit, tstart=tstart,
                               tfinal=tfinal)
    else:
        yield from _step_generic(stepper, u0, maxit=maxit, tstart=tstart,
                                 tfinal=tfinal)


def _step_fused(plan: FusedPlan, stepper: Stepper, u0: torch.Tensor, *,
                maxit: int | None, tstart: float,
                tfinal: float | None) -> Iterator[StepCompleted]:
    _lib.require_device(u0)
    if u0.dtype != plan.dtype or u0.numel() != plan.m:
        raise ValueError("u0 does not match the grid of the bound operator")
    npt = plan.np_dtype
    if tfinal is None:
        tfinal = float(np.finfo(npt).max)

    m = 0
    t = npt(tstart)
    tf = npt(tfinal)
    u = u0
    ctrl = plan.new_control(u0, t, tf, maxit)
    k1 = torch.empty_like(u0)
    k2 = torch.empty_like(u0)

    yield StepCompleted(t=t, tfinal=tf, dt=npt(0.0), iteration=m, u=u)

    while True:
        if stepper.checkpoint is not None:
            save(stepper.checkpoint, m, {"m": m, "t": t, "u": u})

        if t >= tf:
            break
        if maxit is not None and m >= maxit:
            break

        u_new = torch.empty_like(u)
        plan.steps(u, u_new, k1, k2, ctrl, 1)
        c = plan.read_control(ctrl)
        if c.status != _lib.PSA_OK:
            raise ValueError(f"Time step is not finite: {c.dt!r}.")
        if not c.active:
            break

        u = u_new
        m += 1
        t = npt(c.t)
        yield StepCompleted(t=t, tfinal=tf, dt=npt(c.dt), iteration=m, u=u)


def _step_generic(stepper: Stepper, u0: torch.Tensor, *, maxit: int | None,
                  tstart: float, tfinal: float | None) -> Iterator[StepCompleted]:
    npt = _np_dtype(u0.dtype)
    if tfinal is None:
        tfinal = float(np.finfo(npt).max)

    m = 0
    t = npt(tstart)
    tf = npt(tfinal)
    u = u0

    yield StepCompleted(t=t, tfinal=tf, dt=npt(0.0), iteration=m, u=u)

    while True:
        if stepper.checkpoint is not None:
            save(stepper.checkpoint, m, {"m": m, "t": t, "u": u})

        if t >= tf:
            break
        if maxit is not None and m >= maxit:
            break

        dt = npt(float(stepper.predict_timestep(t, u)))
        if tf != np.inf:
            dt_min = npt(tf - t)
            dt = npt((dt if dt < dt_min else dt_min) + npt(1.0e-15))

        if not np.isfinite(dt):
            raise ValueError(f"Time step is not finite: {dt!r}.")

        u = advance(stepper, dt, t, u)

        m += 1
        t = npt(t + dt)

        yield StepCompleted(t=t, tfinal=tf, dt=dt, iteration=m, u=u)


def evolve(stepper: Stepper, u0: torch.Tensor, *, maxit: int | None = None,
           tstart: float = 0.0, tfinal: float | None = None,
           chunk: int = 64, tiles: tuple[int, int] | None = None) -> StepCompleted:
    """Run :func:`step` to completion without per-step host synchronisation
    and return the final :class:`StepCompleted`.

    Steps are enqueued ``chunk`` at a time (one chunk of exactly ``maxit``
    steps when only ``maxit`` bounds the run) and the device-side control
    block is read back once per chunk. ``u0`` is not modified. ``tiles``
    as in :func:`fused_plan`.
    """
    plan = fused_plan(stepper, tiles=tiles)
    if plan is None:
        event = None
        for event in step(stepper, u0, maxit=maxit, tstart=tstart, tfinal=tfinal):
            pass
        assert event is not None
        return event

    _lib.require_device(u0)
    npt = plan.np_dtype
    if tfinal is None:
        tfinal = float(np.finfo(npt).max)
    tf = npt(tfinal)
    u = u0.clone()
    k1 = torch.empty_like(u0)
    k2 = torch.empty_like(u0)
    ctrl = plan.new_control(u0, npt(tstart), tf, maxit)

    bounded = maxit is not None and tfinal >= float(np.finfo(npt).max)
    c = plan.read_control(ctrl)
    while True:
        n = int(maxit) - int(c.m) if bounded else chunk
        if maxit is not None:
            n = min(n, int(maxit) - int(c.m))
        if n <= 0:
            break
        plan.steps(u, u, k1, k2, ctrl, n)
        c = plan.read_control(ctrl)
        if c.status != _lib.PSA_OK:
            raise ValueError(f"Time step is not finite: {c.dt!r}.")
        if not c.active:
            break

    return StepCompleted(t=npt(c.t), tfinal=tf, dt=npt(c.dt), iteration=int(c.m), u=u)


# }}}


# {{{ advance


@singledispatch
def advance(stepper: Stepper, dt: Any, t: Any, u: torch.Tensor) -> torch.Tensor:
    """Advance ``u`` from ``t`` to ``t + dt``."""
    raise NotImplementedError(type(stepper).__name__)


@advance.register(ForwardEuler)
def _advance_forward_euler(stepper: ForwardEuler, dt: Any, t: Any,
                           u: torch.Tensor) -> torch.Tensor:
    return u + dt * stepper.source(t, u)


@advance.register(SSPRK33)
def _advance_ssprk33(stepper: SSPRK33, dt: Any, t: Any, u: torch.Tensor) -> torch.Tensor:
    sp = _scheme_plan(stepper.source)
    # Lax-Friedrichs stages need a control block for their speeds: the
    # operator path below computes them per stage
    if sp is not None and sp[0].flux == _lib.PSA_FLUX_LAX_FRIEDRICHS:
        sp = None
    if sp is not None and isinstance(u, torch.Tensor) and u.is_cuda:
        desc, keep = sp
        plan = FusedPlan(desc=desc, dtype=u.dtype, device=u.device, m=u.numel(),
                         dt_mode=_lib.PSA_DT_FIXED, theta=1.0, cfl_num=0.0,
                         fixed_dt=0.0, keepalive=keep)
        dt_dev = torch.as_tensor(dt, dtype=u.dtype, device=u.device).reshape(())
        out = torch.empty_like(u)
        plan.advance(u, out, torch.empty_like(u), torch.empty_like(u),
                     dt_dev.contiguous())
        return out

    fn = stepper.source
    k1 = u + dt * fn(t, u)
    k2 = 3.0 / 4.0 * u + 1.0 / 4.0 * (k1 + dt * fn(t + dt, k1))
    return 1.0 / 3.0 * u + 2.0 / 3.0 * (k2 + dt * fn(t + 0.5 * dt, k2))


# }}}


# {{{ fixed time step


def predict_timestep_from_maxit(tfinal: float, maxit: int) -> tuple[int, float]:
    """``timestepping.py:242-248``."""
    dt = tfinal / maxit + 1.0e-15
    return maxit, dt


def predict_maxit_from_timestep(tfinal: float, dt: float) -> tuple[int, float]:
    """``timestepping.py:251-261``."""
    maxit = int(tfinal / dt)
    dt = tfinal / maxit + 1.0e-15
    return maxit, dt


# }}}


# {{{ adjoint


def adjoint_step(stepper: Stepper, p0: torch.Tensor, *, maxit: int,
                 apply_boundary: Callable[[Any, torch.Tensor, torch.Tensor],
                                          torch.Tensor] | None = None,
                 ) -> Iterator[AdjointStepCompleted]:
    """Discrete adjoint sweep over the checkpoints of a forward run
    (``timestepping.py:155-215``): ``p <- J(u_m)^T p`` for ``m = maxit - 1 ..
    0`` with ``J`` the Jacobian of one SSP-RK3 step at the checkpointed state
    and the recorded ``dt = t_{m+1} - t_m`` held constant. The transposed
    Jacobian is applied matrix-free by the library's reverse stage kernels."""
    from pyshocks_amd.adjoint import ssprk33_step_vjp

    if stepper.checkpoint is None:
        raise ValueError("Adjoint time stepping requires a checkpoint.")
    sp = _scheme_plan(stepper.source)
    if not isinstance(stepper, SSPRK33) or sp is None:
        raise NotImplementedError(
            "adjoint_step needs an SSPRK33 stepper with a BoundOperator source")

    chk = load(stepper.checkpoint, maxit)
    assert chk["m"] == maxit

    t = tfinal = chk["t"]
    p = p0
    if apply_boundary is not None:
        p = apply_boundary(chk["t"], chk["u"], p)

    npt = _np_dtype(p.dtype)
    yield AdjointStepCompleted(t=t, tfinal=tfinal, dt=npt(0.0), iteration=maxit,
                               u=chk["u"], p=p)

    for m in range(maxit - 1, -1, -1):
        chk = load(stepper.checkpoint, m)
        dt = npt(t - chk["t"])
        assert chk["m"] == m

        p = ssprk33_step_vjp(stepper.source, chk["u"], p, dt)

        if apply_boundary is not None:
            p = apply_boundary(chk["t"], chk["u"], p)

        t = chk["t"]
        yield AdjointStepCompleted(t=t, tfinal=tfinal, dt=dt, iteration=m,
                                   u=chk["u"], p=p)


# }}}
