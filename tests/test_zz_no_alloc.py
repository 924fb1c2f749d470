"""psa_ssprk33_solve_static (no allocation, static device scratch) against
psa_control_init + psa_ssprk33_steps (caller buffers) on the same inputs.

Both run the same kernels, so the states must agree bit for bit, and the
result record must carry the control block's t, dt, step count and status.
Calls are made back to back on one GPU, the later ones on smaller arrays, so
that the static array holds the previous call's contents when the next one
starts.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import exact as E
from pyshocks_amd import _lib, make_uniform_cell_grid
from pyshocks_amd.burgers import LaxFriedrichs, Rusanov
from pyshocks_amd.reconstruction import WENOJS53
from pyshocks_amd.scalar import PeriodicBoundary
from pyshocks_amd.timestepping import fused_plan, make_ssprk33

pytestmark = pytest.mark.gpu


def _plan(scheme, n: int, dtype: torch.dtype):
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype, device="cuda")
    stepper = make_ssprk33(scheme, grid, PeriodicBoundary(), theta=1.0)
    plan = fused_plan(stepper)
    assert plan is not None
    return plan


def _static(plan, u0: torch.Tensor, tfinal: float, nsteps: int):
    out = torch.empty_like(u0)
    res = torch.zeros(ctypes.sizeof(_lib.PsaStaticResult), dtype=torch.uint8,
                      device=u0.device)
    _lib.call("psa_ssprk33_solve_static", ctypes.byref(plan.desc), _lib.ptr(u0),
              _lib.ptr(out), 0.0, float(tfinal), -1, plan.dt_mode, plan.theta,
              plan.cfl_num, plan.fixed_dt, int(nsteps), _lib.ptr(res),
              _lib.stream_of(u0))
    torch.cuda.synchronize()
    rec = _lib.PsaStaticResult.from_buffer_copy(res.cpu().numpy().tobytes())
    return out, rec


def _buffers(plan, u0: torch.Tensor, tfinal: float, nsteps: int):
    out = torch.empty_like(u0)
    k1 = torch.empty_like(u0)
    k2 = torch.empty_like(u0)
    ctrl = plan.new_control(u0, 0.0, tfinal, None)
    plan.steps(u0, out, k1, k2, ctrl, nsteps)
    torch.cuda.synchronize()
    return out, plan.read_control(ctrl)


CASES = [
    # (scheme, n, dtype, tfinal, nsteps): whole-step kernel, then the stage
    # kernels with their Lax-Friedrichs slots, then smaller arrays
    (Rusanov(rec=WENOJS53(eps=1e-6)), 8192, torch.float64, 0.3, 8000),
    (LaxFriedrichs(rec=WENOJS53(eps=1e-6)), 3000, torch.float64, 0.3, 3000),
    (Rusanov(rec=WENOJS53(eps=1e-6)), 1000, torch.float32, 0.5, 2000),
    (Rusanov(rec=WENOJS53(eps=1e-6)), 256, torch.float64, 0.5, 600),
]


def test_static_solve_matches_caller_buffers() -> None:
    for scheme, n, dtype, tfinal, nsteps in CASES:
        print(f"[no_alloc] {type(scheme).__name__} n={n} {dtype} t={tfinal}", flush=True)
        plan = _plan(scheme, n, dtype)
        u0 = torch.tensor(E.exact_cell_averages(E.c1_initial_condition(), n),
                          dtype=dtype, device="cuda")
        ref, ctrl = _buffers(plan, u0, tfinal, nsteps)
        got, rec = _static(plan, u0, tfinal, nsteps)
        assert ctrl.status == _lib.PSA_OK and rec.status == _lib.PSA_OK
        assert ctrl.m > 10 and ctrl.m < nsteps, (ctrl.m, nsteps)
        assert rec.m == ctrl.m
        assert rec.t == ctrl.t and rec.dt == ctrl.dt
        a = ref.cpu().numpy()
        b = got.cpu().numpy()
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b), float(np.max(np.abs(a - b)))


def test_static_solve_in_place_and_zero_steps() -> None:
    plan = _plan(Rusanov(rec=WENOJS53(eps=1e-6)), 512, torch.float64)
    u0 = torch.tensor(E.exact_cell_averages(E.c1_initial_condition(), 512),
                      device="cuda")
    ref, ctrl = _buffers(plan, u0, 0.2, 400)
    u = u0.clone()
    res = torch.zeros(ctypes.sizeof(_lib.PsaStaticResult), dtype=torch.uint8,
                      device="cuda")
    _lib.call("psa_ssprk33_solve_static", ctypes.byref(plan.desc), _lib.ptr(u),
              _lib.ptr(u), 0.0, 0.2, -1, plan.dt_mode, plan.theta, plan.cfl_num,
              plan.fixed_dt, 400, _lib.ptr(res), _lib.stream_of(u))
    torch.cuda.synchronize()
    assert torch.equal(u, ref)
    # no steps: the output is the input, the record says t = tstart, m = 0
    out, rec = _static(plan, u0, 0.2, 0)
    assert torch.equal(out, u0)
    assert rec.m == 0 and rec.t == 0.0 and rec.status == _lib.PSA_OK
