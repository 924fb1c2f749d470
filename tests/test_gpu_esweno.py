"""The ES-WENO scheme for Burgers (``burgers/schemes.py:202-256``) on the
GPU against the oracle: the bound parameters, the dissipative flux of either
weight, face fluxes and operators, solves through a shock by the fused
stages and steps (every schedule; the whole-step kernel equal to the stage
kernels bit for bit) and by the generic step loop over the operator. FP64
to 1e-12 relative, FP32 to 1e-5 (l2) against FP64."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _host(t) -> np.ndarray:
    return t.detach().double().cpu().numpy()


def _relmax(a, b) -> float:
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _bound(torch, n: int, dtype=None):
    from pyshocks_amd import bind, make_uniform_cell_grid
    from pyshocks_amd.burgers import ESWENO32
    from pyshocks_amd.reconstruction import ESWENO32 as ESWENO32Rec
    from pyshocks_amd.scalar import PeriodicBoundary

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=2, dtype=dtype or torch.float64)
    bc = PeriodicBoundary()
    return bind(ESWENO32(rec=ESWENO32Rec()), grid, bc), grid, bc


def test_bind_parameters(torch_mod) -> None:
    from oracle import weno_np as W

    sch, _, _ = _bound(torch_mod, 300)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, 300, nghosts=2)
    eps, delta = W.es_weno_parameters(ogrid, np.ones_like(ogrid.x))
    assert abs(sch.rec.eps - eps) < 1e-14 * eps and sch.rec.delta == delta


@pytest.mark.parametrize("weight", [0, 1])
@pytest.mark.parametrize("n", [3, 1000])
def test_dissipation_vs_oracle(torch_mod, weight: int, n: int) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd.burgers import esweno_dissipation
    from pyshocks_amd.reconstruction import ESWENO32

    rng = np.random.default_rng(n + weight)
    u = rng.uniform(-1, 1, size=n)
    f = rng.uniform(-1, 1, size=n + 1)
    rec = ESWENO32(eps=1e-4, delta=1e-3)
    got = _host(esweno_dissipation(rec, torch.tensor(u, device="cuda"),
                                   torch.tensor(f, device="cuda"), weight))
    om = W.es_weno_weights(W.WENOJS32, u, eps=1e-4)[weight, :]
    ref = f + W.esweno_dissipation(u, om, 1e-3)
    assert _relmax(got, ref) < 1e-13


@pytest.mark.parametrize("n", [7, 4096, 70001])
def test_flux_and_operator_vs_oracle(torch_mod, n: int) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import apply_operator, numerical_flux

    sch, grid, bc = _bound(torch, n)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    osch = W.Scheme("esweno32", eps=sch.rec.eps, rec="esweno32", delta=sch.rec.delta)
    u = np.random.default_rng(n).uniform(-1, 1, size=n + 4)
    ud = torch.tensor(u, device="cuda")
    F = _host(numerical_flux(sch, grid, bc, 0.0, ud))
    assert _relmax(F, W.numerical_flux(osch, ogrid, u)) < 1e-12
    L = _host(apply_operator(sch, grid, bc, 0.0, ud))
    Lr = W.apply_operator(osch, ogrid, 0.0, u)
    assert _relmax(L[2:-2], Lr[2:-2]) < 1e-12


@pytest.mark.parametrize(("dtype_name", "tfinal"), [("float64", 0.3), ("float32", 0.1)])
def test_solve_through_shock(torch_mod, dtype_name: str, tfinal: float) -> None:
    """FP64 through the shock (it forms at t = 1 / (2 pi)); FP32 before it,
    where the 1e-5 bar against FP64 measures the arithmetic (past the
    shock, FP32 time steps move it by O(dt) and the l2 difference reaches
    2e-5)."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33

    dtype = getattr(torch, dtype_name)
    n = 256
    sch, grid, bc = _bound(torch, n, dtype)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    x = (np.arange(n + 4) - 2 + 0.5) / n
    u0 = W.apply_boundary_periodic(ogrid, 0.5 + np.sin(2 * np.pi * x))
    osch = W.Scheme("esweno32", eps=sch.rec.eps, rec="esweno32", delta=sch.rec.delta)
    uref, dts, tref = W.solve(osch, ogrid, u0, theta=1.0, tfinal=tfinal)
    st = make_ssprk33(sch, grid, bc, theta=1.0)
    assert fused_plan(st).uses_ring_kernel()
    from pyshocks_amd import apply_operator, predict_timestep
    from pyshocks_amd.timestepping import SSPRK33

    # the generic loop over the operator kernels (a plain callable source)
    generic = SSPRK33(
        predict_timestep=lambda t, u: 1.0 * predict_timestep(sch, grid, bc, t, u),
        source=lambda t, u: apply_operator(sch, grid, bc, t, u), checkpoint=None)
    for stepper in (st, generic):
        ev = evolve(stepper, torch.tensor(u0, device="cuda", dtype=dtype), tfinal=tfinal)
        got = _host(ev.u)[2:-2]
        rel = np.linalg.norm(got - uref[2:-2]) / np.linalg.norm(uref[2:-2])
        if dtype == torch.float64:
            assert ev.iteration == len(dts) and abs(float(ev.t) - tref) < 1e-12
            assert rel < 1e-12
        else:
            assert rel < 1e-5


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("n", [3, 1013, 100_003])
def test_step_schedule_bit_for_bit(torch_mod, dtype_name: str, n: int) -> None:
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33

    dtype = getattr(torch, dtype_name)
    sch, grid, bc = _bound(torch, n, dtype)
    st = make_ssprk33(sch, grid, bc, theta=1.0)
    avg = E.exact_cell_averages(E.large_initial_condition(seed=9), n, nghosts=2)
    out = []
    for sched in (_lib.PSA_SCHEDULE_TILE, _lib.PSA_SCHEDULE_STEP):
        plan = fused_plan(st, tiles=(256, 4))
        plan.desc.schedule = sched
        u = torch.tensor(avg, device="cuda", dtype=dtype)
        v = torch.empty_like(u)
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, float(np.finfo(plan.np_dtype).max), None)
        plan.steps(u, v, k1, k2, ctrl, 3)
        c = plan.read_control(ctrl)
        assert c.m == 3 and c.status == 0
        out.append((v, c.t, c.dt))
    assert out[0][1:] == out[1][1:]
    assert torch.equal(out[0][0], out[1][0])


def test_scheme_requires_es_reconstruction(torch_mod) -> None:
    from pyshocks_amd.burgers import ESWENO32, make_scheme_from_name
    from pyshocks_amd.reconstruction import ESWENO32 as ESWENO32Rec
    from pyshocks_amd.reconstruction import WENOJS32

    with pytest.raises(TypeError):
        ESWENO32(rec=WENOJS32())
    assert isinstance(make_scheme_from_name("esweno32", rec=ESWENO32Rec()), ESWENO32)


def _advection(torch, n: int, vel: np.ndarray):
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.advection import ESWENO32
    from pyshocks_amd.reconstruction import ESWENO32 as ESWENO32Rec
    from pyshocks_amd.scalar import PeriodicBoundary

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    rec = ESWENO32Rec(eps=1.0 / n**2, delta=1.0 / n**2)
    sch = ESWENO32(rec=rec, velocity=torch.tensor(vel, device="cuda"))
    return sch, grid, PeriodicBoundary()


@pytest.mark.parametrize("n", [7, 4096, 70001])
def test_advection_flux_and_operator_vs_oracle(torch_mod, n: int) -> None:
    """advection/schemes.py:135-186 with a variable velocity."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import apply_operator, numerical_flux

    x = (np.arange(n + 4) - 2 + 0.5) / n
    vel = 0.5 + 0.25 * np.sin(2 * np.pi * x)
    sch, grid, bc = _advection(torch, n, vel)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    osch = W.Scheme("advection-esweno32", eps=sch.rec.eps, velocity=vel, rec="esweno32",
                    delta=sch.rec.delta)
    u = np.random.default_rng(n).uniform(-1, 1, size=n + 4)
    ud = torch.tensor(u, device="cuda")
    F = _host(numerical_flux(sch, grid, bc, 0.0, ud))
    assert _relmax(F, W.numerical_flux(osch, ogrid, u)) < 1e-12
    L = _host(apply_operator(sch, grid, bc, 0.0, ud))
    assert _relmax(L[2:-2], W.apply_operator(osch, ogrid, 0.0, u)[2:-2]) < 1e-12


def test_advection_solve_vs_oracle(torch_mod) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33

    n = 512
    x = (np.arange(n + 4) - 2 + 0.5) / n
    vel = 0.5 + 0.25 * np.sin(2 * np.pi * x)
    sch, grid, bc = _advection(torch, n, vel)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    u0 = np.random.default_rng(2).uniform(-1, 1, size=n + 4)
    osch = W.Scheme("advection-esweno32", eps=sch.rec.eps, velocity=vel, rec="esweno32",
                    delta=sch.rec.delta)
    uref, dts, _ = W.solve(osch, ogrid, u0, theta=0.5, maxit=40)
    st = make_ssprk33(sch, grid, bc, theta=0.5)
    assert fused_plan(st) is None
    ev = evolve(st, torch.tensor(u0, device="cuda"), maxit=40)
    got = _host(ev.u)[2:-2]
    assert ev.iteration == len(dts) == 40
    assert np.linalg.norm(got - uref[2:-2]) / np.linalg.norm(uref[2:-2]) < 1e-12


@pytest.mark.parametrize(("scale", "eps"), [(1.0e10, None), (1.0, 1.0e-20), (1.0, 1.0e-23)])
def test_fused_fp32_weight_scale_free(torch_mod, scale: float, eps) -> None:
    """The FP32 ES-WENO weight of the fused kernels stays finite and within
    1e-5 of the FP64 oracle's operator for states of size 1e10 (whose
    differences, raised to the fourth power, overflow FP32) and for an eps
    far below the squared differences (1e-20, 1e-23)."""
    torch = torch_mod
    from dataclasses import replace

    from oracle import exact as E
    from oracle import weno_np as W
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33

    n = 1000
    sch, grid, bc = _bound(torch, n, torch.float32)
    if eps is not None:
        sch = replace(sch, rec=replace(sch.rec, eps=eps))
    st = make_ssprk33(sch, grid, bc, theta=1.0)
    plan = fused_plan(st, tiles=(256, 4))
    assert plan is not None
    u = scale * E.exact_cell_averages(E.large_initial_condition(seed=4), n, nghosts=2)
    u32 = u.astype(np.float32)
    ud = torch.tensor(u32, device="cuda")
    L = torch.empty_like(ud)
    plan.stage(0, None, ud, L)
    got = _host(L)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    osch = W.Scheme("esweno32", eps=sch.rec.eps, rec="esweno32", delta=sch.rec.delta)
    ref = W.apply_operator(osch, ogrid, 0.0, u32.astype(np.float64))
    assert np.all(np.isfinite(got))
    assert _relmax(got[2:-2], ref[2:-2]) < 1e-5
