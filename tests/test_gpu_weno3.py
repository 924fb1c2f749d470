"""WENO-JS3 and ES-WENO3 reconstructions on the GPU against the oracle.

``reconstruction.py:320-434`` and ``weno.py:166-203,285-296``: the
reconstruction itself (zero padding, ES-WENO's tau padding at the array
ends), face fluxes and operators of every flux (an advection velocity array
is reconstructed the same way), fused solves through a shock on grids with
2 ghost cells, and the whole-step kernel against the stage kernels bit for
bit. FP64 to 1e-12 relative, FP32 to 1e-5 (interior l2) against FP64.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL64 = 1.0e-12
RTOL32 = 1.0e-5
RECS = ["wenojs32", "esweno32"]


@pytest.fixture(scope="module")
def torch_mod():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _dev(torch, a, dtype=None):
    return torch.tensor(np.asarray(a), device="cuda", dtype=dtype or torch.float64)


def _host(t) -> np.ndarray:
    return t.detach().double().cpu().numpy()


def _rel(a, b) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _relmax(a, b) -> float:
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _rec(name: str):
    from pyshocks_amd.reconstruction import make_reconstruction_from_name

    return make_reconstruction_from_name(name, eps=1e-6)


def _scheme(kind: str, rec: str, vel=None):
    from pyshocks_amd import advection, burgers

    r = _rec(rec)
    cls = {"rusanov": burgers.Rusanov, "godunov": burgers.Godunov,
           "lf": burgers.LaxFriedrichs, "eo": burgers.EngquistOsher}
    if kind == "advection":
        return advection.Godunov(rec=r, velocity=vel)
    return cls[kind](rec=r)


@pytest.mark.parametrize("rec", RECS)
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("n", [3, 64, 1001])
def test_reconstruct_vs_oracle(torch_mod, rec: str, dtype_name: str, n: int) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.reconstruction import reconstruct

    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(n)
    f = rng.uniform(-1, 1, size=n + 4)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=2, dtype=dtype)
    ul, ur = reconstruct(_rec(rec), grid, None, _dev(torch, f, dtype), None, None)
    ulr, urr = W.reconstruct(f, 1e-6, rec)
    if dtype == torch.float64:
        assert _relmax(_host(ul), ulr) < RTOL64 and _relmax(_host(ur), urr) < RTOL64
    else:
        assert _rel(_host(ul), ulr) < RTOL32 and _rel(_host(ur), urr) < RTOL32


@pytest.mark.parametrize("rec", RECS)
@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo", "advection"])
@pytest.mark.parametrize("n", [7, 1000, 70001])
def test_flux_and_operator_vs_oracle(torch_mod, rec: str, kind: str, n: int) -> None:
    """Face fluxes of u as given and the operator of the ghost-filled u (every
    entry), 2 ghost cells; advection with a velocity array."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import apply_operator, make_uniform_cell_grid, numerical_flux
    from pyshocks_amd.scalar import PeriodicBoundary

    rng = np.random.default_rng(n + 7)
    g = 2
    u = rng.uniform(-1, 1, size=n + 2 * g)
    vel = 0.5 + 0.25 * np.sin(2 * np.pi * np.arange(n + 2 * g) / n)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=g)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=g)
    sch = _scheme(kind, rec, _dev(torch, vel))
    osch = W.Scheme(kind, eps=1e-6, velocity=vel, rec=rec)
    F = _host(numerical_flux(sch, grid, PeriodicBoundary(), 0.0, _dev(torch, u)))
    assert _relmax(F, W.numerical_flux(osch, ogrid, u)) < RTOL64
    L = _host(apply_operator(sch, grid, PeriodicBoundary(), 0.0, _dev(torch, u)))
    assert _relmax(L, W.apply_operator(osch, ogrid, 0.0, u)) < RTOL64


@pytest.mark.parametrize("rec", RECS)
@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo"])
def test_fused_burgers_through_shock(torch_mod, rec: str, kind: str) -> None:
    """N = 256, 2 ghost cells, u0 = 0.5 + sin(2 pi x), Courant 0.5, to t = 0.5,
    each step schedule, against the oracle's solve."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33

    n, g = 256, 2
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=g)
    x = (np.arange(n + 2 * g) - g + 0.5) / n
    u0 = W.apply_boundary_periodic(ogrid, 0.5 + np.sin(2 * np.pi * x))
    uref, dts, tref = W.solve(W.Scheme(kind, eps=1e-6, rec=rec), ogrid, u0, theta=1.0,
                              tfinal=0.5)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=g)
    st = make_ssprk33(_scheme(kind, rec), grid, PeriodicBoundary(), theta=1.0)
    assert fused_plan(st).uses_ring_kernel() == (kind != "lf")
    for sched in (_lib.PSA_SCHEDULE_DEFAULT, _lib.PSA_SCHEDULE_TILE,
                  _lib.PSA_SCHEDULE_PIPELINE):
        plan = fused_plan(st)
        plan.desc.schedule = sched
        u = _dev(torch, u0)
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, 0.5, None)
        plan.steps(u, u, k1, k2, ctrl, len(dts) + 2)
        c = plan.read_control(ctrl)
        assert c.status == 0 and c.m == len(dts)
        assert abs(c.t - tref) < RTOL64 * 0.5
        assert _rel(_host(u)[g:-g], uref[g:-g]) < RTOL64
    ev = evolve(st, _dev(torch, u0), tfinal=0.5)
    assert ev.iteration == len(dts)
    assert _rel(_host(ev.u)[g:-g], uref[g:-g]) < RTOL64


@pytest.mark.parametrize("rec", RECS)
@pytest.mark.parametrize("const", [True, False])
def test_fused_advection(torch_mod, rec: str, const: bool) -> None:
    """Advection with a constant velocity (whole-step kernel) or a velocity
    array (stage kernels, the velocity reconstructed by the same method)."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33

    n, g = 5000, 3
    rng = np.random.default_rng(3)
    x = (np.arange(n + 2 * g) - g + 0.5) / n
    vel = np.ones(n + 2 * g) if const else 0.5 + 0.25 * np.sin(2 * np.pi * x)
    u0 = rng.uniform(-1, 1, size=n + 2 * g)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=g)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=g)
    st = make_ssprk33(_scheme("advection", rec, _dev(torch, vel)), grid,
                      PeriodicBoundary(), theta=0.5)
    assert fused_plan(st).uses_step_kernel() == const
    ev = evolve(st, _dev(torch, u0), maxit=40)
    u, _, _ = W.solve(W.Scheme("advection", eps=1e-6, velocity=vel, rec=rec), ogrid, u0,
                      theta=0.5, maxit=40)
    assert ev.iteration == 40
    assert _rel(_host(ev.u)[g:-g], u[g:-g]) < RTOL64


@pytest.mark.parametrize("rec", RECS)
@pytest.mark.parametrize("kind", ["rusanov", "eo"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("tiles", [(256, 4), (256, 2)])
@pytest.mark.parametrize("n", [3, 37, 1013, 100_003])
def test_step_schedule_bit_for_bit(torch_mod, rec: str, kind: str, dtype_name: str,
                                   tiles, n: int) -> None:
    """The whole-step kernel computes what the stage kernels compute, bit for
    bit, with the third-order reconstructions and 2 ghost cells (grids
    smaller than the halo included)."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33

    dtype = getattr(torch, dtype_name)
    g = 2
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=g, dtype=dtype)
    st = make_ssprk33(_scheme(kind, rec), grid, PeriodicBoundary(), theta=1.0)
    avg = E.exact_cell_averages(E.large_initial_condition(seed=6), n)
    u0 = np.concatenate([avg[3 - g:3], avg[3:-3], avg[-3:-3 + g]])
    out = []
    for sched in (_lib.PSA_SCHEDULE_TILE, _lib.PSA_SCHEDULE_STEP):
        plan = fused_plan(st, tiles=tiles)
        plan.desc.schedule = sched
        u = _dev(torch, u0, dtype)
        v = torch.empty_like(u)
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, float(np.finfo(plan.np_dtype).max), None)
        plan.steps(u, v, k1, k2, ctrl, 3)
        c = plan.read_control(ctrl)
        assert c.m == 3 and c.status == 0
        out.append((v, c.t, c.dt))
    assert out[0][1:] == out[1][1:]
    assert torch.equal(out[0][0], out[1][0])


@pytest.mark.parametrize("rec", RECS)
def test_fp32_vs_fp64_oracle(torch_mod, rec: str) -> None:
    torch = torch_mod
    from oracle import exact as E
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import evolve, make_ssprk33

    n = 4096
    u0 = E.exact_cell_averages(E.c1_initial_condition(), n)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=torch.float32)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    st = make_ssprk33(_scheme("rusanov", rec), grid, PeriodicBoundary(), theta=1.0)
    ev = evolve(st, _dev(torch, u0, torch.float32), maxit=50)
    u, _, _ = W.solve(W.Scheme("rusanov", eps=1e-6, rec=rec), ogrid, u0, theta=1.0,
                      maxit=50)
    assert _rel(_host(ev.u)[3:-3], u[3:-3]) < RTOL32


def test_unsupported_combinations(torch_mod) -> None:
    """The adjoint takes WENO-JS5 only; one ghost cell is too few."""
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.adjoint import ssprk33_step_vjp
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import BoundOperator

    grid = make_uniform_cell_grid(0.0, 1.0, 40, nghosts=2)
    u = torch.ones(44, dtype=torch.float64, device="cuda")
    src = BoundOperator(_scheme("rusanov", "wenojs32"), grid, PeriodicBoundary())
    with pytest.raises(NotImplementedError):
        ssprk33_step_vjp(src, u, u, 0.01)
