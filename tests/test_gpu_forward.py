"""GPU parity of the forward hot path against the CPU oracle and the golden
fixtures. Tolerances: 1e-12 relative in FP64, 1e-5 relative (interior l2
against the FP64 oracle) in FP32."""

from __future__ import annotations

import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

RTOL64 = 1.0e-12
RTOL32 = 1.0e-5


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _dev(torch, a: np.ndarray, dtype=None):
    dtype = dtype or torch.float64
    return torch.tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _host(t) -> np.ndarray:
    return t.detach().double().cpu().numpy()


def _rel(a: np.ndarray, b: np.ndarray) -> float:
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def _relmax(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# {{{ weno.py kernels


@pytest.mark.parametrize("name", ["sine", "random"])
def test_weno_fields(torch_mod, name: str) -> None:
    torch = torch_mod
    from pyshocks_amd import weno

    g = np.load(GOLDEN / "weno_fields.npz")
    u = _dev(torch, g[f"u_{name}"])
    s = weno.weno_js_53_coefficients()

    beta = _host(weno.weno_smoothness(s, u))
    uhat = _host(weno.weno_interp(s, u))
    assert _relmax(beta, g[f"beta_{name}"]) < RTOL64
    assert _relmax(uhat, g[f"uhat_{name}"]) < RTOL64
    for eps in (6, 12):
        omega = _host(weno.weno_js_weights(s, u, eps=10.0**-eps))
        assert _relmax(omega, g[f"omega_{name}_eps{eps}"]) < RTOL64


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("name", ["sine", "random"])
def test_reconstruct(torch_mod, name: str, dtype_name: str) -> None:
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.reconstruction import WENOJS53, reconstruct
    from pyshocks_amd.schemes import BoundaryType

    dtype = getattr(torch, dtype_name)
    g = np.load(GOLDEN / "weno_fields.npz")
    u = _dev(torch, g[f"u_{name}"], dtype)
    grid = make_uniform_cell_grid(0.0, 1.0, 64, nghosts=3, dtype=dtype)
    for eps in (6, 12):
        rec = WENOJS53(eps=10.0**-eps)
        ul, ur = reconstruct(rec, grid, BoundaryType.Periodic, u, u, u)
        tol = RTOL64 if dtype == torch.float64 else RTOL32
        assert _rel(_host(ul), g[f"ul_{name}_eps{eps}"]) < tol
        assert _rel(_host(ur), g[f"ur_{name}_eps{eps}"]) < tol


# }}}


# {{{ fluxes and operators


def _scheme(kind: str, torch, vel, eps=1e-6):
    from pyshocks_amd import advection, burgers
    from pyshocks_amd.reconstruction import WENOJS53

    rec = WENOJS53(eps=eps)
    if kind == "rusanov":
        return burgers.Rusanov(rec=rec)
    if kind == "godunov":
        return burgers.Godunov(rec=rec)
    if kind == "lf":
        return burgers.LaxFriedrichs(rec=rec)
    if kind == "eo":
        return burgers.EngquistOsher(rec=rec)
    return advection.Godunov(rec=rec, velocity=vel)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "advection"])
def test_numerical_flux_and_operator(torch_mod, kind: str) -> None:
    torch = torch_mod
    from pyshocks_amd import apply_boundary, apply_operator, make_uniform_cell_grid, numerical_flux
    from pyshocks_amd.scalar import PeriodicBoundary

    g = np.load(GOLDEN / "operators_n32.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 32, nghosts=3)
    u = _dev(torch, g["u"])
    vel = _dev(torch, g["velocity"])
    bc = PeriodicBoundary()
    sch = _scheme(kind, torch, vel)

    F = _host(numerical_flux(sch, grid, bc, 0.0, u))
    assert _relmax(F, g[f"flux_{kind}"]) < RTOL64

    L = _host(apply_operator(sch, grid, bc, 0.0, u))
    # every entry, ghosts included
    assert _relmax(L, g[f"op_{kind}"]) < RTOL64

    ub = _host(apply_boundary(bc, grid, 0.0, u))
    np.testing.assert_array_equal(ub, g["boundary"])


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo", "advection"])
@pytest.mark.parametrize("n", [7, 1000, 70001])
def test_operator_random_sizes(torch_mod, kind: str, n: int) -> None:
    """Ragged sizes (tile remainders, a tile smaller than the halo)."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import apply_operator, make_uniform_cell_grid
    from pyshocks_amd.scalar import PeriodicBoundary

    rng = np.random.default_rng(n)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = rng.uniform(-1, 1, size=n + 6)
    vel = rng.uniform(-1, 1, size=n + 6)
    sch = _scheme(kind, torch, _dev(torch, vel))
    L = _host(apply_operator(sch, grid, PeriodicBoundary(), 0.0, _dev(torch, u)))
    Lref = W.apply_operator(W.Scheme(kind, eps=1e-6, velocity=vel), ogrid, 0.0, u)
    assert _relmax(L, Lref) < RTOL64


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo", "advection"])
def test_predict_timestep(torch_mod, kind: str) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid, predict_timestep
    from pyshocks_amd.scalar import PeriodicBoundary

    n = 300
    rng = np.random.default_rng(2)
    u = rng.uniform(-2, 2, size=n + 6)
    vel = rng.uniform(-1, 1, size=n + 6)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    sch = _scheme(kind, torch, _dev(torch, vel))
    dt = float(predict_timestep(sch, grid, PeriodicBoundary(), 0.0, _dev(torch, u)))
    ref = float(W.predict_timestep(W.Scheme(kind, eps=1e-6, velocity=vel), ogrid, 0.0, u))
    assert dt == ref


# }}}


# {{{ fused SSP-RK3


def _stepper(torch, kind: str, grid, vel=None, theta=1.0, dt=None, checkpoint=None):
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import make_ssprk33

    sch = _scheme(kind, torch, vel)
    return make_ssprk33(sch, grid, PeriodicBoundary(), theta=theta, dt=dt,
                        checkpoint=checkpoint)


def test_fused_c1_burgers_rusanov(torch_mod) -> None:
    """N = 256, u0 = 0.5 + sin(2 pi x), WENO-JS5 eps = 1e-6, Rusanov, SSP-RK3,
    Courant 0.5 (theta = 1), to t = 0.5 (past the shock)."""
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan, step

    g = np.load(GOLDEN / "burgers_rusanov_c1.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 256, nghosts=3)
    stepper = _stepper(torch, "rusanov", grid)
    assert fused_plan(stepper) is not None

    u0 = _dev(torch, g["u0"])
    states, dts = {}, []
    event = None
    for event in step(stepper, u0, tfinal=0.5):
        if event.iteration in (1, 10):
            states[event.iteration] = _host(event.u)
        if event.iteration > 0:
            dts.append(float(event.dt))
    assert event is not None
    assert len(dts) == len(g["dts"])
    np.testing.assert_allclose(dts[:-1], g["dts"][:-1], rtol=RTOL64, atol=0)
    # the last step is clipped to tfinal - t + 1e-15, where t is a sum of
    # 368 steps that each agree with the oracle's to RTOL64: t (and so the
    # clipped step) agrees to RTOL64 * tfinal
    tol_t = RTOL64 * float(g["tfinal"])
    assert abs(dts[-1] - g["dts"][-1]) < tol_t
    assert abs(float(event.t) - float(g["tfinal"])) < tol_t
    i = np.s_[3:-3]
    for k in (1, 10):
        assert _rel(states[k][i], g[f"u{k}"][i]) < RTOL64
    assert _rel(_host(event.u)[i], g["ufinal"][i]) < RTOL64
    # ghost cells hold the periodic images
    uf = _host(event.u)
    np.testing.assert_array_equal(uf[:3], uf[256:259])
    np.testing.assert_array_equal(uf[-3:], uf[3:6])


def test_fused_burgers_godunov(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import evolve

    g = np.load(GOLDEN / "burgers_godunov_n128.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 128, nghosts=3)
    event = evolve(_stepper(torch, "godunov", grid), _dev(torch, g["u0"]), tfinal=0.5)
    assert event.iteration == len(g["dts"])
    assert _rel(_host(event.u)[3:-3], g["ufinal"][3:-3]) < RTOL64


def test_fused_advection(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import step

    g = np.load(GOLDEN / "advection_n128.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 128, nghosts=3)
    vel = torch.ones(134, dtype=torch.float64, device="cuda")
    event = None
    for event in step(_stepper(torch, "advection", grid, vel=vel, theta=0.5),
                      _dev(torch, g["u0"]), tfinal=0.25):
        pass
    assert event is not None
    assert event.iteration == len(g["dts"])
    assert _rel(_host(event.u)[3:-3], g["ufinal"][3:-3]) < RTOL64


@pytest.mark.parametrize("n", [200, 20000])
def test_fused_advection_velocity_array(torch_mod, n: int) -> None:
    """A non-constant velocity array takes the reconstructed-velocity path
    (n = 20000 also takes the interior-tile fast path)."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import evolve

    rng = np.random.default_rng(11)
    x = (np.arange(n + 6) - 2.5) / n
    vel = 0.5 + 0.25 * np.sin(2 * np.pi * x)
    u0 = rng.uniform(-1, 1, size=n + 6)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ev = evolve(_stepper(torch, "advection", grid, vel=_dev(torch, vel), theta=0.5),
                _dev(torch, u0), maxit=40)
    u, dts, _ = W.solve(W.Scheme("advection", eps=1e-6, velocity=vel), ogrid, u0,
                        theta=0.5, maxit=40)
    assert ev.iteration == 40
    assert _rel(_host(ev.u)[3:-3], u[3:-3]) < RTOL64


@pytest.mark.parametrize(("kind", "tiles"), [
    ("rusanov", (256, 1)), ("rusanov", (256, 2)), ("rusanov", (256, 4)),
    ("godunov", (256, 4)), ("lf", (256, 4)), ("lf", (256, 2)), ("eo", (256, 4)),
    ("eo", (256, 2)),
])
def test_fused_large_sine_step(torch_mod, kind: str, tiles) -> None:
    """N = 2^16, sum of 8 sinusoids plus a step, 100 steps at Courant 0.5,
    for each tile configuration of the stage kernels."""
    torch = torch_mod
    from oracle import exact as E
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import evolve

    n = 1 << 16
    u0 = E.exact_cell_averages(E.large_initial_condition(seed=0), n)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ev = evolve(_stepper(torch, kind, grid), _dev(torch, u0), maxit=100, tiles=tiles)
    u, dts, t = W.solve(W.Scheme(kind, eps=1e-6), ogrid, u0, theta=1.0, maxit=100)
    assert ev.iteration == 100
    assert abs(float(ev.t) - t) <= 1e-12 * t
    assert _rel(_host(ev.u)[3:-3], u[3:-3]) < RTOL64


@pytest.mark.parametrize("kind", ["rusanov", "lf", "eo"])
def test_fused_fp32_vs_fp64_oracle(torch_mod, kind: str) -> None:
    torch = torch_mod
    from oracle import exact as E
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import evolve

    n = 4096
    u0 = E.exact_cell_averages(E.c1_initial_condition(), n)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=torch.float32)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ev = evolve(_stepper(torch, kind, grid), _dev(torch, u0, torch.float32),
                maxit=50)
    u, _, _ = W.solve(W.Scheme(kind, eps=1e-6), ogrid, u0, theta=1.0, maxit=50)
    assert ev.u.dtype == torch.float32
    assert _rel(_host(ev.u)[3:-3], u[3:-3]) < RTOL32


def test_generic_path_matches_fused(torch_mod) -> None:
    """An opaque source callable takes the reference's host loop; it must
    agree with the fused path."""
    torch = torch_mod
    from pyshocks_amd import apply_operator, make_uniform_cell_grid
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import SSPRK33, evolve, fused_plan, step

    g = np.load(GOLDEN / "burgers_rusanov_c1.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 256, nghosts=3)
    fused = _stepper(torch, "rusanov", grid)
    sch = fused.source.scheme
    bc = PeriodicBoundary()
    generic = SSPRK33(predict_timestep=fused.predict_timestep,
                      source=lambda t, u: apply_operator(sch, grid, bc, t, u),
                      checkpoint=None)
    assert fused_plan(generic) is None
    u0 = _dev(torch, g["u0"])
    event = None
    for event in step(generic, u0, tfinal=0.5):
        pass
    ev2 = evolve(fused, u0, tfinal=0.5)
    assert event.iteration == ev2.iteration
    assert _rel(_host(event.u)[3:-3], _host(ev2.u)[3:-3]) < RTOL64


def test_nonfinite_timestep_raises(torch_mod) -> None:
    """``timestepping.py:144-145``: with tfinal = inf no clipping happens and
    a zero state gives dt = inf."""
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import step

    grid = make_uniform_cell_grid(0.0, 1.0, 64, nghosts=3)
    u0 = torch.zeros(70, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="not finite"):
        for _ in step(_stepper(torch, "rusanov", grid), u0, tfinal=float("inf"),
                      maxit=5):
            pass


def test_checkpoint_and_maxit(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.checkpointing import InMemoryCheckpoint, load
    from pyshocks_amd.timestepping import step

    g = np.load(GOLDEN / "burgers_rusanov_c1.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 256, nghosts=3)
    chk = InMemoryCheckpoint(basename="Iteration")
    stepper = _stepper(torch, "rusanov", grid, checkpoint=chk)
    events = list(step(stepper, _dev(torch, g["u0"]), maxit=10))
    assert [e.iteration for e in events] == list(range(11))
    for m in (0, 5, 10):
        c = load(chk, m)
        assert c["m"] == m
    assert _rel(_host(load(chk, 10)["u"])[3:-3], g["u10"][3:-3]) < RTOL64
    with pytest.raises(KeyError):
        load(chk, 11)


def test_advance_matches_step(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.timestepping import advance

    g = np.load(GOLDEN / "burgers_rusanov_c1.npz")
    grid = make_uniform_cell_grid(0.0, 1.0, 256, nghosts=3)
    stepper = _stepper(torch, "rusanov", grid)
    u1 = advance(stepper, float(g["dts"][0]), 0.0, _dev(torch, g["u0"]))
    assert _rel(_host(u1)[3:-3], g["u1"][3:-3]) < RTOL64


# }}}


# {{{ norms


@pytest.mark.parametrize("p", [1, 2, np.inf, "tvd"])
def test_norms(torch_mod, p) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid, norm, rnorm

    n = 5000
    rng = np.random.default_rng(4)
    u = rng.uniform(-1, 1, size=n + 6)
    v = rng.uniform(-1, 1, size=n + 6)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    for weighted in (False, True):
        got = float(norm(grid, _dev(torch, u), p=p, weighted=weighted))
        ref = W.norm(ogrid, u, p=p, weighted=weighted)
        assert abs(got - ref) <= 1e-13 * abs(ref)
    got = float(rnorm(grid, _dev(torch, u), _dev(torch, v), p=p, weighted=True))
    ref = W.rnorm(ogrid, u, v, p=p, weighted=True)
    assert abs(got - ref) <= 1e-13 * abs(ref)


# }}}


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo", "advection"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("tiles", [(256, 4), (256, 2)])
def test_pipelined_schedule_bit_for_bit(torch_mod, kind: str, dtype_name: str, tiles) -> None:
    """The persistent pipelined stage kernels compute exactly what the
    one-workgroup-per-tile kernels compute (partial last tile, several
    steps, grid smaller than the tile count)."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan

    dtype = getattr(torch, dtype_name)
    n = 3_000_017
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    vel = torch.ones(grid.n, dtype=dtype, device="cuda") if kind == "advection" else None
    st = _stepper(torch, kind, grid, vel=vel, theta=0.5 if kind == "advection" else 1.0)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=3), n),
                      device="cuda").to(dtype)
    out = []
    for sched in (_lib.PSA_SCHEDULE_TILE, _lib.PSA_SCHEDULE_PIPELINE):
        plan = fused_plan(st, tiles=tiles)
        plan.desc.schedule = sched
        u = u0.clone()
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, float(np.finfo(plan.np_dtype).max), 5)
        plan.steps(u, u, k1, k2, ctrl, 5)
        c = plan.read_control(ctrl)
        assert c.m == 5 and c.status == 0
        out.append((u, c.t))
    assert out[0][1] == out[1][1]
    assert torch.equal(out[0][0], out[1][0])


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("tiles", [(256, 4), (256, 2)])
@pytest.mark.parametrize("n", [3, 16, 37, 1013, 100_003])
def test_step_schedule_bit_for_bit(torch_mod, kind: str, dtype_name: str, tiles,
                                   n: int) -> None:
    """The whole-step kernel (the three stages of a tile in LDS, the halo
    cells recomputed by the neighbouring tiles) computes exactly what the
    stage kernels compute: the state with its ghost images, t, m and dt, over
    in-place and out-of-place calls, a ragged last tile, a grid smaller than
    one tile (n = 3 and 16: smaller than the 9-cell halo, which then wraps
    around the periodic grid several times), and steps enqueued past tfinal
    (which leave the state as is)."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan

    dtype = getattr(torch, dtype_name)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    vel = torch.ones(grid.n, dtype=dtype, device="cuda") if kind == "advection" else None
    st = _stepper(torch, kind, grid, vel=vel, theta=0.5 if kind == "advection" else 1.0)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=5), n),
                      device="cuda").to(dtype)

    def run(sched: int):
        plan = fused_plan(st, tiles=tiles)
        plan.desc.schedule = sched
        big = float(np.finfo(plan.np_dtype).max)
        u = u0.clone()
        v = torch.empty_like(u)
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, big, None)
        plan.steps(u, u, k1, k2, ctrl, 3)   # in place, several steps
        plan.steps(u, v, k1, k2, ctrl, 1)   # out of place, one step
        plan.steps(v, v, k1, k2, ctrl, 1)   # in place, one step
        plan.steps(v, u, k1, k2, ctrl, 2)   # out of place, two steps
        c = plan.read_control(ctrl)
        assert c.m == 7 and c.status == 0 and c.active == 1
        # a run that ends inside the enqueued steps: two steps, one clipped
        # to tfinal, then three that are not taken
        w = torch.empty_like(u)
        ctrl2 = plan.new_control(u, 0.0, 2.5 * c.dt, None)
        plan.steps(u, w, k1, k2, ctrl2, 6)
        c2 = plan.read_control(ctrl2)
        assert c2.m == 3 and c2.active == 0 and c2.status == 0
        return u, w, (c.t, c.dt, c.m), (c2.t, c2.dt, c2.m)

    a = run(_lib.PSA_SCHEDULE_TILE)
    b = run(_lib.PSA_SCHEDULE_STEP)
    assert a[2] == b[2] and a[3] == b[3]
    assert torch.equal(a[0], b[0])
    assert torch.equal(a[1], b[1])


def _schedule_run(torch, plan, u0, sched: int):
    """The call sequence of test_step_schedule_bit_for_bit on one schedule:
    in place and out of place, one and several steps, then a run whose
    control block stops inside the enqueued steps; the states, (t, dt, m)
    of both control blocks and their max|u| fields."""
    plan.desc.schedule = sched
    big = float(np.finfo(plan.np_dtype).max)
    u = u0.clone()
    v = torch.empty_like(u)
    k1, k2 = torch.empty_like(u), torch.empty_like(u)
    ctrl = plan.new_control(u, 0.0, big, None)
    plan.steps(u, u, k1, k2, ctrl, 3)
    plan.steps(u, v, k1, k2, ctrl, 1)
    plan.steps(v, v, k1, k2, ctrl, 1)
    plan.steps(v, u, k1, k2, ctrl, 2)
    c = plan.read_control(ctrl)
    assert c.m == 7 and c.status == 0 and c.active == 1
    w = torch.empty_like(u)
    ctrl2 = plan.new_control(u, 0.0, 1.5 * c.dt, None)
    plan.steps(u, w, k1, k2, ctrl2, 8)
    c2 = plan.read_control(ctrl2)
    assert 1 <= c2.m < 8 and c2.active == 0 and c2.status == 0
    # a call whose first step is not taken leaves the state as given
    x = torch.full_like(u, 7.0)
    plan.steps(w, x, k1, k2, ctrl2, 2)
    c3 = plan.read_control(ctrl2)
    assert c3.m == c2.m and c3.active == 0
    assert torch.equal(x, w)
    fields = [(cc.t, cc.dt, cc.m, cc.umax_bits, cc.umax_slot[0][0]) for cc in (c, c2, c3)]
    return u, w, fields


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("n", [8, 37, 256, 1000, 1024])
def test_ring_schedule_bit_for_bit(torch_mod, kind: str, dtype_name: str, n: int) -> None:
    """The ring schedule (one workgroup holding the periodic state in LDS for
    every step of a call) computes exactly what the whole-step kernel
    computes: states with their ghost images, t, dt, m and the max|u|
    fields of the control block, over the call sequence of _schedule_run,
    from 8 cells (a stencil that wraps within a thread's neighbours) to
    PSA_RING_MAX_CELLS; the default schedule takes the ring."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan

    dtype = getattr(torch, dtype_name)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    vel = torch.ones(grid.n, dtype=dtype, device="cuda") if kind == "advection" else None
    st = _stepper(torch, kind, grid, vel=vel, theta=0.5 if kind == "advection" else 1.0)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=6), n),
                      device="cuda").to(dtype)
    plan = fused_plan(st)
    assert plan.uses_ring_kernel()
    a = _schedule_run(torch, plan, u0, _lib.PSA_SCHEDULE_STEP)
    b = _schedule_run(torch, plan, u0, _lib.PSA_SCHEDULE_RING)
    d = _schedule_run(torch, plan, u0, _lib.PSA_SCHEDULE_DEFAULT)
    assert a[2] == b[2] == d[2]
    for k in (0, 1):
        assert torch.equal(a[k], b[k]) and torch.equal(b[k], d[k])


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("n", [1500, 20000])
def test_one_cell_per_thread_tiles_bit_for_bit(torch_mod, kind: str, dtype_name: str,
                                               n: int) -> None:
    """Below 2^17 cells the whole-step kernel runs one cell per thread by
    default (tiles 256 x 1); it computes what the 256 x 2 and 256 x 4 tiles
    compute, bit for bit, over the call sequence of _schedule_run (a last
    tile that is partial, a control block that stops inside the call)."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan

    dtype = getattr(torch, dtype_name)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    vel = torch.ones(grid.n, dtype=dtype, device="cuda") if kind == "advection" else None
    st = _stepper(torch, kind, grid, vel=vel, theta=0.5 if kind == "advection" else 1.0)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=8), n),
                      device="cuda").to(dtype)
    runs = []
    for tiles in (None, (256, 1), (256, 2), (256, 4)):
        plan = fused_plan(st, tiles=tiles)
        assert plan.uses_step_kernel()
        runs.append(_schedule_run(torch, plan, u0, _lib.PSA_SCHEDULE_STEP))
    for r in runs[1:]:
        assert r[2] == runs[0][2]
        assert torch.equal(r[0], runs[0][0]) and torch.equal(r[1], runs[0][1])


@pytest.mark.parametrize("rec", ["wenojs32", "esweno32", "muscl", "muscls", "esweno-flux"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_ring_schedule_reconstructions(torch_mod, rec: str, dtype_name: str) -> None:
    """The ring schedule against the whole-step kernel for the third-order,
    MUSCL and slope-limited MUSCL reconstructions (Rusanov) and the ES-WENO
    scheme, bit for bit."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, bind, burgers, make_uniform_cell_grid
    from pyshocks_amd.limiters import make_limiter_from_name
    from pyshocks_amd.reconstruction import ESWENO32, make_reconstruction_from_name
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33

    dtype = getattr(torch, dtype_name)
    n = 300
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=2, dtype=dtype)
    bc = PeriodicBoundary()
    if rec == "esweno-flux":
        sch = bind(burgers.ESWENO32(rec=ESWENO32()), grid, bc)
    elif rec in ("muscl", "muscls"):
        lm = make_limiter_from_name("minmod", theta=1.0, variant=1)
        sch = burgers.Rusanov(rec=make_reconstruction_from_name(rec, lm=lm))
    else:
        sch = burgers.Rusanov(rec=make_reconstruction_from_name(rec, eps=1e-6))
    st = make_ssprk33(sch, grid, bc, theta=1.0)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=8), n, nghosts=2),
                      device="cuda").to(dtype)
    plan = fused_plan(st)
    assert plan.uses_ring_kernel()
    a = _schedule_run(torch, plan, u0, _lib.PSA_SCHEDULE_STEP)
    b = _schedule_run(torch, plan, u0, _lib.PSA_SCHEDULE_RING)
    assert a[2] == b[2]
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_ring_schedule_mixes_with_the_others(torch_mod) -> None:
    """Calls on the ring, whole-step and stage schedules alternate on one
    control block through the end of the run and give what the stage
    schedule gives alone."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan

    n = 512
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    st = _stepper(torch, "rusanov", grid)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=4), n),
                      device="cuda")
    plan = fused_plan(st)
    R, S, T = _lib.PSA_SCHEDULE_RING, _lib.PSA_SCHEDULE_STEP, _lib.PSA_SCHEDULE_TILE
    out = {}
    for name, seq in {"tile": [T] * 5, "mixed": [R, S, T, R, R]}.items():
        u = u0.clone()
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, 17.5 * 0.5 / n / float(u0[3:-3].abs().max()), None)
        for sched in seq:
            plan.desc.schedule = sched
            plan.steps(u, u, k1, k2, ctrl, 4)
        c = plan.read_control(ctrl)
        assert c.status == 0 and c.active == 0
        out[name] = (u, c.t, c.m, c.dt)
    assert out["tile"][1:] == out["mixed"][1:]
    assert torch.equal(out["tile"][0], out["mixed"][0])


def test_schedules_mix_on_one_control_block(torch_mod) -> None:
    """Calls of the whole-step and stage schedules alternate on one control
    block (each leaves t, m, dt and the max|u| slots as the other expects),
    through the end of the run."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.timestepping import fused_plan

    n = 100_003
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    st = _stepper(torch, "rusanov", grid)
    u0 = torch.tensor(E.exact_cell_averages(E.large_initial_condition(seed=2), n),
                      device="cuda")
    plan = fused_plan(st, tiles=(256, 4))
    seqs = {"tile": [_lib.PSA_SCHEDULE_TILE] * 4,
            "mixed": [_lib.PSA_SCHEDULE_STEP, _lib.PSA_SCHEDULE_TILE,
                      _lib.PSA_SCHEDULE_STEP, _lib.PSA_SCHEDULE_TILE]}
    out = {}
    for name, seq in seqs.items():
        u = u0.clone()
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        # about 10 steps of dt ~ 0.5 h / max|u|, then steps that are not taken
        ctrl = plan.new_control(u, 0.0, 10.5 * 0.5 / n / float(u0[3:-3].abs().max()), None)
        for sched in seq:
            plan.desc.schedule = sched
            plan.steps(u, u, k1, k2, ctrl, 4)
        c = plan.read_control(ctrl)
        assert c.status == 0 and c.active == 0
        out[name] = (u, c.t, c.m, c.dt)
    assert out["tile"][1:] == out["mixed"][1:]
    assert torch.equal(out["tile"][0], out["mixed"][0])
