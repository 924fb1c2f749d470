"""Dirichlet and Neumann ghost fills on the GPU against the oracle.

``scalar.py:333-511``: the fills themselves (either side, both), the
operator of a state with such a fill (the ghost cells taken as given, every
reconstruction and flux), a solve through the reference's host loop
(time-dependent data) and through the fused whole-step kernel (constant
gradients or values, filled by the kernel) against the oracle's solve. FP64
to 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL64 = 1.0e-12


@pytest.fixture(scope="module")
def torch_mod():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _dev(torch, a):
    return torch.tensor(np.asarray(a), device="cuda", dtype=torch.float64)


def _host(t) -> np.ndarray:
    return t.detach().double().cpu().numpy()


def _relmax(a, b) -> float:
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _bcs(kind: str):
    """(GPU boundary, oracle fill) pairs."""
    from oracle import weno_np as W
    from pyshocks_amd.scalar import make_dirichlet_boundary, make_neumann_boundary

    if kind == "dirichlet":
        ga = (lambda t, x: 0.5 + 0.25 * t + 0 * x)
        gb = (lambda t, x: -0.25 + 0 * x)
        bc = make_dirichlet_boundary(ga, gb)

        def fill(g, t, v):
            return W.apply_boundary_dirichlet(g, W.apply_boundary_dirichlet(g, v, t, -1, ga),
                                              t, +1, gb)
        return bc, fill
    gl = (lambda t: 0.3)
    gr = (lambda t: -0.2 + t)
    bc = make_neumann_boundary(gl, gr)

    def fill(g, t, v):
        return W.apply_boundary_neumann(g, W.apply_boundary_neumann(g, v, t, -1, gl), t,
                                        +1, gr)
    return bc, fill


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
def test_fills_vs_oracle(torch_mod, kind: str) -> None:
    torch = torch_mod
    from pyshocks_amd import apply_boundary, make_uniform_cell_grid
    from oracle import weno_np as W

    n = 40
    grid = make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    u = np.random.default_rng(2).uniform(-1, 1, size=n + 6)
    bc, fill = _bcs(kind)
    got = _host(apply_boundary(bc, grid, 0.4, _dev(torch, u)))
    ref = fill(ogrid, 0.4, u)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
@pytest.mark.parametrize(("flux", "rec"), [("rusanov", "wenojs53"), ("lf", "wenojs53"),
                                           ("eo", "wenojs32"), ("godunov", "esweno32"),
                                           ("rusanov", "muscl-minmod-1.0"),
                                           ("advection", "wenojs53")])
@pytest.mark.parametrize("n", [7, 1000])
def test_operator_with_boundary(torch_mod, kind: str, flux: str, rec: str, n: int) -> None:
    """Every entry of the operator of the filled state (``schemes.py:339-346``
    with the ghost cells as the fill leaves them)."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import advection, apply_operator, burgers, make_uniform_cell_grid
    from pyshocks_amd.limiters import make_limiter_from_name
    from pyshocks_amd.reconstruction import make_reconstruction_from_name

    g = 3
    grid = make_uniform_cell_grid(-1.0, 1.0, n, nghosts=g)
    ogrid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=g)
    rng = np.random.default_rng(n + len(flux))
    u = rng.uniform(-1, 1, size=n + 2 * g)
    vel = 0.5 + 0.25 * np.sin(np.pi * np.arange(n + 2 * g) / n)
    if rec.startswith("muscl"):
        r = make_reconstruction_from_name("muscl", lm=make_limiter_from_name("minmod",
                                                                             theta=1.0))
    else:
        r = make_reconstruction_from_name(rec, eps=1e-6)
    cls = {"rusanov": burgers.Rusanov, "lf": burgers.LaxFriedrichs,
           "eo": burgers.EngquistOsher, "godunov": burgers.Godunov}
    sch = (advection.Godunov(rec=r, velocity=_dev(torch, vel)) if flux == "advection"
           else cls[flux](rec=r))
    bc, fill = _bcs(kind)
    L = _host(apply_operator(sch, grid, bc, 0.2, _dev(torch, u)))
    osch = W.Scheme(flux, eps=0.0 if rec.startswith("muscl") else 1e-6, velocity=vel,
                    rec=rec)
    Lref = W.apply_operator(osch, ogrid, 0.2, u, bc=fill)
    assert _relmax(L, Lref) < RTOL64


def test_solve_with_dirichlet_boundary(torch_mod) -> None:
    """A Burgers Riemann problem with Dirichlet data through the reference's
    host loop (the stepper has no fused plan), against the oracle's solve."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.burgers import Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33

    n = 200
    grid = make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    ogrid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    bc, fill = _bcs("dirichlet")
    x = ogrid.x
    u0 = np.where(x < 0.0, 0.5, -0.25)
    st = make_ssprk33(Rusanov(rec=WENOJS53(eps=1e-6)), grid, bc, theta=1.0)
    assert fused_plan(st) is None
    ev = evolve(st, _dev(torch, u0), tfinal=0.3)
    uref, dts, t = W.solve(W.Scheme("rusanov", eps=1e-6), ogrid, u0, theta=1.0,
                           tfinal=0.3, bc=fill)
    assert ev.iteration == len(dts)
    i = np.s_[3:-3]
    assert _relmax(_host(ev.u)[i], uref[i]) < RTOL64


def _constant_bcs(kind: str):
    """(GPU boundary with constant g, its callable twin, oracle fill)."""
    from oracle import weno_np as W
    from pyshocks_amd.scalar import make_dirichlet_boundary, make_neumann_boundary

    if kind == "dirichlet":
        a, b = 0.5, -0.25
        bc = make_dirichlet_boundary(a, b)
        twin = make_dirichlet_boundary(lambda t, x: a + 0 * x, lambda t, x: b + 0 * x)

        def fill(g, t, v):
            return W.apply_boundary_dirichlet(
                g, W.apply_boundary_dirichlet(g, v, t, -1, lambda t, x: a + 0 * x), t, +1,
                lambda t, x: b + 0 * x)
        return bc, twin, fill
    a, b = 0.3, -0.2
    bc = make_neumann_boundary(a, b)
    twin = make_neumann_boundary(lambda t: a, lambda t: b)

    def fill(g, t, v):
        return W.apply_boundary_neumann(g, W.apply_boundary_neumann(g, v, t, -1, lambda t: a),
                                        t, +1, lambda t: b)
    return bc, twin, fill


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
@pytest.mark.parametrize("flux", ["rusanov", "godunov"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize(("n", "maxit", "tfinal"), [(200, None, 0.3), (5000, 25, None)])
def test_fused_steps_with_constant_fill(torch_mod, kind: str, flux: str, dtype_name: str,
                                        n: int, maxit, tfinal) -> None:
    """Constant Neumann gradients or Dirichlet values: the whole-step kernel
    fills the ghost cells of every stage input itself, so ``step`` stays on
    the device; against the oracle's solve with the same fill (FP64 1e-12,
    FP32 1e-5 of the FP64 oracle), every dt and the step count. N = 5000 has
    many tiles between the two boundary tiles."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.burgers import Godunov, Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33, step

    dtype = getattr(torch, dtype_name)
    grid = make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3, dtype=dtype)
    ogrid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    bc, _, fill = _constant_bcs(kind)
    x = ogrid.x
    u0 = np.where(x < 0.0, 0.5, -0.25) + 0.2 * np.sin(3 * np.pi * x)
    sch = (Rusanov if flux == "rusanov" else Godunov)(rec=WENOJS53(eps=1e-6))
    st = make_ssprk33(sch, grid, bc, theta=1.0)
    plan = fused_plan(st)
    assert plan is not None and plan.uses_step_kernel()
    uref, dts, t = W.solve(W.Scheme(flux, eps=1e-6), ogrid, u0, theta=1.0, tfinal=tfinal,
                           maxit=maxit, bc=fill)
    events = list(step(st, torch.tensor(u0, dtype=dtype, device="cuda"), maxit=maxit,
                       tfinal=tfinal))
    ev = events[-1]
    assert ev.iteration == len(dts)
    got = _host(ev.u)
    i = np.s_[3:-3]
    if dtype == torch.float64:
        np.testing.assert_allclose([float(e.dt) for e in events[1:]], dts, rtol=1e-12)
        assert _relmax(got[i], uref[i]) < RTOL64
    else:
        assert float(np.linalg.norm(got[i] - uref[i]) / np.linalg.norm(uref[i])) < 1e-5
    # the returned state carries the fill of its own interior in its ghost cells
    want = fill(ogrid, float(ev.t), got)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14 if dtype == torch.float64
                               else 1e-6)
    # evolve (one call of many steps) ends in the same state
    ev2 = evolve(st, torch.tensor(u0, dtype=dtype, device="cuda"), maxit=maxit,
                 tfinal=tfinal)
    assert ev2.iteration == ev.iteration
    assert torch.equal(ev2.u, ev.u)


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
def test_fused_fill_equals_host_loop(torch_mod, kind: str) -> None:
    """The same constant boundary given as callables takes the reference's
    host loop (three operator calls per step, the fill kernels before each):
    the two paths agree to rounding, step for step."""
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.burgers import Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.timestepping import evolve, fused_plan, make_ssprk33

    n = 600
    grid = make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    bc, twin, _ = _constant_bcs(kind)
    x = _host(grid.x)
    u0 = _dev(torch, np.where(x < 0.0, 0.5, -0.25) + 0.2 * np.cos(2 * np.pi * x))
    sch = Rusanov(rec=WENOJS53(eps=1e-6))
    fused = make_ssprk33(sch, grid, bc, theta=1.0)
    host = make_ssprk33(sch, grid, twin, theta=1.0)
    assert fused_plan(fused) is not None
    assert fused_plan(host) is None
    a = evolve(fused, u0, tfinal=0.2)
    b = evolve(host, u0, tfinal=0.2)
    assert a.iteration == b.iteration
    assert _relmax(_host(a.u)[3:-3], _host(b.u)[3:-3]) < RTOL64
