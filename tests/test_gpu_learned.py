"""The learned-weight solvers on the GPU: the fused operator against the CPU
oracle, the WENO-JS fallback bit for bit, fused against generic stepping,
batched against single-instance. The weights are FP32 (softmax in FP32) and
the hidden layers BF16, so the oracle comparison bounds the bulk of the
cells at the FP32-weight level and allows rare BF16 rounding-boundary flips
(see ``test_gpu_mlp.py``)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_np as M
from oracle import weno_np as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _weights(p, tau):
    from pyshocks_amd.learned import MLPWeights

    return MLPWeights.from_numpy(p, tau)


def _params(seed=12, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(M.NPARAM) * scale).astype(np.float32)


def _u0(n, dtype=np.float64):
    x = (np.arange(n + 6) - 3 + 0.5) / n
    u = 0.5 + np.sin(2 * np.pi * x) + 0.3 * np.sin(10 * np.pi * x + 1.0) + 0.5 * (x > 0.4)
    u[:3] = u[n:n + 3]
    u[-3:] = u[3:6]
    return u.astype(dtype)


def _scheme(kind, rec):
    from pyshocks_amd.burgers import EngquistOsher, Godunov, LaxFriedrichs, Rusanov

    return {"rusanov": Rusanov, "godunov": Godunov, "lf": LaxFriedrichs,
            "eo": EngquistOsher}[kind](rec=rec)


def _stepper(torch, kind, grid, rec):
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import make_ssprk33

    return make_ssprk33(_scheme(kind, rec), grid, PeriodicBoundary(), theta=1.0)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("sigma", [0.0, 1.01])
def test_learned_operator_vs_oracle(torch_mod, kind, dtype_name, sigma):
    torch = torch_mod
    from pyshocks_amd import apply_operator, make_uniform_cell_grid
    from pyshocks_amd.reconstruction import LearnedWENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary

    dtype = getattr(torch, dtype_name)
    n = 3000
    p = _params()
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    from pyshocks_amd.learned import MLPWeights

    rec = LearnedWENOJS53(eps=1e-6, weights=MLPWeights.from_numpy(p, 20.0, sigma))
    sch = _scheme(kind, rec)
    u = _u0(n, np.float64 if dtype_name == "float64" else np.float32)
    L = apply_operator(sch, grid, PeriodicBoundary(), 0.0,
                       torch.tensor(u, device="cuda")).cpu().double().numpy()[3:-3]
    og = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ref = M.learned_apply_operator(p, 20.0, kind, og, u, 1e-6, sigma=sigma)[3:-3]
    if sigma > 0:
        # most cells take the ideal weights, some still the network's
        reg = M.regimes(M.stencils(u), 1e-6, 20.0, sigma)
        assert np.mean(reg == 1) > 0.5 and np.sum(reg == 2) >= 10
    scale = np.max(np.abs(ref))
    d = np.abs(L - ref) / scale
    tol = 1e-5 if dtype_name == "float64" else 2e-4
    assert np.mean(d < tol) > 0.99, np.sort(d)[-10:]
    assert np.max(d) < 5e-2
    # conservation: the two tiles sharing a face compute the same flux
    assert abs(np.sum(L)) * grid.h < (1e-12 if dtype_name == "float64" else 1e-4) * scale


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_fallback_is_wenojs_bit_for_bit(torch_mod, kind, dtype_name):
    """tau = 0 flags every cell with any variation: the learned solver must
    then reproduce the WENO-JS solver exactly (FP64: bit for bit)."""
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.reconstruction import LearnedWENOJS53, WENOJS53
    from pyshocks_amd.timestepping import evolve

    dtype = getattr(torch, dtype_name)
    n = 5000
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    u0 = torch.tensor(_u0(n), device="cuda").to(dtype)
    plain = evolve(_stepper(torch, kind, grid, WENOJS53(eps=1e-6)), u0, maxit=30,
                   tiles=(256, 2))
    learned = evolve(_stepper(torch, kind, grid,
                              LearnedWENOJS53(eps=1e-6, weights=_weights(_params(), 0.0))),
                     u0, maxit=30, tiles=(256, 2))
    assert learned.iteration == plain.iteration == 30
    if dtype_name == "float64":
        assert torch.equal(learned.u[3:-3], plain.u[3:-3])
        assert float(learned.t) == float(plain.t)
    else:
        # the same formulas; in FP32 the compiler contracts some of them into
        # fused multiply-adds differently in the two kernels, a few ulps per
        # step over 30 steps
        a, b = learned.u[3:-3].double(), plain.u[3:-3].double()
        assert float((a - b).abs().max()) <= 1e-5 * float(b.abs().max())


@pytest.mark.parametrize("tiles", [(256, 2), (256, 4)])
def test_learned_fused_vs_generic(torch_mod, tiles):
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.reconstruction import LearnedWENOJS53
    from pyshocks_amd.timestepping import SSPRK33, evolve, step

    n = 2048
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rec = LearnedWENOJS53(eps=1e-6, weights=_weights(_params(), 20.0))
    st = _stepper(torch, "rusanov", grid, rec)
    u0 = torch.tensor(_u0(n), device="cuda")
    fused = evolve(st, u0, maxit=10, tiles=tiles)
    generic = SSPRK33(predict_timestep=lambda t, u: st.predict_timestep(t, u),
                      source=lambda t, u: st.source(t, u), checkpoint=None)
    ev = None
    for ev in step(generic, u0, maxit=10):
        pass
    assert ev is not None and ev.iteration == 10
    a, b = fused.u.cpu().numpy()[3:-3], ev.u.cpu().numpy()[3:-3]
    assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("kind", ["rusanov", "eo"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("tiles", [(256, 2), (256, 4)])
def test_learned_step_schedule_matches_stage_kernels(torch_mod, dtype_name, tiles, kind):
    """The whole-step kernel with learned weights (three stages of a tile in
    LDS) against the three stage kernels, over several tiles and a ragged
    last one."""
    torch = torch_mod
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.reconstruction import LearnedWENOJS53
    from pyshocks_amd.timestepping import fused_plan

    dtype = getattr(torch, dtype_name)
    n = 100_003
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    rec = LearnedWENOJS53(eps=1e-6, weights=_weights(_params(), 20.0))
    st = _stepper(torch, kind, grid, rec)
    u0 = torch.tensor(_u0(n), device="cuda").to(dtype)
    out = []
    for sched in (_lib.PSA_SCHEDULE_TILE, _lib.PSA_SCHEDULE_STEP):
        plan = fused_plan(st, tiles=tiles)
        plan.desc.schedule = sched
        u = u0.clone()
        k1, k2 = torch.empty_like(u), torch.empty_like(u)
        ctrl = plan.new_control(u, 0.0, float(np.finfo(plan.np_dtype).max), None)
        plan.steps(u, u, k1, k2, ctrl, 20)
        c = plan.read_control(ctrl)
        assert c.m == 20 and c.status == 0
        out.append((u, c.t))
    if dtype_name == "float64":
        assert out[0][1] == out[1][1]
        assert torch.equal(out[0][0], out[1][0])
    else:
        a, b = out[1][0].double(), out[0][0].double()
        assert float((a - b).norm()) <= 5e-4 * float(b.norm())


@pytest.mark.parametrize("kind", ["rusanov", "eo"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_batched_learned_equals_single_instance(torch_mod, dtype_name, kind):
    torch = torch_mod
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances
    from pyshocks_amd.reconstruction import LearnedWENOJS53
    from pyshocks_amd.timestepping import evolve

    dtype = getattr(torch, dtype_name)
    w = _weights(_params(), 20.0)
    _, ics = random_instances(4, seed=8)
    ns = np.array([5, 300, 1500, 4100])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    solver = BatchedSolver(ns, dtype=dtype, tiles=(256, 4), grid=2, mlp=w, time_f64=False,
                           team_cells=None, scheme=kind)
    solver.set_states(states)
    res = solver.solve(0.04)
    assert np.all(res.status == 0)
    for i, n in enumerate(ns):
        grid = make_uniform_cell_grid(0.0, 1.0, int(n), nghosts=3, dtype=dtype)
        st = _stepper(torch, kind, grid, LearnedWENOJS53(eps=1e-6, weights=w))
        ev = evolve(st, torch.tensor(states[i], device="cuda").to(dtype), tfinal=0.04,
                    tiles=(256, 4))
        assert res.steps[i] == ev.iteration
        got = solver.state(i).cpu().double().numpy()[3:-3]
        ref = ev.u.cpu().double().numpy()[3:-3]
        # FP32: the two kernels round the state differently in the last ulp
        # (contraction), which moves a few BF16 inputs across a rounding
        # boundary over the ~100 steps
        tol = 1e-10 if dtype == torch.float64 else 5e-4
        assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_batched_learned_step_schedule(torch_mod, dtype_name):
    """The batched solver with learned weights: one sweep per step against
    one sweep per stage."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    dtype = getattr(torch, dtype_name)
    w = _weights(_params(), 20.0)
    _, ics = random_instances(4, seed=8)
    ns = np.array([5, 300, 1500, 4100])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    out = []
    for sched in ("tile", "step"):
        solver = BatchedSolver(ns, dtype=dtype, tiles=(256, 4), mlp=w, schedule=sched)
        solver.set_states(states)
        res = solver.solve(0.04)
        assert np.all(res.status == 0)
        out.append((solver.u.clone(), res))
    np.testing.assert_array_equal(out[0][1].steps, out[1][1].steps)
    if dtype_name == "float64":
        assert torch.equal(out[0][0], out[1][0])
    else:
        a, b = out[1][0].double(), out[0][0].double()
        assert float((a - b).norm()) <= 5e-4 * float(b.norm())


def test_ideal_network_is_fifth_order_on_smooth_data(torch_mod):
    """All-zero parameters and no fallback give the linear 5th-order
    scheme: on smooth advection-free Burgers data before the shock its error
    is below WENO-JS's with eps = 1e-6."""
    torch = torch_mod
    from oracle import exact as E
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.reconstruction import LearnedWENOJS53, WENOJS53
    from pyshocks_amd.timestepping import evolve

    ic = E.SineStepIC(c=0.5, amplitudes=(0.3,), wavenumbers=(1,), phases=(0.0,))
    n, t = 128, 0.3
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u0 = torch.tensor(E.exact_cell_averages(ic, n), device="cuda")
    h = 1.0 / n
    f = h * np.arange(0, n + 1)
    exact = np.diff(E.hopf_lax_primitive(ic, f, t)) / h
    errs = {}
    for name, rec in [("js", WENOJS53(eps=1e-6)),
                      ("ideal", LearnedWENOJS53(eps=1e-6,
                                                weights=_weights(M.ideal_params(), 1e30)))]:
        ev = evolve(_stepper(torch, "rusanov", grid, rec), u0, tfinal=t)
        errs[name] = np.sum(np.abs(ev.u.cpu().numpy()[3:-3] - exact)) * h
    assert errs["ideal"] < errs["js"], errs


def test_training_records(torch_mod):
    torch = torch_mod
    from pyshocks_amd.learned import make_training_samples, training_ics

    ics = training_ics(2, seed=77)
    rec = make_training_samples(ics, n=256, times=(0.02, 0.3), tau=20.0)
    assert rec.shape == (2 * 2 * 2 * 256, 12)
    r = rec.cpu().numpy()
    assert np.all(np.isfinite(r))
    assert np.all(np.abs(r[:, :5]) <= 1.0)
    np.testing.assert_allclose(r[:, 8:11].sum(axis=1), 1.0, atol=1e-6)
    used = r[:, 11] > 0
    assert 0.5 < used.mean() < 1.0  # shock cells are left out
    # on used (smooth) records the ideal combination beats the worst candidate
    ideal = r[used, 5:8] @ np.array([0.1, 0.6, 0.3])
    worst = np.max(np.abs(r[used, 5:8]), axis=1)
    assert np.median(np.abs(ideal) / worst) < 0.2


def test_trained_network_gain(torch_mod):
    """The shipped network against plain WENO-JS and against the untrained
    network (ideal weights with the same shock fallback) on unseen seeded
    instances, FP64, before and after shock formation."""
    torch = torch_mod
    from pyshocks_amd.learned import MLPWeights, l1_errors, training_ics

    w = MLPWeights.load()
    ideal = MLPWeights.from_numpy(np.zeros(M.NPARAM, np.float32), w.tau)
    ics = training_ics(16, seed=3000)
    ns = np.round(np.exp(np.random.default_rng(3000).uniform(np.log(1024), np.log(4096),
                                                             16))).astype(np.int64)
    for t in (0.1, 0.5):
        e_plain, _ = l1_errors(ics, ns, t)
        e_ideal, _ = l1_errors(ics, ns, t, mlp=ideal)
        e_learned, _ = l1_errors(ics, ns, t, mlp=w)
        assert e_learned.mean() < 0.85 * e_plain.mean(), (t, e_learned.mean(), e_plain.mean())
        assert e_learned.mean() < 0.95 * e_ideal.mean(), (t, e_learned.mean(), e_ideal.mean())
        # FP32 state keeps the gain
        e32, _ = l1_errors(ics, ns, t, mlp=w, dtype=torch.float32)
        assert e32.mean() < 0.9 * e_plain.mean()
