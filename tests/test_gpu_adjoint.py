"""GPU parity of the discrete adjoint (reverse stage kernels) against dense
complex-step Jacobians of the CPU oracle, central finite differences of the
GPU forward solve, and torch autograd. Tolerance: 1e-12 relative (FP64) for
the Jacobian products; FD agreement to 1e-4 relative at h = 1e-7."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1.0e-12
KINDS = ["rusanov", "godunov", "eo", "advection"]
#: with Lax-Friedrichs (single instances: the stage-wise reverse with the
#: derivative through its global speed)
KINDS_LF = KINDS + ["lf"]


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _host(t):
    return t.detach().double().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _source(torch, kind: str, n: int, velocity: float = 1.0):
    from pyshocks_amd import advection, burgers, make_uniform_cell_grid
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import BoundOperator

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rec = WENOJS53(eps=1e-6)
    if kind == "rusanov":
        sch = burgers.Rusanov(rec=rec)
    elif kind == "godunov":
        sch = burgers.Godunov(rec=rec)
    elif kind == "eo":
        sch = burgers.EngquistOsher(rec=rec)
    elif kind == "lf":
        sch = burgers.LaxFriedrichs(rec=rec)
    else:
        vel = torch.full((n + 6,), velocity, dtype=torch.float64, device="cuda")
        sch = advection.Godunov(rec=rec, velocity=vel)
    return BoundOperator(sch, grid, PeriodicBoundary()), grid


def _data(n: int, seed: int):
    rng = np.random.default_rng(seed)
    x = (np.arange(n + 6) - 2.5) / n
    u = 0.3 + np.sin(2 * np.pi * x) + 0.2 * rng.standard_normal(n + 6)
    lam = rng.standard_normal(n + 6)
    return u, lam


@pytest.mark.parametrize("kind", KINDS_LF)
def test_operator_vjp_vs_complex_step(torch_mod, kind: str) -> None:
    torch = torch_mod
    from oracle import adjoint_cs as A
    from oracle import weno_np as W
    from pyshocks_amd.adjoint import operator_vjp

    n = 40
    u, mu = _data(n, 1)
    src, _ = _source(torch, kind, n)
    got = _host(operator_vjp(src, _dev(torch, u), _dev(torch, mu)))
    J = A.operator_jacobian(kind, W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3),
                            u, eps=1e-6)
    ref = J.T @ mu[3:-3]
    assert _rel(got[3:-3], ref) < RTOL
    np.testing.assert_array_equal(got[:3], 0.0)
    np.testing.assert_array_equal(got[-3:], 0.0)


@pytest.mark.parametrize("kind", KINDS_LF)
@pytest.mark.parametrize("n", [40, 600])
def test_step_vjp_vs_complex_step(torch_mod, kind: str, n: int) -> None:
    torch = torch_mod
    from oracle import adjoint_cs as A
    from oracle import weno_np as W
    from pyshocks_amd.adjoint import ssprk33_step_vjp

    u, lam = _data(n, 2)
    dt = 0.4 / n
    src, _ = _source(torch, kind, n)
    got = _host(ssprk33_step_vjp(src, _dev(torch, u), _dev(torch, lam), dt))
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    J = A.step_jacobian(kind, grid, u, dt, eps=1e-6)
    ref = J.T @ lam[3:-3]
    assert _rel(got[3:-3], ref) < RTOL


def test_operator_vjp_dot_product(torch_mod) -> None:
    """<J v, w> = <v, J^T w> with J v from the complex-step oracle."""
    torch = torch_mod
    from oracle import adjoint_cs as A
    from oracle import weno_np as W
    from pyshocks_amd.adjoint import operator_vjp

    n = 64
    u, w = _data(n, 3)
    v = np.random.default_rng(4).standard_normal(n + 6)
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    base = W.apply_boundary_periodic(grid, u).astype(np.complex128)
    z = W.apply_boundary_periodic(grid, base + 1e-30j * W.apply_boundary_periodic(grid, v))
    Jv = A.apply_operator_cs("rusanov", grid, z, eps=1e-6).imag / 1e-30
    src, _ = _source(torch, "rusanov", n)
    JTw = _host(operator_vjp(src, _dev(torch, u), _dev(torch, w)))
    lhs = float(Jv[3:-3] @ w[3:-3])
    rhs = float(v[3:-3] @ JTw[3:-3])
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adjoint_step_reference_api(torch_mod) -> None:
    """``adjoint_step`` over checkpoints of ``step`` (adaptive dt held
    constant per step) equals the product of dense oracle Jacobians."""
    torch = torch_mod
    from oracle import adjoint_cs as A
    from oracle import weno_np as W
    from pyshocks_amd.checkpointing import InMemoryCheckpoint
    from pyshocks_amd.timestepping import adjoint_step, make_ssprk33, step

    n = 32
    u0, _ = _data(n, 5)
    src, grid = _source(torch, "rusanov", n)
    chk = InMemoryCheckpoint(basename="Iteration")
    stepper = make_ssprk33(src.scheme, grid, src.bc, theta=1.0, checkpoint=chk)
    events = list(step(stepper, _dev(torch, u0), maxit=6))
    uT = events[-1].u
    p = None
    for ev in adjoint_step(stepper, uT.clone(), maxit=6):
        p = ev.p
    # oracle: p0 = J_0^T ... J_5^T u_T
    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    lam = _host(uT)[3:-3]
    states = [_host(e.u) for e in events]
    for m in range(5, -1, -1):
        dt = float(events[m + 1].t - events[m].t)
        J = A.step_jacobian("rusanov", ogrid, states[m], dt, eps=1e-6)
        lam = J.T @ lam
    assert _rel(_host(p)[3:-3], lam) < 1e-11


_FORWARD_TO_DISK = """
import sys
import numpy as np
import torch
from pyshocks_amd import make_uniform_cell_grid
from pyshocks_amd.burgers import Rusanov
from pyshocks_amd.checkpointing import NumpyCheckpoint
from pyshocks_amd.reconstruction import WENOJS53
from pyshocks_amd.scalar import PeriodicBoundary
from pyshocks_amd.timestepping import make_ssprk33, step

n = int(sys.argv[2])
u0 = torch.tensor(np.load(sys.argv[3]), device="cuda")
grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
chk = NumpyCheckpoint(basename="Iteration", dirname=sys.argv[1])
st = make_ssprk33(Rusanov(rec=WENOJS53(eps=1e-6)), grid, PeriodicBoundary(), theta=1.0,
                  checkpoint=chk)
for _ in step(st, u0, maxit=6):
    pass
"""


def test_adjoint_resumes_from_disk(torch_mod, tmp_path) -> None:
    """The forward run writes its checkpoints as npz files in another
    process; ``adjoint_step`` over those files gives the gradient of the
    in-memory run bit for bit (``checkpointing.py:209-254``)."""
    import subprocess
    import sys

    torch = torch_mod
    from pyshocks_amd.checkpointing import InMemoryCheckpoint, NumpyCheckpoint
    from pyshocks_amd.timestepping import adjoint_step, make_ssprk33, step

    n = 48
    u0, _ = _data(n, 8)
    np.save(tmp_path / "u0.npy", u0)
    r = subprocess.run([sys.executable, "-c", _FORWARD_TO_DISK, str(tmp_path), str(n),
                        str(tmp_path / "u0.npy")], timeout=180, capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (tmp_path / "Iteration_000000006.npz").exists()

    src, grid = _source(torch, "rusanov", n)
    mem = InMemoryCheckpoint(basename="Iteration")
    st_mem = make_ssprk33(src.scheme, grid, src.bc, theta=1.0, checkpoint=mem)
    uT = list(step(st_mem, _dev(torch, u0), maxit=6))[-1].u
    st_disk = make_ssprk33(src.scheme, grid, src.bc, theta=1.0,
                           checkpoint=NumpyCheckpoint(basename="Iteration",
                                                      dirname=tmp_path))
    out = []
    for st in (st_mem, st_disk):
        p = None
        for ev in adjoint_step(st, uT.clone(), maxit=6):
            p = ev.p
        out.append(p)
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("kind", KINDS_LF)
def test_checkpointed_adjoint_matches_step_vjps(torch_mod, kind: str) -> None:
    torch = torch_mod
    from pyshocks_amd.adjoint import checkpointed_solve, ssprk33_step_vjp
    from pyshocks_amd.timestepping import advance, make_ssprk33

    n, nsteps, every = 300, 23, 5
    u0, lam = _data(n, 6)
    dt = 0.3 / n
    src, grid = _source(torch, kind, n)
    solve = checkpointed_solve(src, _dev(torch, u0), nsteps, dt, every=every)

    stepper = make_ssprk33(src.scheme, grid, src.bc, dt=dt)
    states = [_dev(torch, u0)]
    for _ in range(nsteps):
        states.append(advance(stepper, dt, 0.0, states[-1]))
    assert _rel(_host(solve.u_final)[3:-3], _host(states[-1])[3:-3]) < RTOL

    g = _dev(torch, lam)
    ref = g
    for s in range(nsteps - 1, -1, -1):
        ref = ssprk33_step_vjp(src, states[s], ref, dt)
    got = solve.gradient(g)
    assert _rel(_host(got)[3:-3], _host(ref)[3:-3]) < RTOL
    np.testing.assert_array_equal(_host(got)[:3], 0.0)


@pytest.mark.parametrize("kind", ["rusanov", "lf"])
def test_gradient_vs_central_fd(torch_mod, kind: str) -> None:
    """J = 0.5 ||u(T) - u_target||^2 dx; <grad J, d> against central
    differences of the GPU forward solve with h = 1e-7 on 8 seeded unit
    directions. The rounding noise of the forward solve (~1e-15 per cell)
    divided by 2h bounds the FD accuracy; with an O(0.1) mismatch it stays
    near 1e-5 of the directional derivative, hence the 1e-4 tolerance."""
    torch = torch_mod
    from pyshocks_amd.adjoint import checkpointed_solve

    n, nsteps, every = 1024, 200, 50
    rng = np.random.default_rng(7)
    x = (np.arange(n + 6) - 2.5) / n
    u0 = 0.5 + np.sin(2 * np.pi * x)
    target = u0 + 0.2 * np.sin(4 * np.pi * x + 1.0)
    dt = 0.5 * (1.0 / n) / 1.5
    h = 1.0 / n
    src, _ = _source(torch, kind, n)

    def objective(u):
        s = checkpointed_solve(src, u, nsteps, dt, every=every)
        d = (s.u_final - tgt)[3:-3]
        return 0.5 * float(torch.sum(d * d)) * h, s

    tgt = _dev(torch, target)
    J0, solve = objective(_dev(torch, u0))
    lam = (solve.u_final - tgt) * h
    grad = _host(solve.gradient(lam))
    eps = 1.0e-7
    for k in range(8):
        d = rng.standard_normal(n + 6)
        d[:3] = 0.0
        d[-3:] = 0.0
        d /= np.linalg.norm(d)
        Jp, _ = objective(_dev(torch, u0 + eps * d))
        Jm, _ = objective(_dev(torch, u0 - eps * d))
        fd = (Jp - Jm) / (2 * eps)
        ad = float(grad @ d)
        assert abs(fd - ad) <= 1e-4 * max(abs(ad), 1e-12), (k, fd, ad)


def test_gradient_vs_c_reverse_mode(torch_mod) -> None:
    """The checkpointed GPU gradient through the shock of u0 = 0.5 +
    sin(2 pi x) against the C reverse mode of the oracle (its own forward
    pass in the reference's expression order, whose states differ from the
    GPU's by rounding)."""
    torch = torch_mod
    from oracle import build as oracle_build
    from oracle import ref_c
    from pyshocks_amd.adjoint import checkpointed_solve

    if not ref_c.LIB.exists():
        oracle_build.build(verbose=False)
    n, nsteps, every = 1024, 300, 50
    x = (np.arange(n + 6) - 2.5) / n
    u0 = 0.5 + np.sin(2 * np.pi * x)
    target = u0 + 0.2 * np.sin(4 * np.pi * x + 1.0)
    dt = 0.5 * (1.0 / n) / 1.5
    src, _ = _source(torch, "rusanov", n)
    solve = checkpointed_solve(src, _dev(torch, u0), nsteps, dt, every=every)
    tgt = _dev(torch, target)
    grad = _host(solve.gradient((solve.u_final - tgt) * (1.0 / n)))
    cgrad, cu, _ = ref_c.adjoint_rusanov(u0, target, n=n, g=3, dx=1.0 / n, eps=1e-6, dt=dt,
                                         nsteps=nsteps, every=every)
    fwd_rel = _rel(_host(solve.u_final)[3:-3], cu[3:-3])
    rel = _rel(grad[3:-3], cgrad[3:-3])
    print(f"[adjoint] N={n} {nsteps} steps: u(T) rel {fwd_rel:.3e}, gradient rel {rel:.3e}")
    assert fwd_rel < RTOL
    assert rel < RTOL


def test_autograd_function(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd.adjoint import checkpointed_solve, ssprk33_solve

    n, nsteps = 256, 30
    u0, lam = _data(n, 8)
    dt = 0.3 / n
    src, _ = _source(torch, "rusanov", n)
    u = _dev(torch, u0).requires_grad_(True)
    out = ssprk33_solve(u, src, nsteps, dt, 7)
    loss = torch.sum(out * _dev(torch, lam))
    loss.backward()
    ref = checkpointed_solve(src, _dev(torch, u0), nsteps, dt, every=7).gradient(
        _dev(torch, lam))
    assert _rel(_host(u.grad)[3:-3], _host(ref)[3:-3]) < RTOL


@pytest.mark.parametrize("kind", KINDS)
def test_checkpointed_adjoint_schedules_agree(torch_mod, kind: str) -> None:
    """The forward and recompute passes as whole-step kernels (the stage
    states written by the step) give the stage kernels' final state and
    gradient bit for bit (several tiles, a partial last window)."""
    torch = torch_mod
    from pyshocks_amd import _lib
    from pyshocks_amd.adjoint import CheckpointedSolve

    n, nsteps, every = 5000, 23, 5
    u0, lam = _data(n, 7)
    dt = 0.3 / n
    src, _ = _source(torch, kind, n)
    out = []
    for sched in (_lib.PSA_SCHEDULE_TILE, _lib.PSA_SCHEDULE_STEP):
        solve = CheckpointedSolve(src, _dev(torch, u0), nsteps, dt, every=every,
                                  schedule=sched)
        out.append((solve.u_final.clone(), solve.gradient(_dev(torch, lam))))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])


@pytest.mark.parametrize("kind", KINDS)
def test_batched_adjoint_equals_single(torch_mod, kind: str) -> None:
    """A batch of instances (own states, step sizes and cotangents) in one
    launch per step gives every instance's single-solve state and gradient
    bit for bit (a partial last window)."""
    torch = torch_mod
    from pyshocks_amd.adjoint import BatchedCheckpointedSolve, CheckpointedSolve

    n, nsteps, every, batch = 1000, 37, 8, 5
    src, _ = _source(torch, kind, n)
    # smooth states (the upwind flux without an entropy fix blows up on the
    # rough data of _data over tens of steps)
    x = (np.arange(n + 6) - 2.5) / n
    rng = np.random.default_rng(100)
    states = [0.3 + np.sin(2 * np.pi * (x + rng.uniform())) for _ in range(batch)]
    lams = [rng.standard_normal(n + 6) for _ in range(batch)]
    dts = [0.1 / n * (1.0 + 0.1 * b) for b in range(batch)]
    u0 = torch.tensor(np.stack(states), dtype=torch.float64, device="cuda")
    lam = torch.tensor(np.stack(lams), dtype=torch.float64, device="cuda")
    bs = BatchedCheckpointedSolve(src, u0, nsteps, dts, every)
    gb = bs.gradient(lam)
    assert bool(torch.isfinite(bs.u_final).all()) and bool(torch.isfinite(gb).all())
    for b in range(batch):
        s = CheckpointedSolve(src, u0[b].contiguous(), nsteps, dts[b], every)
        assert torch.equal(bs.u_final[b], s.u_final)
        assert torch.equal(gb[b], s.gradient(lam[b].contiguous()))


@pytest.mark.parametrize("kind", KINDS)
def test_reverse_step_forms_agree(torch_mod, kind: str) -> None:
    """A batch whose reverse step is a grid of more than 512 workgroups
    (8 instances of 16384 cells) runs each cell's two reconstructions in one
    thread, a single instance splits them between two threads: the same
    gradients bit for bit."""
    torch = torch_mod
    from pyshocks_amd.adjoint import BatchedCheckpointedSolve, CheckpointedSolve

    n, nsteps, batch = 16384, 6, 8
    src, _ = _source(torch, kind, n)
    x = (np.arange(n + 6) - 2.5) / n
    rng = np.random.default_rng(102)
    u0 = torch.tensor(np.stack([0.3 + np.sin(2 * np.pi * (x + rng.uniform()))
                                for _ in range(batch)]), dtype=torch.float64, device="cuda")
    lam = torch.tensor(rng.standard_normal((batch, n + 6)), dtype=torch.float64,
                       device="cuda")
    dts = [0.3 / n * (1.0 + 0.05 * b) for b in range(batch)]
    gb = BatchedCheckpointedSolve(src, u0, nsteps, dts, nsteps).gradient(lam)
    assert bool(torch.isfinite(gb).all())
    for b in (0, batch - 1):
        s = CheckpointedSolve(src, u0[b].contiguous(), nsteps, dts[b], nsteps)
        assert torch.equal(gb[b], s.gradient(lam[b].contiguous()))


@pytest.mark.parametrize("kind", KINDS_LF)
def test_one_window_keeps_the_forward_states(torch_mod, kind: str) -> None:
    """With every >= nsteps the forward pass keeps the stage states and the
    reverse sweep recomputes nothing: the same final state and gradient, bit
    for bit, as windows recomputed from checkpoints, and the same gradient
    again on a second reverse sweep (single solves on both schedules, and a
    batch)."""
    torch = torch_mod
    from pyshocks_amd import _lib
    from pyshocks_amd.adjoint import BatchedCheckpointedSolve, CheckpointedSolve

    n, nsteps = 3000, 23
    # smooth data (the upwind flux without an entropy fix blows up on the
    # rough data of _data, and NaN compares unequal to itself)
    x = (np.arange(n + 6) - 2.5) / n
    rng = np.random.default_rng(11)
    u0 = 0.3 + np.sin(2 * np.pi * x) + 0.01 * np.sin(14 * np.pi * x)
    lam = rng.standard_normal(n + 6)
    dt = 0.3 / n
    src, _ = _source(torch, kind, n)
    for sched in (_lib.PSA_SCHEDULE_TILE, _lib.PSA_SCHEDULE_STEP):
        out = []
        for every in (5, nsteps, nsteps + 4):
            solve = CheckpointedSolve(src, _dev(torch, u0), nsteps, dt, every=every,
                                      schedule=sched)
            g = solve.gradient(_dev(torch, lam))
            assert bool(torch.isfinite(g).all())
            assert torch.equal(g, solve.gradient(_dev(torch, lam)))
            out.append((solve.u_final.clone(), g))
        for o in out[1:]:
            assert torch.equal(out[0][0], o[0])
            assert torch.equal(out[0][1], o[1])
    if kind == "lf":
        return
    x = (np.arange(n + 6) - 2.5) / n
    rng = np.random.default_rng(101)
    u0b = torch.tensor(np.stack([0.3 + np.sin(2 * np.pi * (x + rng.uniform()))
                                 for _ in range(3)]), dtype=torch.float64, device="cuda")
    lamb = torch.tensor(rng.standard_normal((3, n + 6)), dtype=torch.float64, device="cuda")
    dts = [0.1 / n * (1.0 + 0.1 * b) for b in range(3)]
    out = []
    for every in (8, nsteps):
        bs = BatchedCheckpointedSolve(src, u0b, nsteps, dts, every)
        out.append((bs.u_final.clone(), bs.gradient(lamb)))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])


# {{{ learned weights


def _learned_source(torch, kind: str, n: int, p, tau: float, sigma: float):
    from pyshocks_amd import burgers, make_uniform_cell_grid
    from pyshocks_amd.learned import MLPWeights
    from pyshocks_amd.reconstruction import LearnedWENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import BoundOperator

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rec = LearnedWENOJS53(eps=1e-6, weights=MLPWeights.from_numpy(p, tau, sigma))
    cls = {"rusanov": burgers.Rusanov, "godunov": burgers.Godunov,
           "eo": burgers.EngquistOsher}[kind]
    return BoundOperator(cls(rec=rec), grid, PeriodicBoundary()), grid


def _learned_case(n: int, seed: int):
    """Parameters with BF16-representable weights (so that the oracle's
    surrogate has the kernels' operands) and a state with smooth parts, a
    step and noise: shock, smooth and learned cells all occur."""
    from oracle import mlp_np as M

    rng = np.random.default_rng(seed)
    p = (0.3 * rng.standard_normal(M.NPARAM)).astype(np.float32)
    P = M.unpack(p)
    p = M.pack(M.bf16(P["W1"]), P["b1"], M.bf16(P["W2"]), P["b2"], M.bf16(P["W3"]), P["b3"])
    x = (np.arange(n + 6) - 2.5) / n
    u = 0.5 + np.sin(2 * np.pi * x) + 0.3 * (x > 0.5) + 0.02 * rng.standard_normal(n + 6)
    return p, u, rng.standard_normal(n + 6)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo"])
def test_learned_operator_vjp_vs_complex_step(torch_mod, kind: str) -> None:
    """J_L(u)^T mu of the learned operator against the dense complex-step
    Jacobian of the oracle's surrogate (shock cells WENO-JS, smooth cells
    the ideal weights, learned cells the network with BF16 weights and
    unrounded activations), FP64, every regime present."""
    torch = torch_mod
    from oracle import adjoint_cs as A
    from oracle import mlp_np as M
    from oracle import weno_np as W
    from pyshocks_amd.adjoint import operator_vjp

    n, tau, sigma = 48, 20.0, 1.5
    for seed in (1, 2, 3):
        p, u, mu = _learned_case(n, seed)
        ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
        uf = W.apply_boundary_periodic(ogrid, u)
        reg = M.regimes(M.stencils(uf), 1e-6, tau, sigma)[3:-3]
        assert np.all(np.bincount(reg, minlength=3) > 0)
        src, _ = _learned_source(torch, kind, n, p, tau, sigma)
        got = _host(operator_vjp(src, _dev(torch, uf), _dev(torch, mu)))[3:-3]
        J = A.learned_operator_jacobian(p, tau, sigma, kind, ogrid, uf, eps=1e-6)
        ref = J.T @ mu[3:-3]
        assert _rel(got, ref) < 1e-11, (seed, _rel(got, ref))


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo"])
def test_learned_adjoint_tau0_equals_plain(torch_mod, kind: str) -> None:
    """With tau = 0 every varying cell is in the shock regime: the learned
    checkpointed adjoint reproduces the WENO-JS one bit for bit."""
    torch = torch_mod
    from pyshocks_amd.adjoint import CheckpointedSolve

    n, nsteps, every = 1000, 40, 9
    p, _, _ = _learned_case(n, 5)
    # a sine with a step and small noise: every cell varies, and the
    # reference's Godunov flux (no entropy fix at sonic points) stays finite
    # on it, which it does not on O(1) cell-to-cell noise
    rng = np.random.default_rng(6)
    x = (np.arange(n + 6) - 2.5) / n
    u0 = 0.5 + np.sin(2 * np.pi * x) + 0.3 * (x > 0.4) + 0.01 * rng.standard_normal(n + 6)
    lam = rng.standard_normal(n + 6)
    dt = 0.2 / n
    src_l, _ = _learned_source(torch, kind, n, p, 0.0, 0.0)
    src_p, _ = _source(torch, kind, n)
    sl = CheckpointedSolve(src_l, _dev(torch, u0), nsteps, dt, every)
    sp = CheckpointedSolve(src_p, _dev(torch, u0), nsteps, dt, every)
    assert bool(torch.isfinite(sp.u_final[3:-3]).all())
    assert torch.equal(sl.u_final[3:-3], sp.u_final[3:-3])
    gl = sl.gradient(_dev(torch, lam))
    gp = sp.gradient(_dev(torch, lam))
    assert torch.equal(gl, gp)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo"])
def test_learned_adjoint_consistency(torch_mod, kind: str) -> None:
    """The learned checkpointed adjoint (whole-step recompute, fused reverse
    step) equals the chain of single-step VJPs (stage kernels) bit for bit,
    and a batch of instances equals the single solves."""
    torch = torch_mod
    from pyshocks_amd.adjoint import (BatchedCheckpointedSolve, CheckpointedSolve,
                                      ssprk33_step_vjp)
    from pyshocks_amd.timestepping import advance, make_ssprk33

    n, nsteps, every = 700, 12, 5
    p, _, _ = _learned_case(n, 7)
    src, grid = _learned_source(torch, kind, n, p, 20.0, 1.5)
    x = (np.arange(n + 6) - 2.5) / n
    rng = np.random.default_rng(8)
    states = [0.5 + np.sin(2 * np.pi * (x + rng.uniform())) + 0.3 * (x > rng.uniform())
              for _ in range(3)]
    lams = [rng.standard_normal(n + 6) for _ in range(3)]
    dt = 0.2 / n
    stepper = make_ssprk33(src.scheme, grid, src.bc, dt=dt)
    grads = []
    for b in range(3):
        solve = CheckpointedSolve(src, _dev(torch, states[b]), nsteps, dt, every=every)
        g = solve.gradient(_dev(torch, lams[b]))
        traj = [_dev(torch, states[b])]
        for _ in range(nsteps):
            traj.append(advance(stepper, dt, 0.0, traj[-1]))
        ref = _dev(torch, lams[b])
        for s in range(nsteps - 1, -1, -1):
            ref = ssprk33_step_vjp(src, traj[s], ref, dt)
        assert bool(torch.isfinite(g).all())
        assert torch.equal(g[3:-3], ref[3:-3]), float((g - ref).abs().max())
        grads.append(g)
    bs = BatchedCheckpointedSolve(src, _dev(torch, np.stack(states)), nsteps, dt, every)
    gb = bs.gradient(_dev(torch, np.stack(lams)))
    for b in range(3):
        assert torch.equal(gb[b], grads[b])


# }}}
