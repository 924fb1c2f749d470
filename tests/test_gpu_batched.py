"""GPU tests of the batched persistent solver: every instance must take the
same steps as the single-instance fused solve and agree with it to 1e-12
(FP64; the tile arithmetic is shared, but each kernel's floating-point
contraction may differ by an ulp) or 1e-5 (FP32), match the oracle to 1e-12
(FP64), graph replays must reproduce the eager result, and the queue /
step-limit logic must terminate correctly."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _single(torch, n: int, u0: np.ndarray, dtype, tfinal: float, scheme: str = "rusanov"):
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.burgers import EngquistOsher, Godunov, LaxFriedrichs, Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import evolve, make_ssprk33

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, dtype=dtype)
    cls = {"rusanov": Rusanov, "godunov": Godunov, "eo": EngquistOsher,
           "lf": LaxFriedrichs}[scheme]
    sch = cls(rec=WENOJS53(eps=1e-6))
    st = make_ssprk33(sch, grid, PeriodicBoundary(), theta=1.0)
    u = torch.tensor(u0, device="cuda").to(dtype)
    return evolve(st, u, tfinal=tfinal, tiles=(256, 4))


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("scheme", ["rusanov", "godunov", "eo", "lf"])
def test_batched_equals_single_instance(torch_mod, dtype_name: str, scheme: str) -> None:
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    dtype = getattr(torch, dtype_name)
    _, ics = random_instances(6, seed=3)
    ns = np.array([3, 37, 128, 1000, 3000, 4100])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    # time in the state dtype, as the single-instance path and the reference
    solver = BatchedSolver(ns, dtype=dtype, scheme=scheme, tiles=(256, 4), grid=3,
                           time_f64=False)
    solver.set_states(states)
    res = solver.solve(0.05)
    assert np.all(res.status == 0)
    for i, n in enumerate(ns):
        ev = _single(torch, int(n), states[i], dtype, 0.05, scheme)
        assert res.steps[i] == ev.iteration
        assert res.t[i] == float(ev.t)
        got = solver.state(i).cpu().double().numpy()[3:-3]
        ref = ev.u.cpu().double().numpy()[3:-3]
        tol = 1e-12 if dtype == torch.float64 else 1e-5
        assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("scheme", ["rusanov", "eo", "lf"])
def test_batched_vs_oracle(torch_mod, scheme: str) -> None:
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _, ics = random_instances(3, seed=4)
    ns = np.array([64, 200, 513])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    solver = BatchedSolver(ns, dtype=torch.float64, scheme=scheme)
    solver.set_states(states)
    res = solver.solve(0.3)
    for i, n in enumerate(ns):
        grid = W.make_uniform_cell_grid(0.0, 1.0, int(n), nghosts=3)
        u, dts, t = W.solve(W.Scheme(scheme, eps=1e-6), grid, states[i], theta=1.0,
                            tfinal=0.3)
        assert res.steps[i] == len(dts)
        got = solver.state(i).cpu().numpy()[3:-3]
        assert np.linalg.norm(got - u[3:-3]) / np.linalg.norm(u[3:-3]) < 1e-12


def test_batched_graph_replay(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    ns, ics = random_instances(40, nmin=16, nmax=2048, seed=5)
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    solver = BatchedSolver(ns, dtype=torch.float32)
    solver.set_states(states)
    u0 = solver.u.clone()
    eager = solver.solve(0.02)
    ref = solver.u.clone()
    graph = solver.capture(u0, 0.02)
    for _ in range(3):
        solver.u.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(solver.u, ref)
        res = solver.result()
        np.testing.assert_array_equal(res.steps, eager.steps)


def test_batched_step_limit_terminates(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    ns, ics = random_instances(5, nmin=32, nmax=256, seed=6)
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    solver = BatchedSolver(ns, dtype=torch.float64, grid=2)
    solver.desc.max_steps = 5
    solver.set_states(states)
    res = solver.solve(10.0)
    np.testing.assert_array_equal(res.steps, 5)
    np.testing.assert_array_equal(res.status, -1)


def test_autotune_picks_a_launchable_config(torch_mod) -> None:
    torch = torch_mod
    from pyshocks_amd import autotune, funcs, make_uniform_cell_grid
    from pyshocks_amd.batched import exact_averages_np, random_instances
    from pyshocks_amd.burgers import Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33

    grid = make_uniform_cell_grid(0.0, 1.0, 1 << 16, nghosts=3)
    st = make_ssprk33(Rusanov(rec=WENOJS53(eps=1e-6)), grid, PeriodicBoundary())
    plan = fused_plan(st)
    u0 = funcs.cell_averages(funcs.random_sine_step(0), grid)
    best, times = autotune.tune_stage(plan, u0, steps=2)
    assert best in times and len(times) >= 3
    assert fused_plan(st).desc.block == best[0]

    ns, ics = random_instances(16, nmin=64, nmax=4096, seed=7)
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    bbest, btimes = autotune.tune_batch(ns, states, tfinal=0.005)
    assert bbest in btimes


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("scheme", ["rusanov", "lf"])
def test_teams_match_single_workgroup_bit_for_bit(torch_mod, dtype_name: str,
                                                  scheme: str) -> None:
    """Splitting an instance's tiles among a team changes who computes a
    tile, not what is computed (Lax-Friedrichs: nor the stage speeds the
    team reduces)."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    dtype = getattr(torch, dtype_name)
    _, ics = random_instances(5, seed=9)
    ns = np.array([20000, 50, 3000, 9000, 700])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    out = []
    for team_cells in (0, None, 2048):
        solver = BatchedSolver(ns, dtype=dtype, tiles=(256, 4), team_cells=team_cells,
                               scheme=scheme)
        # 0: one workgroup each; None: teams sized by work (a batch this
        # small spreads its larger instances over many workgroups)
        assert (solver.team.max() > 1) == (team_cells != 0)
        assert solver.team.max() <= solver.resident_workgroups()
        solver.set_states(states)
        res = solver.solve(0.02)
        assert np.all(res.status == 0)
        out.append((solver.u.clone(), res))
    for o in out[1:]:
        assert torch.equal(out[0][0], o[0])
        np.testing.assert_array_equal(out[0][1].steps, o[1].steps)
        np.testing.assert_array_equal(out[0][1].t, o[1].t)


def test_work_teams_respect_an_explicit_grid(torch_mod) -> None:
    """Teams sized by work fit the grid: none when an explicit grid exceeds
    the resident workgroups (a team's members must all be resident), at
    most ``grid`` members for a smaller explicit grid; every form gives the
    same bits."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _, ics = random_instances(5, seed=12)
    ns = np.array([20000, 50, 3000, 9000, 700])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    resident = BatchedSolver(ns, dtype=torch.float64, tiles=(256, 4)).resident_workgroups()
    assert resident > 0
    out = []
    for grid, most in ((0, None), (2 * resident, 1), (3, 3)):
        solver = BatchedSolver(ns, dtype=torch.float64, tiles=(256, 4), grid=grid)
        if most is None:
            assert solver.team.max() > 3
        else:
            assert solver.team.max() == most
        solver.set_states(states)
        res = solver.solve(0.02)
        assert np.all(res.status == 0)
        out.append(solver.u.clone())
    assert torch.equal(out[0], out[1]) and torch.equal(out[0], out[2])


def test_team_wait_limit_terminates(torch_mod) -> None:
    """A wait limit of zero makes a member that has to wait at a barrier
    abort its team: the launch still ends, team instances report
    PSA_ERR_HIP or finish, single-workgroup instances are unaffected."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _, ics = random_instances(4, seed=10)
    ns = np.array([30000, 100, 200, 12000])
    solver = BatchedSolver(ns, dtype=torch.float32, tiles=(256, 4), team_cells=2048,
                           team_wait_s=1e-12)
    solver.set_states([exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)])
    res = solver.solve(0.01)
    teamed = solver.team > 1
    assert np.all(np.isin(res.status[teamed], [0, -3]))
    assert np.all(res.status[~teamed] == 0)


def test_time_f64_reaches_tfinal(torch_mod) -> None:
    """FP32 states with FP64 time end at tfinal (to the FP32 rounding of the
    last step) instead of drifting with FP32 sums of t."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _, ics = random_instances(3, seed=11)
    ns = np.array([4000, 16000, 1000])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    r = {}
    for tf64 in (True, False):
        solver = BatchedSolver(ns, dtype=torch.float32, time_f64=tf64)
        solver.set_states(states)
        r[tf64] = solver.solve(0.3)
        assert np.all(r[tf64].status == 0)
    assert np.max(np.abs(r[True].t - 0.3)) < 1e-9
    assert np.max(np.abs(r[False].t - 0.3)) > np.max(np.abs(r[True].t - 0.3))


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("scheme", ["rusanov", "godunov", "eo"])
@pytest.mark.parametrize("team_cells", [0, None, 2048])
def test_step_schedule_matches_stage_sweeps(torch_mod, dtype_name: str, scheme: str,
                                            team_cells) -> None:
    """One sweep per step (the three stages of a tile in LDS, the state
    alternating between u and k1) computes what one sweep per stage
    computes, bit for bit: team and single-workgroup instances, odd and even
    step counts, instances smaller than a tile."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    dtype = getattr(torch, dtype_name)
    _, ics = random_instances(6, seed=21)
    ns = np.array([3, 37, 1000, 9000, 20000, 700])
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    out = []
    for sched in ("tile", "step"):
        solver = BatchedSolver(ns, dtype=dtype, scheme=scheme, tiles=(256, 4),
                               team_cells=team_cells, schedule=sched)
        solver.set_states(states)
        res = solver.solve(0.02)
        assert np.all(res.status == 0)
        out.append((solver.u.clone(), res))
    assert len(set(int(k) % 2 for k in out[0][1].steps)) == 2  # odd and even counts
    np.testing.assert_array_equal(out[0][1].steps, out[1][1].steps)
    np.testing.assert_array_equal(out[0][1].t, out[1][1].t)
    assert torch.equal(out[0][0], out[1][0])


def test_kernel_info_and_autotune_skip_spills(torch_mod) -> None:
    """psa_batch_kernel_info reports the selected kernel's resources, and the
    autotuner never picks a candidate whose kernel spills."""
    torch = torch_mod
    from pyshocks_amd import autotune
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    ns, ics = random_instances(8, nmin=64, nmax=2048, seed=9)
    states = [exact_averages_np(ic, int(n)) for ic, n in zip(ics, ns)]
    for dtype in (torch.float64, torch.float32):
        for tiles in autotune.BATCH_CANDIDATES:
            info = BatchedSolver(ns, dtype=dtype, tiles=tiles).kernel_info()
            assert 0 < info["vgprs"] <= 512
            assert info["lds_bytes"] > 0 and info["private_bytes"] >= 0
        best, times = autotune.tune_batch(ns, states, dtype=dtype, tfinal=0.005)
        assert BatchedSolver(ns, dtype=dtype, tiles=best).kernel_info()["private_bytes"] == 0
        for cand in times:
            assert BatchedSolver(ns, dtype=dtype, tiles=cand).kernel_info()["private_bytes"] == 0
