"""The batched persistent solver with the other reconstructions (WENO-JS3,
ES-WENO3, MUSCL with flux and slope limiters): every instance takes the
steps of the single-instance fused solve and agrees with it to 1e-12 (FP64)
or 1e-5 (FP32), matches the oracle to 1e-12, and the step schedule equals
the stage sweeps bit for bit (teams included)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RECS = ["wenojs32", "esweno32", "muscl-minmod", "muscls-vanleer"]


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _rec(name: str):
    from pyshocks_amd.limiters import make_limiter_from_name
    from pyshocks_amd.reconstruction import make_reconstruction_from_name

    if name.startswith("muscl"):
        kind, lm = name.split("-")
        return make_reconstruction_from_name(kind, lm=make_limiter_from_name(lm, theta=1.0))
    return make_reconstruction_from_name(name, eps=1e-6)


def _oname(name: str) -> str:
    return f"{name}-1.0" if name.startswith("muscl") else name


def _single(torch, n: int, u0: np.ndarray, dtype, tfinal: float, scheme: str, rec):
    from pyshocks_amd import make_uniform_cell_grid
    from pyshocks_amd.burgers import EngquistOsher, Godunov, Rusanov
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import evolve, make_ssprk33

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=2, dtype=dtype)
    cls = {"rusanov": Rusanov, "godunov": Godunov, "eo": EngquistOsher}[scheme]
    st = make_ssprk33(cls(rec=rec), grid, PeriodicBoundary(), theta=1.0)
    u = torch.tensor(u0, device="cuda").to(dtype)
    return evolve(st, u, tfinal=tfinal, tiles=(256, 4))


@pytest.mark.parametrize("rec_name", RECS)
@pytest.mark.parametrize("scheme", ["rusanov", "eo"])
@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_batched_equals_single_instance(torch_mod, rec_name: str, scheme: str,
                                        dtype_name: str) -> None:
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    dtype = getattr(torch, dtype_name)
    rec = _rec(rec_name)
    _, ics = random_instances(5, seed=7)
    ns = np.array([3, 37, 256, 1000, 4100])
    states = [exact_averages_np(ic, int(n), nghosts=2) for ic, n in zip(ics, ns)]
    solver = BatchedSolver(ns, dtype=dtype, scheme=scheme, nghosts=2, grid=3,
                           time_f64=False, rec=rec)
    solver.set_states(states)
    res = solver.solve(0.05)
    assert np.all(res.status == 0)
    for i, n in enumerate(ns):
        ev = _single(torch, int(n), states[i], dtype, 0.05, scheme, rec)
        assert res.steps[i] == ev.iteration
        assert res.t[i] == float(ev.t)
        got = solver.state(i).cpu().double().numpy()[2:-2]
        ref = ev.u.cpu().double().numpy()[2:-2]
        tol = 1e-12 if dtype == torch.float64 else 1e-5
        assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("rec_name", RECS)
def test_batched_vs_oracle(torch_mod, rec_name: str) -> None:
    """Power-of-two grids, whose cell centres are exact binary fractions
    (the slope-limited MUSCL of the reference divides by their differences)."""
    torch = torch_mod
    from oracle import weno_np as W
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _, ics = random_instances(3, seed=4)
    ns = np.array([64, 256, 512])
    states = [exact_averages_np(ic, int(n), nghosts=2) for ic, n in zip(ics, ns)]
    solver = BatchedSolver(ns, dtype=torch.float64, scheme="rusanov", nghosts=2,
                           rec=_rec(rec_name))
    solver.set_states(states)
    res = solver.solve(0.3)
    for i, n in enumerate(ns):
        grid = W.make_uniform_cell_grid(0.0, 1.0, int(n), nghosts=2)
        u, dts, _ = W.solve(W.Scheme("rusanov", eps=1e-6, rec=_oname(rec_name)), grid,
                            states[i], theta=1.0, tfinal=0.3)
        assert res.steps[i] == len(dts)
        got = solver.state(i).cpu().numpy()[2:-2]
        assert np.linalg.norm(got - u[2:-2]) / np.linalg.norm(u[2:-2]) < 1e-12


@pytest.mark.parametrize("rec_name", ["esweno32", "muscl-minmod", "muscls-vanleer"])
@pytest.mark.parametrize("team_cells", [0, 2048])
def test_step_schedule_matches_stage_sweeps(torch_mod, rec_name: str, team_cells) -> None:
    """team_cells 0: one workgroup per instance, also for the 20000-cell one
    (the teamless path of these reconstructions on a large instance); 2048:
    teams of up to 10 workgroups."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _, ics = random_instances(6, seed=21)
    ns = np.array([3, 37, 1000, 9000, 20000, 700])
    states = [exact_averages_np(ic, int(n), nghosts=2) for ic, n in zip(ics, ns)]
    out = []
    for sched in ("tile", "step"):
        solver = BatchedSolver(ns, dtype=torch.float64, scheme="rusanov", nghosts=2,
                               tiles=(256, 2), team_cells=team_cells, schedule=sched,
                               rec=_rec(rec_name))
        solver.set_states(states)
        res = solver.solve(0.02)
        assert np.all(res.status == 0)
        out.append((solver.u.clone(), res))
    np.testing.assert_array_equal(out[0][1].steps, out[1][1].steps)
    np.testing.assert_array_equal(out[0][1].t, out[1][1].t)
    assert torch.equal(out[0][0], out[1][0])


def test_refusals(torch_mod) -> None:
    """Learned weights with another reconstruction, tiles other than 256 x 4
    or 256 x 2, and the Koren limiter without a slope form are refused
    before any launch (when the solver sizes its teams from the descriptor's
    kernel, or at the latest when it solves)."""
    torch = torch_mod
    from pyshocks_amd.batched import BatchedSolver

    ns = np.array([64, 128])
    with pytest.raises(ValueError, match="learned"):
        BatchedSolver(ns, dtype=torch.float64, nghosts=2, rec=_rec("wenojs32"),
                      mlp=object())
    for team_cells in (None, 0):
        with pytest.raises(ValueError, match="256"):
            solver = BatchedSolver(ns, dtype=torch.float64, nghosts=2, tiles=(512, 2),
                                   rec=_rec("muscl-minmod"), team_cells=team_cells)
            solver.set_states([np.zeros(68), np.zeros(132)])
            solver.solve(0.1)
        with pytest.raises(NotImplementedError, match="Koren"):
            solver = BatchedSolver(ns, dtype=torch.float64, nghosts=2,
                                   rec=_rec("muscls-koren"), team_cells=team_cells)
            solver.set_states([np.zeros(68), np.zeros(132)])
            solver.solve(0.1)
