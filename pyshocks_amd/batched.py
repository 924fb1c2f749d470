"""Batched request front end: many independent periodic Burgers instances of
uneven size advanced in one persistent launch.

Each instance ``i`` is a uniform periodic grid of ``ncells[i]`` cells on
``[a, b]`` (3 ghost cells per side) with its own initial state and final time,
advanced by the fused WENO-JS5 / Rusanov (or Godunov) / SSP-RK3 step (or its
learned-weight form, ``mlp``, or another reconstruction, ``rec``: WENO-JS3,
ES-WENO3 and MUSCL with flux or slope limiters) with its own CFL time step -- the loop of ``timestepping.py:89-152`` per instance, as
the reference harness would run it for each request separately. The instances
are packed into three device arrays (state and two stage buffers) at 16-byte
aligned offsets; the library runs them longest-first from an atomic queue,
one workgroup per instance, or a team of workgroups for an instance with more
work than the mean per resident workgroup (``psa_batch_solve``). The launch
can be captured in a HIP graph and replayed (:meth:`BatchedSolver.capture`).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Sequence

import numpy as np
import torch

from pyshocks_amd import _lib

ALIGN = 64  # elements; keeps every instance 16-byte aligned in FP32 and FP64


def work_teams(ncells: np.ndarray, grid: int, cap: np.ndarray | int,
               share: float = 1.0) -> np.ndarray:
    """Team sizes from work: on one domain an instance's CFL step count grows
    like its cell count, so its work (cell-updates) like the square. An
    instance with more than ``share`` times the batch's mean work per
    workgroup of a ``grid`` of workgroups would still run on one workgroup
    after the others ran out of instances; it is split among
    ``ceil(work / (share * mean))`` members (at most ``cap``), which leaves
    the small instances, whose barriers would cost more than they save, on
    one workgroup each."""
    work = np.asarray(ncells, dtype=np.float64) ** 2
    mean = share * work.sum() / grid
    return np.clip(np.ceil(work / mean), 1, cap).astype(np.int32)


@dataclass
class BatchResult:
    t: np.ndarray
    """Final time of each instance."""
    steps: np.ndarray
    """Steps taken by each instance."""
    status: np.ndarray
    """Per-instance status (0 = reached its final time)."""


class BatchedSolver:
    """Packed storage and launch parameters of a batch of instances."""

    def __init__(self, ncells: Sequence[int], *, dtype: torch.dtype = torch.float32,
                 scheme: str = "rusanov", eps: float = 1.0e-6, theta: float = 1.0,
                 a: float = 0.0, b: float = 1.0, nghosts: int = 3,
                 device: torch.device | str = "cuda",
                 tiles: tuple[int, int] | None = None, grid: int = 0,
                 mlp: Any = None, time_f64: bool = True,
                 team_cells: int | str | None = None, max_team: int = 32,
                 team_wait_s: float = 60.0, schedule: str = "step",
                 rec: Any = None) -> None:
        """``rec`` is a reconstruction object of the GPU path
        (:mod:`pyshocks_amd.reconstruction`; ``None``: WENO-JS5 with ``eps``);
        its own ``eps`` replaces the argument. The reconstructions other than
        WENO-JS5 take 2 ghost cells, tiles 256 x 4 or 256 x 2 and no learned
        weights. ``team_cells`` ``None`` (or ``"work"``) sizes teams by work
        (:meth:`_work_teams`), a count splits every instance above that many
        cells, ``0`` runs one workgroup per instance. A team's workgroups
        must all be resident, so with an explicit ``grid`` no team is larger
        than ``grid`` workgroups."""
        if scheme not in _lib.BATCH_FLUX:
            raise NotImplementedError(
                f"the batched solver takes Burgers 'rusanov', 'godunov', 'lf' or 'eo', "
                f"not {scheme!r}")
        rec_kind = _lib.PSA_REC_WENOJS53
        if rec is not None:
            from pyshocks_amd.reconstruction import rec_code

            rec_kind = rec_code(rec)
            eps = float(getattr(rec, "eps", eps))
        gmin = 3 if rec_kind == _lib.PSA_REC_WENOJS53 else 2
        if team_cells == "work":
            team_cells = None
        elif team_cells is not None and (not isinstance(team_cells, (int, np.integer))
                                         or isinstance(team_cells, bool) or team_cells < 0):
            raise ValueError(f"team_cells must be a cell count (>= 0), 'work' or None, "
                             f"not {team_cells!r}")
        if nghosts < gmin:
            raise ValueError(f"the reconstruction needs nghosts >= {gmin}")
        if rec_kind != _lib.PSA_REC_WENOJS53 and mlp is not None:
            raise ValueError("learned weights replace the WENO-JS5 weights only")
        if schedule not in ("step", "tile"):
            raise ValueError(f"schedule must be 'step' or 'tile', not {schedule!r}")
        ncells = np.asarray(ncells, dtype=np.int64)
        if ncells.ndim != 1 or ncells.size == 0 or np.any(ncells < nghosts):
            raise ValueError("every instance needs at least nghosts interior cells")

        self.dtype = dtype
        self.device = torch.device(device)
        self.g = nghosts
        self.ncells = ncells
        self.m = ncells + 2 * nghosts
        self.h = (b - a) / ncells.astype(np.float64)
        self.a = a
        self.b = b
        sizes = ((self.m + ALIGN - 1) // ALIGN) * ALIGN
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.total = int(np.sum(sizes))

        self.u = torch.zeros(self.total, dtype=dtype, device=self.device)
        self.k1 = torch.zeros_like(self.u)
        self.k2 = torch.zeros_like(self.u)

        dev = self.device
        self._offsets = torch.tensor(self.offsets, dtype=torch.int64, device=dev)
        self._ncells = torch.tensor(ncells.astype(np.int32), dtype=torch.int32, device=dev)
        self._dx = torch.tensor(self.h, dtype=torch.float64, device=dev)
        # longest first: the step count of an instance grows like its cell count
        order = np.argsort(-ncells, kind="stable").astype(np.int32)
        self._order = torch.tensor(order, dtype=torch.int32, device=dev)
        self._tfinal = torch.zeros(ncells.size, dtype=torch.float64, device=dev)
        self._counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self.t_out = torch.zeros(ncells.size, dtype=torch.float64, device=dev)
        self.steps_out = torch.zeros(ncells.size, dtype=torch.int64, device=dev)
        self.status_out = torch.zeros(ncells.size, dtype=torch.int32, device=dev)

        self.desc = _lib.PsaBatchDesc()
        self.desc.dtype = _lib.dtype_code(dtype)
        self.desc.flux = _lib.BATCH_FLUX[scheme]
        self.desc.g = nghosts
        self.desc.grid = int(grid)
        self.desc.ninst = int(ncells.size)
        self.desc.offsets = self._offsets.data_ptr()
        self.desc.ncells = self._ncells.data_ptr()
        self.desc.dx = self._dx.data_ptr()
        self.desc.tfinal = self._tfinal.data_ptr()
        self.desc.order = self._order.data_ptr()
        self.desc.eps = float(eps)
        self.desc.theta = float(theta)
        self.desc.max_steps = 4_000_000
        # FP64 time keeps FP32 instances on schedule over ~10^5 steps; False
        # accumulates t in the state dtype as the reference's step loop does
        self.desc.time_f64 = 1 if time_f64 else 0
        # one sweep per step with the three stages of a tile in LDS, or one
        # sweep per stage (the same values)
        self.desc.schedule = _lib.PSA_SCHEDULE_STEP if schedule == "step" \
            else _lib.PSA_SCHEDULE_TILE
        self.rec = rec
        self.desc.rec = rec_kind
        if rec_kind in (_lib.PSA_REC_MUSCL, _lib.PSA_REC_MUSCLS):
            from pyshocks_amd.limiters import limiter_code

            self.desc.limiter, self.desc.limiter_theta = limiter_code(rec.lm)
            self.desc.muscl_atol = float(rec.atol)
        self.mlp = mlp
        if mlp is not None:
            # learned weights (pyshocks_amd.learned.MLPWeights)
            self.desc.mlp = mlp.ptr
            self.desc.mlp_tau = float(mlp.tau)
            self.desc.mlp_sigma = float(mlp.sigma)
        if tiles is None and rec_kind == _lib.PSA_REC_WENOJS53:
            from pyshocks_amd.autotune import lookup_batch

            tiles = lookup_batch(self.desc.dtype, self.desc.flux, mlp is not None)
        if tiles is not None:
            self.desc.block, self.desc.cells_per_thread = int(tiles[0]), int(tiles[1])
        self._set_teams(order, team_cells, max_team, team_wait_s)

    #: work share (of the mean work per resident workgroup) above which
    #: :meth:`_work_teams` splits an instance
    WORK_SHARE = 1.0

    def resident_workgroups(self) -> int:
        """Workgroups of the selected kernel the device holds at once (the
        persistent grid when ``grid`` is 0, and the largest team)."""
        wg = ctypes.c_int(0)
        _lib.call("psa_batch_resident_workgroups", ctypes.byref(self.desc), ctypes.byref(wg))
        return int(wg.value)

    def _work_teams(self, cap: np.ndarray) -> np.ndarray:
        """:func:`work_teams` on the persistent grid (none if an explicit
        ``grid`` exceeds what the device holds at once)."""
        resident = self.resident_workgroups()
        grid = self.desc.grid or resident
        if grid > resident:
            return np.ones(self.ninst, dtype=np.int32)
        return work_teams(self.ncells, grid, cap, self.WORK_SHARE)

    def _set_teams(self, order: np.ndarray, team_cells: int | None, max_team: int,
                   wait_s: float) -> None:
        """Instances above ``team_cells`` cells (``None``: the work rule of
        :meth:`_work_teams`) are split among a team of ``ceil(m / team_cells)``
        workgroups (at most ``max_team``, at most one per tile) that meet at a
        device-wide barrier after every stage."""
        tile = (self.desc.block or 256) * (self.desc.cells_per_thread or 4)
        ntiles = (self.m + tile - 1) // tile
        cap = np.minimum(max_team, ntiles)
        if self.desc.grid > 0:
            cap = np.minimum(cap, self.desc.grid)
        if team_cells is None:
            team = self._work_teams(cap)
        elif team_cells:
            team = np.clip((self.m + team_cells - 1) // team_cells, 1, cap).astype(np.int32)
        else:
            team = np.ones(self.ninst, dtype=np.int32)
        self.team = team
        if np.all(team == 1):
            self.desc.team = None
            return
        slot_inst = np.repeat(order, team[order]).astype(np.int32)
        slot_rank = np.concatenate([np.arange(team[i]) for i in order]).astype(np.int32)
        dev = self.device
        self._team = torch.tensor(team, dtype=torch.int32, device=dev)
        self._slot_inst = torch.tensor(slot_inst, dtype=torch.int32, device=dev)
        self._slot_rank = torch.tensor(slot_rank, dtype=torch.int32, device=dev)
        self._team_state = torch.zeros(self.ninst * _lib.TEAM_STATE_BYTES, dtype=torch.uint8,
                                       device=dev)
        self.desc.team = self._team.data_ptr()
        self.desc.slot_inst = self._slot_inst.data_ptr()
        self.desc.slot_rank = self._slot_rank.data_ptr()
        self.desc.nslots = int(slot_inst.size)
        self.desc.max_team = int(team.max())
        self.desc.team_state = self._team_state.data_ptr()
        self.desc.team_wait_s = float(wait_s)

    @property
    def ninst(self) -> int:
        return int(self.ncells.size)

    def state(self, i: int) -> torch.Tensor:
        """View of instance ``i`` (``ncells[i] + 2 g`` cells)."""
        off = int(self.offsets[i])
        return self.u[off: off + int(self.m[i])]

    def set_states(self, states: Sequence[np.ndarray | torch.Tensor]) -> None:
        """Copy initial states (each of ``ncells[i] + 2 g`` cells) into the
        packed array with one host-to-device transfer."""
        np_dtype = np.float64 if self.dtype == torch.float64 else np.float32
        host = np.zeros(self.total, dtype=np_dtype)
        for i, s in enumerate(states):
            s = s.detach().cpu().numpy() if isinstance(s, torch.Tensor) else np.asarray(s)
            if s.size != self.m[i]:
                raise ValueError(f"instance {i}: expected {self.m[i]} cells, got {s.size}")
            off = int(self.offsets[i])
            host[off: off + s.size] = s
        self.u.copy_(torch.from_numpy(host))

    def set_tfinal(self, tfinal: float | Sequence[float]) -> None:
        tf = np.array(np.broadcast_to(np.asarray(tfinal, dtype=np.float64), (self.ninst,)))
        self._tfinal.copy_(torch.from_numpy(tf))

    def launch(self) -> None:
        """Enqueue the batched solve on the current stream (no sync)."""
        _lib.call("psa_batch_solve", ctypes.byref(self.desc), self.u.data_ptr(),
                  self.k1.data_ptr(), self.k2.data_ptr(), self._counter.data_ptr(),
                  self.t_out.data_ptr(), self.steps_out.data_ptr(),
                  self.status_out.data_ptr(), _lib.stream_of(self.u))

    def kernel_info(self) -> dict[str, int]:
        """Registers per lane, private (spill) bytes per lane and static LDS
        bytes per workgroup of the kernel :meth:`launch` runs."""
        vgprs = ctypes.c_int(0)
        private = ctypes.c_longlong(0)
        lds = ctypes.c_longlong(0)
        _lib.call("psa_batch_kernel_info", ctypes.byref(self.desc), ctypes.byref(vgprs),
                  ctypes.byref(private), ctypes.byref(lds))
        return {"vgprs": int(vgprs.value), "private_bytes": int(private.value),
                "lds_bytes": int(lds.value)}

    def solve(self, tfinal: float | Sequence[float]) -> BatchResult:
        """Advance every instance from t = 0 to ``tfinal`` (per instance)."""
        self.set_tfinal(tfinal)
        self.launch()
        return self.result()

    def result(self) -> BatchResult:
        return BatchResult(t=self.t_out.cpu().numpy(), steps=self.steps_out.cpu().numpy(),
                           status=self.status_out.cpu().numpy())

    def capture(self, u0: torch.Tensor, tfinal: float | Sequence[float]) -> torch.cuda.CUDAGraph:
        """A graph that restores the packed initial state ``u0`` (a device
        copy, no memset) and runs the batched solve; replay it with
        ``graph.replay()``."""
        if u0.shape != self.u.shape or u0.dtype != self.dtype:
            raise ValueError("u0 must be a packed state of this batch")
        graph = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        # warm up outside capture (a zero-length solve) so that lazy
        # initialisation is not recorded
        self.set_tfinal(0.0)
        with torch.cuda.stream(stream):
            self.u.copy_(u0)
            self.launch()
        torch.cuda.current_stream(self.device).wait_stream(stream)
        torch.cuda.synchronize(self.device)
        self.set_tfinal(tfinal)
        with torch.cuda.graph(graph):
            self.u.copy_(u0)
            self.launch()
        return graph

    def cell_updates(self) -> float:
        """Sum over instances of interior cells x steps of the last solve."""
        return float(np.sum(self.ncells * self.steps_out.cpu().numpy()))


# {{{ the uneven-size workload


def random_instances(ninst: int, *, nmin: int = 128, nmax: int = 65536,
                     seed: int = 0) -> tuple[np.ndarray, list]:
    """``ninst`` sizes drawn log-uniformly from ``[nmin, nmax]`` and initial
    conditions ``c + sum_{k=1..8} A_k sin(2 pi k x + phi_k)`` with
    ``c ~ U[0, 1]``, ``A_k ~ U[-1, 1] / k`` and ``phi_k ~ U[0, 2 pi)``."""
    from pyshocks_amd.funcs import SineStep

    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=ninst))).astype(np.int64)
    ics = []
    for _ in range(ninst):
        c = rng.uniform(0.0, 1.0)
        amps = rng.uniform(-1.0, 1.0, size=8) / np.arange(1, 9)
        phis = rng.uniform(0.0, 2.0 * np.pi, size=8)
        ics.append(SineStep(c=float(c), amplitudes=tuple(float(x) for x in amps),
                            wavenumbers=tuple(range(1, 9)),
                            phases=tuple(float(x) for x in phis)))
    return ns, ics


def exact_averages_np(ic, n: int, nghosts: int = 3) -> np.ndarray:
    """Exact cell averages of a sine-sum IC on a uniform periodic grid of
    ``n`` cells on ``[0, 1]`` (ghosts included), in FP64 on the host."""
    h = 1.0 / n
    f = h * np.arange(-nghosts, n + nghosts + 1, dtype=np.float64)
    w = ic.c * f
    for a, k, p in zip(ic.amplitudes, ic.wavenumbers, ic.phases):
        w = w + a * (np.cos(p) - np.cos(2.0 * np.pi * k * f + p)) / (2.0 * np.pi * k)
    return (w[1:] - w[:-1]) / h


# }}}
