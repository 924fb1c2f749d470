"""Benchmark of the fused WENO-JS5 / Rusanov / SSP-RK3 hot path on one GPU.

Workload (the large single-instance configuration): periodic inviscid
Burgers on [0, 1], N = 2^24 cells, FP64, WENO-JS5 (eps = 1e-6, p = 2) with
the Rusanov flux, SSP-RK3 at Courant 0.5 (theta = 1), u0 = exact cell
averages of a seeded sum of 8 sinusoids (wavenumbers 1..32) plus a step. A
bench step is one SSP-RK3 time step over all N cells (with the device-side
CFL time step); the workload's 100 steps are
the default ``--steps``. ``value`` is cell-updates/s = N * steps / wall
time with the state resident in HBM.

A step runs as one whole-step kernel (the three stages of each tile in LDS,
2 words of HBM traffic per cell-update) where it applies. Also reported:
the roofline of the dominant kernel (the whole-step kernel is VALU-bound:
executed flops per cell-update from the committed rocprofv3 counters over
its event-timed duration, against the vector peak, with its HBM figures
beside them; the stage kernels' HBM roofline), and the CPU restatement of
the same step loop on the host cores, on a bounded sample.

The other Burgers fluxes run the same workload with ``--scheme godunov``,
``--scheme eo`` (Engquist-Osher, whole-step kernel) or ``--scheme lf``
(Lax-Friedrichs: the stage kernels, each reducing the global max|u| that the
next stage's flux uses), ``--rec wenojs32|esweno32`` the third-order
reconstructions, ``--rec muscl --limiter <name>`` MUSCL with a flux limiter,
``--rec muscls --limiter <name>`` MUSCL with a slope limiter and ``--scheme
esweno32`` the ES-WENO scheme (upwind flux plus its dissipation).

Usage: python bench.py [--gpus 1] [--steps 100] [--warmup 10]
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "WENO5-RK3 cell-updates/s + adjoint grads/s, learned vs plain at equal L1 err"
HBM_PEAK_GBS = 8000.0
# MI355X published vector peaks (TFLOP/s): FP64 78.6, FP32 157.3
VALU_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
# instruction issue: 256 CUs x 4 SIMDs, one wave64 VALU instruction per 4
# cycles each, at the 2.4 GHz peak engine clock
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4

# words moved per cell by each stage kernel: stage 1 reads u^n and writes
# u^(1); stages 2 and 3 read u^n and u^(k) and write the next state
STAGE_WORDS = {1: 2, 2: 3, 3: 3}
REC_NAMES = {"wenojs53": "WENO-JS5", "wenojs32": "WENO-JS3", "esweno32": "ES-WENO3",
             "muscl": "MUSCL", "muscls": "MUSCL (slope limiter)"}
# the whole-step kernel reads u^n and writes u^(n+1) (its halo re-reads, 16
# of every 1024 cells, are not algorithmic traffic)
WHOLE_STEP_WORDS = 2


def parse_args(argv: list[str] | None = None) -> argparse.Namespace:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--n", type=int, default=1 << 24)
    p.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    p.add_argument("--scheme", choices=["rusanov", "godunov", "lf", "eo", "esweno32",
                                        "advection"],
                   default="rusanov")
    p.add_argument("--rec", choices=["wenojs53", "wenojs32", "esweno32", "muscl", "muscls"],
                   default="wenojs53", help="the reconstruction (WENO-JS5 for the metric)")
    p.add_argument("--limiter", default="minmod",
                   help="MUSCL's limiter (minmod with theta 1, mc, superbee, vanleer, ...)")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--block", type=int, default=0)
    p.add_argument("--cells-per-thread", type=int, default=0)
    p.add_argument("--event-steps", type=int, default=20,
                   help="steps of the per-kernel event-timed pass")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target duration of the CPU baseline sample (0: skip)")
    p.add_argument("--workload", choices=["single", "batched", "adjoint", "learned"],
                   default="single",
                   help="single: the large single instance (default); batched: "
                        "4096 uneven Burgers instances in one persistent graph "
                        "launch; adjoint: dJ/du0 at N = 16384, 2000 steps")
    p.add_argument("--no-adjoint", action="store_true",
                   help="skip the adjoint gradients/s figure of the default run")
    p.add_argument("--no-c1", action="store_true",
                   help="skip the small-workload time-to-solution figures (N = 256, "
                        "t = 0.5) of the default run")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the 4096-instance batched figure of the default run")
    p.add_argument("--no-learned", action="store_true",
                   help="skip the learned-vs-plain equal-error figure of the default run")
    p.add_argument("--ninst", type=int, default=4096)
    p.add_argument("--schedule", choices=["default", "tile", "pipeline", "step"], default="default",
                   help="stage-kernel schedule (one workgroup per tile, or the persistent "
                        "pipelined grid)")
    p.add_argument("--learned", action="store_true",
                   help="single workload with the shipped learned weights")
    p.add_argument("--team-cells", type=int, default=None,
                   help="batched and learned workloads: cells per team member (0: no "
                        "teams; unset: the solver's default, teams sized by work)")
    p.add_argument("--max-team", type=int, default=32)
    p.add_argument("--adjoint-batch", type=int, default=64,
                   help="adjoint workload: instances per batched gradient launch (64, "
                        "the directions of the finite-difference check)")
    p.add_argument("--adjoint-every", type=int, default=0,
                   help="adjoint workload: checkpoint interval (0: one window over the "
                        "horizon, every stage state kept by the forward pass)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain device copy of N FP64 values (N * 8 "
                        "bytes read and written per launch), the reference "
                        "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = p.parse_args(argv)
    if args.scheme == "esweno32":
        # burgers/schemes.py:214-216: the scheme takes its own reconstruction
        args.rec = "esweno32"
    return args


def copy_calibration(args: argparse.Namespace) -> int:
    import torch

    a = torch.randn(args.n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(1, args.steps)):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": True, "n": args.n,
                      "bytes_read_per_launch": 8 * args.n,
                      "bytes_written_per_launch": 8 * args.n}), flush=True)
    return 0


def make_problem(args: argparse.Namespace):
    import torch

    from pyshocks_amd import funcs, make_uniform_cell_grid
    from pyshocks_amd.advection import Godunov as AdvGodunov
    from pyshocks_amd.burgers import EngquistOsher, Godunov, LaxFriedrichs, Rusanov
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33

    dtype = torch.float64 if args.dtype == "f64" else torch.float32
    grid = make_uniform_cell_grid(0.0, 1.0, args.n, nghosts=3, dtype=dtype,
                                  device="cuda")
    rec = make_rec(args)
    if args.learned:
        from pyshocks_amd.reconstruction import LearnedWENOJS53

        rec = LearnedWENOJS53(eps=1.0e-6)
    if args.scheme == "rusanov":
        scheme, theta = Rusanov(rec=rec), 1.0
    elif args.scheme == "godunov":
        scheme, theta = Godunov(rec=rec), 1.0
    elif args.scheme == "lf":
        scheme, theta = LaxFriedrichs(rec=rec), 1.0
    elif args.scheme == "eo":
        scheme, theta = EngquistOsher(rec=rec), 1.0
    elif args.scheme == "esweno32":
        from pyshocks_amd import bind
        from pyshocks_amd.burgers import ESWENO32

        scheme = bind(ESWENO32(rec=rec), grid, PeriodicBoundary())
        theta = 1.0
    else:
        vel = torch.ones(grid.n, dtype=dtype, device="cuda")
        scheme, theta = AdvGodunov(rec=rec, velocity=vel), 0.5
    stepper = make_ssprk33(scheme, grid, PeriodicBoundary(), theta=theta)
    plan = fused_plan(stepper)
    assert plan is not None
    plan.desc.block = args.block
    plan.desc.cells_per_thread = args.cells_per_thread
    plan.desc.schedule = {"default": 0, "tile": 1, "pipeline": 2, "step": 3}[args.schedule]
    u0 = funcs.cell_averages(funcs.random_sine_step(args.seed), grid)
    return grid, plan, u0, theta


def gpu_leg(args: argparse.Namespace) -> dict:
    import numpy as np
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    grid, plan, u0, theta = make_problem(args)
    n = args.n
    itemsize = 8 if args.dtype == "f64" else 4

    u = u0.clone()
    k1 = torch.empty_like(u)
    k2 = torch.empty_like(u)
    ctrl = plan.new_control(u, 0.0, float(np.finfo(plan.np_dtype).max), None)

    if args.warmup > 0:
        plan.steps(u, u, k1, k2, ctrl, args.warmup)
    torch.cuda.synchronize()

    # the timed region, bracketed by events on the launch stream as well
    r0 = torch.cuda.Event(enable_timing=True)
    r1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    r0.record()
    plan.steps(u, u, k1, k2, ctrl, args.steps)
    r1.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    region_ms = r0.elapsed_time(r1)

    c = plan.read_control(ctrl)
    if c.status != _lib.PSA_OK or c.m != args.warmup + args.steps:
        raise RuntimeError(f"step control ended early: status {c.status} m {c.m}")
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError("non-finite state after the timed steps")

    whole_step = plan.uses_step_kernel()

    # per-kernel event timing of the stage kernels on the launch stream
    nev = max(1, args.event_steps)
    events = {s: [] for s in (1, 2, 3)}
    bufs = {1: (None, u, k1), 2: (u, k1, k2), 3: (u, k2, u)}
    for _ in range(nev):
        for s in (1, 2, 3):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            un, uk, out = bufs[s]
            e0.record()
            plan.stage(s, un, uk, out, ctrl=ctrl)
            e1.record()
            events[s].append((e0, e1))
    torch.cuda.synchronize()
    stage_ms = {s: float(np.mean([a.elapsed_time(b) for a, b in events[s]]))
                for s in events}
    stage_dominant = max(stage_ms, key=lambda s: stage_ms[s])
    stage_bytes = STAGE_WORDS[stage_dominant] * n * itemsize
    stage_gbs = stage_bytes / (stage_ms[stage_dominant] * 1.0e-3) / 1.0e9
    if whole_step:
        # one whole-step launch per step (the prologue launch of the call is
        # inside the region, ~0.01 % of it)
        kernel_ms = region_ms / args.steps
        dominant = "k_step"
        algo_bytes = WHOLE_STEP_WORDS * n * itemsize
    else:
        kernel_ms = stage_ms[stage_dominant]
        dominant = f"k_stage stage {stage_dominant}"
        algo_bytes = stage_bytes
    achieved = algo_bytes / (kernel_ms * 1.0e-3) / 1.0e9

    ms_per_step = elapsed / args.steps * 1.0e3
    return {
        "value": n * args.steps / elapsed,
        "ms_per_step": ms_per_step,
        "region_ms_per_step": region_ms / args.steps,
        "whole_step": whole_step,
        "kernel": dominant,
        "kernel_ms": kernel_ms,
        "algo_bytes": algo_bytes,
        "achieved_gbs": achieved,
        "stage_ms": stage_ms,
        "stage_dominant": stage_dominant,
        "stage_achieved_gbs": stage_gbs,
        "t_final": float(c.t),
        "theta": theta,
    }


def _profile_entry(name: str, key: str):
    """An entry of a counter summary under profiles/ (rocprofv3 measurements
    of the same workload), or None."""
    prof = ROOT / "profiles" / name
    try:
        return json.loads(prof.read_text()).get(key)
    except (OSError, ValueError):
        return None


def roofline_record(args: argparse.Namespace, gpu: dict, traffic, valu) -> dict:
    """Roofline of the dominant kernel. The whole-step kernel moves 2 words
    per cell-update and is bound by the vector ALU: its executed flops (and
    VALU instructions) per cell-update come from the rocprofv3 counters in
    profiles/valu.json, timed live with the event-timed kernel duration; the
    HBM figures are reported beside them. The stage kernels are HBM-bound."""
    hbm_gbs = gpu["achieved_gbs"]
    hbm = {
        "achieved": hbm_gbs,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": hbm_gbs / HBM_PEAK_GBS,
        "algorithmic_bytes_per_launch": gpu["algo_bytes"],
        "traffic": traffic,
    }
    stage = {
        "stage_ms": gpu["stage_ms"],
        "dominant": gpu["stage_dominant"],
        "achieved": gpu["stage_achieved_gbs"],
        "frac": gpu["stage_achieved_gbs"] / HBM_PEAK_GBS,
        "unit": "GB/s",
    }
    if not gpu["whole_step"] or valu is None:
        return {"bound": "hbm", **{k: hbm[k] for k in ("achieved", "peak", "unit", "frac",
                                                        "traffic")},
                "kernel": gpu["kernel"], "kernel_ms": gpu["kernel_ms"],
                "algorithmic_bytes_per_launch": gpu["algo_bytes"], "stage_kernels": stage}
    kernel_s = gpu["kernel_ms"] * 1.0e-3
    flops = valu["flops_per_cell_update"] * args.n
    tflops = flops / kernel_s / 1.0e12
    peak = VALU_PEAK_TFLOPS[args.dtype]
    issue = valu["valu_instr_per_cell_update"] * args.n / 64.0 / kernel_s
    return {
        "bound": "valu",
        "achieved": tflops,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": tflops / peak,
        "traffic": traffic,
        "kernel": gpu["kernel"],
        "kernel_ms": gpu["kernel_ms"],
        "flops_per_cell_update": valu["flops_per_cell_update"],
        "valu_instr_per_cell_update": valu["valu_instr_per_cell_update"],
        "valu_issue_frac": issue / VALU_PEAK_WAVE_INSTR_PER_S,
        "valu_busy_profiled": valu["valu_busy"],
        "hbm": hbm,
        "stage_kernels": stage,
        "note": "whole-step kernel: 2 words of HBM traffic per cell-update (8 for the "
                "stage kernels); executed flops and VALU instructions per cell-update "
                "from the counters in profiles/valu.json",
    }


def _cpu_model() -> str:
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def c1_leg(args: argparse.Namespace, repeats: int = 5) -> dict:
    """The small plain workload: periodic Burgers on [0, 1], N = 256, FP64,
    u0 = 0.5 + sin(2 pi x) (exact cell averages), WENO-JS5 (eps 1e-6) +
    Rusanov, SSP-RK3 at Courant 0.5, to t = 0.5 (past the shock).

    GPU: time-to-solution of one persistent workgroup that runs the whole
    time loop (the batched solver with this one instance), of the
    device-resident step loop (one psa_ssprk33_steps call of 400 steps from
    the control-block set-up; the steps after t = 0.5 are no-ops) and of the
    reference-style host loop (timestepping.step, one host sync per step),
    best of ``repeats``, each final state checked against the NumPy oracle. CPU: the NumPy
    restatement (one thread) and the C restatement on one core and on all
    cores, each the whole solve."""
    import numpy as np
    import torch

    from oracle import build as oracle_build
    from oracle import exact as E
    from oracle import ref_c
    from oracle import weno_np as W
    from pyshocks_amd import _lib, make_uniform_cell_grid
    from pyshocks_amd.burgers import Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import fused_plan, make_ssprk33, step

    _lib.load()
    n, tfinal = 256, 0.5
    u0 = E.exact_cell_averages(E.c1_initial_condition(), n)
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, device="cuda")
    stepper = make_ssprk33(Rusanov(rec=WENOJS53(eps=1e-6)), grid, PeriodicBoundary(),
                           theta=1.0)
    plan = fused_plan(stepper)
    ud = torch.tensor(u0, device="cuda")
    out = torch.empty_like(ud)
    k1, k2 = torch.empty_like(ud), torch.empty_like(ud)

    def device_loop() -> tuple[float, object]:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctrl = plan.new_control(ud, 0.0, tfinal, None)
        plan.steps(ud, out, k1, k2, ctrl, 400)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ctrl

    def host_loop() -> tuple[float, object]:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev = None
        for ev in step(stepper, ud, tfinal=tfinal):
            pass
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ev

    # one persistent workgroup running the whole time loop (the batched
    # solver with one instance, time in the state dtype as the reference)
    from pyshocks_amd.batched import BatchedSolver

    pers = BatchedSolver([n], dtype=torch.float64, time_f64=False, tiles=(256, 2))
    pers.set_states([u0])
    p0 = pers.u.clone()

    def persistent() -> float:
        pers.u.copy_(p0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pers.solve(tfinal)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    persistent()
    pers_s = min(persistent() for _ in range(repeats))
    pers_u = pers.state(0).cpu().numpy()
    pers_steps = int(pers.result().steps[0])

    device_loop()
    host_loop()
    dev_s = min(device_loop()[0] for _ in range(repeats))
    _, ctrl = device_loop()
    c = plan.read_control(ctrl)
    host_s = min(host_loop()[0] for _ in range(repeats))
    _, ev = host_loop()

    ogrid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    t0 = time.perf_counter()
    ref, dts, _ = W.solve(W.Scheme("rusanov", eps=1e-6), ogrid, u0, theta=1.0,
                          tfinal=tfinal)
    numpy_s = time.perf_counter() - t0
    i = ogrid.i_
    got = out.cpu().numpy()
    rel = float(np.linalg.norm(got[i] - ref[i]) / np.linalg.norm(ref[i]))
    rel_p = float(np.linalg.norm(pers_u[i] - ref[i]) / np.linalg.norm(ref[i]))
    if c.m != len(dts) or ev.iteration != len(dts) or rel > 1e-12:
        raise RuntimeError(f"C1 parity: steps {c.m}/{ev.iteration} vs {len(dts)}, "
                           f"rel {rel:.3e}")
    if pers_steps != len(dts) or rel_p > 1e-12:
        raise RuntimeError(f"C1 parity (persistent): steps {pers_steps} vs {len(dts)}, "
                           f"rel {rel_p:.3e}")

    if not ref_c.LIB.exists():
        oracle_build.build(verbose=False)
    threads = ref_c.num_threads()
    cpu = {}
    for cores in (1, threads):
        # 256 cells are too few for a parallel loop: every operator call then
        # pays the thread team's synchronisation, and a busy host makes that
        # slow; repeat only fast runs
        ref_c.set_num_threads(cores)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            ref_c.solve("rusanov", u0, n=n, g=3, dx=1.0 / n, eps=1e-6, theta=1.0,
                        cfl_num=0.5 / n, tfinal=tfinal)
            best = min(best, time.perf_counter() - t0)
            if best > 1.0:
                break
        cpu[cores] = best
    ref_c.set_num_threads(threads)
    cu = n * len(dts)
    return {
        "config": "periodic burgers rusanov WENO-JS5 SSP-RK3, N=256, f64, Courant 0.5, "
                  "u0 = 0.5 + sin(2 pi x), t = 0.5",
        "steps": len(dts),
        "rel_l2_vs_oracle": rel,
        "gpu_persistent_ms": pers_s * 1e3,
        "gpu_device_loop_ms": dev_s * 1e3,
        "gpu_host_loop_ms": host_s * 1e3,
        "gpu_cell_updates_per_s": cu / min(pers_s, dev_s),
        "cpu_numpy_1thread_ms": numpy_s * 1e3,
        "cpu_c_1core_ms": cpu[1] * 1e3,
        f"cpu_c_{threads}cores_ms": cpu[threads] * 1e3,
        "cpu_c_cores": threads,
    }


def cpu_leg_numpy(args: argparse.Namespace, gpu: dict) -> dict:
    """The NumPy oracle (one host thread) on a bounded sample of the same
    workload: the MUSCL and ES-WENO steps, which the C restatement does not
    have."""
    import numpy as np

    from oracle import exact as E
    from oracle import weno_np as W

    n = args.n
    u0 = E.exact_cell_averages(E.large_initial_condition(seed=args.seed), n)
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    if args.scheme == "esweno32":
        eps, delta = W.es_weno_parameters(grid, np.ones_like(grid.x))
        sch = W.Scheme("esweno32", eps=eps, rec="esweno32", delta=delta)
    else:
        sch = W.Scheme(args.scheme, eps=1e-6, rec=f"{args.rec}-{args.limiter}-1.0")
    t0 = time.perf_counter()
    u, _, _ = W.solve(sch, grid, u0, theta=gpu["theta"], maxit=1)
    one = time.perf_counter() - t0
    nsteps = int(max(1, min(50, args.cpu_seconds / max(one, 1e-9))))
    t0 = time.perf_counter()
    W.solve(sch, grid, u, theta=gpu["theta"], maxit=nsteps)
    elapsed = time.perf_counter() - t0
    return {
        "value": n * nsteps / elapsed,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"NumPy restatement of the reference step loop, same workload (N={n}, "
                  f"f64), {nsteps} SSP-RK3 steps after 1 sizing step, {elapsed:.1f} s",
    }


def cpu_leg(args: argparse.Namespace, gpu: dict) -> dict | None:
    """The C restatement of the reference step loop on the host cores, on a
    bounded sample of the same workload."""
    if args.cpu_seconds <= 0:
        return None
    if args.rec in ("muscl", "muscls") or args.scheme == "esweno32":
        return cpu_leg_numpy(args, gpu)
    import numpy as np

    from oracle import build as oracle_build
    from oracle import exact as E
    from oracle import ref_c

    if not ref_c.LIB.exists():
        oracle_build.build(verbose=False)
    threads = ref_c.num_threads()
    n = args.n
    u0 = E.exact_cell_averages(E.large_initial_condition(seed=args.seed), n)
    h = 1.0 / n
    kind = args.scheme
    cfl = 0.5 * h if kind != "advection" else h
    vel = np.ones(n + 6) if kind == "advection" else None

    # one step to size the sample, then enough steps for ~cpu_seconds
    t0 = time.perf_counter()
    u, _, _, _ = ref_c.solve(kind, u0, n=n, g=3, dx=h, eps=1e-6, theta=gpu["theta"],
                             cfl_num=cfl, maxit=1, velocity=vel, rec=args.rec)
    one = time.perf_counter() - t0
    nsteps = int(max(1, min(200, args.cpu_seconds / max(one, 1e-9))))
    t0 = time.perf_counter()
    ref_c.solve(kind, u, n=n, g=3, dx=h, eps=1e-6, theta=gpu["theta"], cfl_num=cfl,
                maxit=nsteps, velocity=vel, rec=args.rec)
    elapsed = time.perf_counter() - t0
    # the same loop on one core, a few steps (~1/threads of the rate)
    steps1 = int(max(1, min(5, 0.25 * args.cpu_seconds * nsteps / max(elapsed, 1e-9)
                            / max(threads, 1))))
    ref_c.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        ref_c.solve(kind, u, n=n, g=3, dx=h, eps=1e-6, theta=gpu["theta"], cfl_num=cfl,
                    maxit=steps1, velocity=vel, rec=args.rec)
        elapsed1 = time.perf_counter() - t0
    finally:
        ref_c.set_num_threads(threads)
    return {
        "value": n * nsteps / elapsed,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"C/OpenMP restatement of the reference step loop, same workload "
                  f"(N={n}, {args.dtype if args.dtype == 'f64' else 'f64'}), "
                  f"{nsteps} SSP-RK3 steps after 1 sizing step, {elapsed:.1f} s",
        "one_core": {"value": n * steps1 / elapsed1, "unit": "cell-updates/s",
                     "cores": 1,
                     "sample": f"the same loop on one core, {steps1} steps, "
                               f"{elapsed1:.1f} s"},
        "cpu_model": _cpu_model(),
    }


def adjoint_leg(args: argparse.Namespace, repeats: int = 3) -> dict:
    """Gradients/s of J = 0.5 ||u(T) - u_target||^2 dx with respect to u0 for
    periodic Burgers Rusanov, N = 16384, FP64, 2000 SSP-RK3 steps of a fixed
    dt (Courant 0.5 on the unperturbed u0 = 0.5 + sin(2 pi x)); the forward
    pass keeps every step's stage states (one window over the horizon: 786 MB
    per instance, 50 GB for the batch of 64), so the reverse sweep recomputes
    nothing (--adjoint-every K: checkpoints every K steps and windows
    recomputed from them); u_target is the solve from u0 + a seeded
    perturbation."""
    import math

    import numpy as np
    import torch

    from pyshocks_amd import funcs, make_uniform_cell_grid
    from pyshocks_amd.adjoint import CheckpointedSolve
    from pyshocks_amd.burgers import Rusanov
    from pyshocks_amd.reconstruction import WENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import BoundOperator

    n, nsteps = 16384, 2000
    every = args.adjoint_every or nsteps
    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, device="cuda")
    src = BoundOperator(Rusanov(rec=WENOJS53(eps=1e-6)), grid, PeriodicBoundary())
    u0 = funcs.cell_averages(funcs.sine_offset(0.5), grid)
    dt = 0.5 * grid.h / float(u0[3:-3].abs().max())
    rng = np.random.default_rng(args.seed)
    pert = funcs.SineStep(c=0.0, amplitudes=tuple(0.01 * rng.uniform(-1, 1, 4)),
                          wavenumbers=(1, 2, 3, 4),
                          phases=tuple(rng.uniform(0, 2 * math.pi, 4)))
    target = CheckpointedSolve(src, u0 + funcs.cell_averages(pert, grid), nsteps, dt,
                               every).u_final
    torch.cuda.synchronize()

    def gradient() -> torch.Tensor:
        solve = CheckpointedSolve(src, u0, nsteps, dt, every)
        lam = (solve.u_final - target) * grid.h
        return solve.gradient(lam)

    gradient()  # warm up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        g = gradient()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if not bool(torch.isfinite(g).all()):
        raise RuntimeError("non-finite gradient")

    # a batch of independent gradients (instance b: u0 plus its own seeded
    # perturbation, J_b against the same target), one launch per step for all
    from pyshocks_amd.adjoint import BatchedCheckpointedSolve

    batch = args.adjoint_batch
    pb = []
    for b in range(batch):
        rb = np.random.default_rng(1000 + b)
        pb.append(funcs.cell_averages(funcs.SineStep(
            c=0.0, amplitudes=tuple(0.001 * rb.uniform(-1, 1, 4)), wavenumbers=(1, 2, 3, 4),
            phases=tuple(rb.uniform(0, 2 * math.pi, 4))), grid))
    ub = torch.stack([u0 + d for d in pb])

    def gradients() -> torch.Tensor:
        solve = BatchedCheckpointedSolve(src, ub, nsteps, dt, every)
        lam = (solve.u_final - target[None, :]) * grid.h
        return solve.gradient(lam)

    gb = gradients()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gb = gradients()
    torch.cuda.synchronize()
    elapsed_b = time.perf_counter() - t0
    if not bool(torch.isfinite(gb).all()):
        raise RuntimeError("non-finite batched gradient")

    # the same gradients through the learned reconstruction (the trained
    # weights; the adjoint of its surrogate), single and batched
    from pyshocks_amd.learned import MLPWeights
    from pyshocks_amd.reconstruction import LearnedWENOJS53

    w = MLPWeights.load()
    src_l = BoundOperator(Rusanov(rec=LearnedWENOJS53(eps=1e-6, weights=w)), grid,
                          PeriodicBoundary())
    target_l = CheckpointedSolve(src_l, u0 + funcs.cell_averages(pert, grid), nsteps, dt,
                                 every).u_final

    def gradient_l() -> torch.Tensor:
        solve = CheckpointedSolve(src_l, u0, nsteps, dt, every)
        return solve.gradient((solve.u_final - target_l) * grid.h)

    def gradients_l() -> torch.Tensor:
        solve = BatchedCheckpointedSolve(src_l, ub, nsteps, dt, every)
        return solve.gradient((solve.u_final - target_l[None, :]) * grid.h)

    gradient_l()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        gl = gradient_l()
    torch.cuda.synchronize()
    elapsed_l = time.perf_counter() - t0
    gradients_l()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gbl = gradients_l()
    torch.cuda.synchronize()
    elapsed_lb = time.perf_counter() - t0
    if not (bool(torch.isfinite(gl).all()) and bool(torch.isfinite(gbl).all())):
        raise RuntimeError("non-finite learned gradient")
    cpu = adjoint_cpu_baseline(args, u0.cpu().numpy(), target.cpu().numpy(), g.cpu().numpy(),
                               n=n, dx=grid.h, dt=dt, nsteps=nsteps, every=50)
    return {"gradients_per_s": repeats / elapsed,
            "ms_per_gradient": elapsed / repeats * 1e3,
            "batched": {"batch": batch, "gradients_per_s": batch / elapsed_b,
                        "ms_per_batch": elapsed_b * 1e3,
                        "note": "independent gradients (own u0) in one launch per step"},
            "learned": {"gradients_per_s": repeats / elapsed_l,
                        "batched_gradients_per_s": batch / elapsed_lb,
                        "tau": w.tau, "sigma": w.sigma,
                        "note": "the same gradients through LearnedWENOJS53 (trained "
                                "weights): the adjoint of the learned solver's surrogate"},
            "cpu_baseline": cpu,
            "config": f"burgers rusanov N={n} f64, {nsteps} steps fixed dt, "
                      + ("every stage state kept by the forward pass" if every >= nsteps
                         else f"checkpoints every {every}")
                      + ", J = 0.5||u(T)-u_target||^2 dx"}


def adjoint_cpu_baseline(args: argparse.Namespace, u0, target, gpu_grad, *, n: int,
                         dx: float, dt: float, nsteps: int, every: int) -> dict | None:
    """The C reverse mode of the oracle on the adjoint workload: the whole
    gradient on all host cores (checked against the GPU's), and a 200-step
    sample on one core, scaled to the full horizon (the cost is linear in
    the steps). The reference itself forms a dense n x n step Jacobian with
    jax.jacfwd (timestepping.py:155-214), O(n^2) work per step, which the
    matrix-free restatement does not reproduce."""
    if args.cpu_seconds <= 0:
        return None
    import numpy as np

    from oracle import build as oracle_build
    from oracle import ref_c

    if not ref_c.LIB.exists():
        oracle_build.build(verbose=False)
    threads = ref_c.num_threads()
    t0 = time.perf_counter()
    cgrad, _, _ = ref_c.adjoint_rusanov(u0, target, n=n, g=3, dx=dx, eps=1e-6, dt=dt,
                                        nsteps=nsteps, every=every)
    full_s = time.perf_counter() - t0
    rel = float(np.linalg.norm(cgrad[3:-3] - gpu_grad[3:-3])
                / np.linalg.norm(cgrad[3:-3]))
    sample = 200
    ref_c.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        ref_c.adjoint_rusanov(u0, target, n=n, g=3, dx=dx, eps=1e-6, dt=dt, nsteps=sample,
                              every=every)
        one_s = (time.perf_counter() - t0) * nsteps / sample
    finally:
        ref_c.set_num_threads(threads)
    return {"value": 1.0 / full_s, "unit": "gradients/s", "cores": threads, "kind": "port",
            "sample": f"C reverse mode of the oracle (forward with checkpoints every {every}, "
                      f"window recompute, three stage transposes per step), the whole "
                      f"{nsteps}-step gradient, {full_s:.1f} s",
            "gradient_rel_l2_vs_gpu": rel,
            "one_core": {"value": 1.0 / one_s, "unit": "gradients/s", "cores": 1,
                         "sample": f"{sample} steps on one core, scaled to {nsteps}"}}


def make_rec(args: argparse.Namespace):
    """The reconstruction object of --rec (with --limiter for MUSCL)."""
    from pyshocks_amd.reconstruction import make_reconstruction_from_name

    if args.rec in ("muscl", "muscls"):
        from pyshocks_amd.limiters import make_limiter_from_name

        return make_reconstruction_from_name(
            args.rec, lm=make_limiter_from_name(args.limiter, theta=1.0, variant=1))
    return make_reconstruction_from_name(args.rec, eps=1.0e-6)


def batched_leg(args: argparse.Namespace) -> dict:
    """cell-updates/s of the uneven-size batch: --ninst Burgers instances with
    N log-uniform in [128, 65536] (seeded), FP32, Rusanov, each to t = 0.3,
    one persistent launch replayed from a HIP graph, tiles autotuned with
    (unless --team-cells says otherwise) the solver's teams sized by work;
    --rec other than WENO-JS5 runs with 2 ghost cells and tiles 256 x 4 or
    256 x 2."""
    import numpy as np
    import torch

    from pyshocks_amd import _lib, autotune
    from pyshocks_amd.batched import BatchedSolver, exact_averages_np, random_instances

    _lib.load()
    rec = None if args.rec == "wenojs53" else make_rec(args)
    g = 3 if rec is None else 2
    ns, ics = random_instances(args.ninst, seed=args.seed)
    states = [exact_averages_np(ic, int(n), nghosts=g) for ic, n in zip(ics, ns)]
    cands = autotune.BATCH_CANDIDATES if rec is None else ((256, 4), (256, 2))
    # the whole batch to t = 0.01 with each candidate's teams: tiles and teams
    # are chosen together
    best, _ = autotune.tune_batch(ns, states, dtype=torch.float32, tfinal=0.01, rec=rec,
                                  candidates=cands, team_cells=args.team_cells)
    solver = BatchedSolver(ns, dtype=torch.float32, team_cells=args.team_cells,
                           max_team=args.max_team, nghosts=g, rec=rec,
                           tiles=None if rec is None else best)
    solver.set_states(states)
    u0 = solver.u.clone()
    graph = solver.capture(u0, 0.3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graph.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res = solver.result()
    if np.any(res.status != 0):
        raise RuntimeError(f"{int(np.sum(res.status != 0))} instances did not finish")
    cu = solver.cell_updates()
    teams = np.bincount(solver.team)
    return {"value": cu / elapsed, "seconds": elapsed, "cell_updates": cu,
            "tiles": list(best), "ninst": len(ns), "total_cells": int(np.sum(ns)),
            "teams": {str(k): int(c) for k, c in enumerate(teams) if k > 1 and c},
            "max_steps": int(res.steps.max()), "ns": ns, "states": states}


def batched_cpu_leg(args: argparse.Namespace, ns, states) -> dict | None:
    """The C restatement on a seeded subset of the batch's instances (FP64,
    to t = 0.3), within about --cpu-seconds."""
    if args.cpu_seconds <= 0 or args.rec != "wenojs53":
        return None
    import numpy as np

    from oracle import build as oracle_build
    from oracle import ref_c

    if not ref_c.LIB.exists():
        oracle_build.build(verbose=False)
    rng = np.random.default_rng(args.seed + 2)
    order = rng.permutation(len(ns))
    done, cells, t0 = 0, 0.0, time.perf_counter()
    for i in order:
        n = int(ns[i])
        _, _, steps, _ = ref_c.solve("rusanov", np.asarray(states[i], dtype=np.float64),
                                     n=n, g=3, dx=1.0 / n, eps=1e-6, theta=1.0,
                                     cfl_num=0.5 / n, tfinal=0.3)
        cells += n * steps
        done += 1
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    elapsed = time.perf_counter() - t0
    return {"value": cells / elapsed, "unit": "cell-updates/s", "cores": ref_c.num_threads(),
            "kind": "port",
            "sample": f"C/OpenMP restatement, {done} seeded instances of the batch "
                      f"(FP64) each to t = 0.3, {elapsed:.1f} s"}


def learned_leg(args: argparse.Namespace, *, ninst: int = 1024, nmin: int = 1024,
                nmax: int = 32768, scales: tuple[float, float] = (1.0, 1.5)) -> dict:
    """Learned-weight batched solver (shipped network) against the plain
    WENO-JS5 FP64 batched solver at equal L1 error: ``ninst`` seeded
    sine-plus-step Burgers instances with N log-uniform in [nmin, nmax],
    t = 0.5, L1 errors against exact Hopf-Lax cell averages; every solver
    with its own tiles (autotuned on the whole batch with the teams each
    tile gives it, spilling kernels skipped).

    The plain solver's error-vs-N curve is measured on the same instances
    with every N scaled by each of ``scales``; its order p fixes the scale
    s* = (E_plain / E_learned)^(1/p) at which the plain solver should match
    the learned error. The plain solver is then run at s* N (and once more
    at a secant refinement if its error misses the learned one by more than
    5 %), and the speedup at equal error is the plain wall time there over
    the learned wall time; the effective throughput is the plain
    cell-updates there per learned second."""
    import numpy as np
    import torch

    from pyshocks_amd import autotune, exact
    from pyshocks_amd.learned import MLPWeights, l1_errors, training_ics

    w = MLPWeights.load()
    rng = np.random.default_rng(5000 + args.seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), ninst))).astype(np.int64)
    ics = training_ics(ninst, seed=5000 + args.seed)
    u0s = [u.cpu().numpy() for u in exact.burgers_cell_averages(ics, ns, 0.0)]
    candidates = ((256, 4), (256, 2), (512, 2), (1024, 2))
    teams = args.team_cells

    def tune(dtype, mlp):
        # the whole batch to t = 0.01 with each candidate's teams
        tiles, _ = autotune.tune_batch(ns, u0s, dtype=dtype, tfinal=0.01, mlp=mlp,
                                       candidates=candidates, team_cells=teams)
        return tiles

    plain_tiles = tune(torch.float64, None)

    def plain(scale: float):
        nsc = np.maximum(np.round(ns * scale).astype(np.int64), 8)
        e, info = l1_errors(ics, nsc, 0.5, dtype=torch.float64, tiles=plain_tiles,
                            team_cells=teams)
        return {"scale": scale, "l1": float(e.mean()), "seconds": info["seconds"],
                "cell_updates": info["cell_updates"]}

    curve = [plain(sc) for sc in scales]
    (s1, e1), (s2, e2) = (curve[0]["scale"], curve[0]["l1"]), (curve[1]["scale"], curve[1]["l1"])
    p = float(np.log(e1 / e2) / np.log(s2 / s1))
    out = {
        "config": f"{ninst} seeded sine+step Burgers instances, N log-uniform "
                  f"[{nmin}, {nmax}], t = 0.5, learned batched (tau {w.tau}, sigma {w.sigma}) "
                  f"vs plain WENO-JS5 FP64 batched, L1 error vs exact Hopf-Lax averages",
        "plain_tiles": list(plain_tiles),
        "team_cells": teams,
        "plain_curve": curve,
        "plain_order": p,
    }
    best = None
    for name, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        tiles = tune(dtype, w)
        e_l, info_l = l1_errors(ics, ns, 0.5, mlp=w, dtype=dtype, tiles=tiles,
                                team_cells=teams)
        el = float(e_l.mean())
        rec = {"tiles": list(tiles), "l1": el, "error_ratio": e1 / el,
               "seconds": info_l["seconds"],
               "cell_updates_per_s": info_l["cell_updates"] / info_l["seconds"]}
        out[f"learned_{name}"] = rec
        if best is None or el < best[1]:
            best = (name, el, info_l["seconds"])

    # the plain solver at the scale that matches the better learned error
    name, el, tl = best
    points = sorted([(c["scale"], c["l1"]) for c in curve])
    s_star = s1 * (e1 / el) ** (1.0 / p) if p > 0 else float("nan")
    matched = []
    if np.isfinite(s_star) and 1.0 <= s_star <= 4.0:
        m = plain(s_star)
        matched.append(m)
        if abs(m["l1"] / el - 1.0) > 0.05:
            # secant in log-log through the measured point nearest in error
            pts = sorted(points + [(m["scale"], m["l1"])], key=lambda q: abs(np.log(q[1] / el)))
            (sa, ea), (sb, eb) = pts[0], pts[1]
            q = float(np.log(ea / eb) / np.log(sb / sa)) if sa != sb else p
            s_next = sa * (ea / el) ** (1.0 / q) if q > 0 else float("nan")
            if np.isfinite(s_next) and 1.0 <= s_next <= 4.0:
                matched.append(plain(s_next))
    out["plain_matched"] = matched
    if matched:
        m = min(matched, key=lambda c: abs(np.log(c["l1"] / el)))
        speedup = m["seconds"] / tl
        out.update({"best": name, "speedup_at_equal_error": speedup,
                    "effective_cell_updates_per_s": m["cell_updates"] / tl,
                    "matched_error_gap": m["l1"] / el - 1.0})
        # gradients at the same error ratio: the learned adjoint at N = 16384
        # against the plain one at the matched N
        n0 = 16384
        n_plain = int(round(m["scale"] * n0))
        g_l = _gradient_rate(n0, w)
        g_p = _gradient_rate(n_plain, None)
        out["gradients_at_equal_error"] = {
            "learned": g_l, "plain": g_p,
            "speedup": g_l["gradients_per_s"] / g_p["gradients_per_s"],
            "note": f"batched gradients (batch 32) of J = 0.5||u(T) - u_target||^2 dx over "
                    f"the horizon of 2000 steps at N = {n0}, Courant 0.5: the learned "
                    f"FP64 adjoint at N = {n0} against the plain one at N = {n_plain} "
                    f"(the scale at which the plain forward solve matches the learned "
                    f"L1 error above)"}
    else:
        out.update({"best": name, "speedup_at_equal_error": None,
                    "effective_cell_updates_per_s": None})
    return out


def _gradient_rate(n: int, weights, batch: int = 32, n_ref: int = 16384,
                   steps_ref: int = 2000, every: int = 0) -> dict:
    """Batched gradients/s of J = 0.5 ||u(T) - u_target||^2 dx for periodic
    Burgers Rusanov FP64 (WENO-JS5, or the learned weights) at N = n, over
    the horizon of steps_ref steps at N = n_ref with Courant number 0.5:
    instance b starts from 0.5 + sin(2 pi x) plus a seeded perturbation of
    amplitude 1e-3 and is compared with the solve from the unperturbed
    state. ``every`` 0 keeps every stage state in the forward pass (one
    window over the horizon)."""
    import math

    import numpy as np
    import torch

    from pyshocks_amd import funcs, make_uniform_cell_grid
    from pyshocks_amd.adjoint import BatchedCheckpointedSolve, CheckpointedSolve
    from pyshocks_amd.burgers import Rusanov
    from pyshocks_amd.reconstruction import WENOJS53, LearnedWENOJS53
    from pyshocks_amd.scalar import PeriodicBoundary
    from pyshocks_amd.timestepping import BoundOperator

    grid = make_uniform_cell_grid(0.0, 1.0, n, nghosts=3, device="cuda")
    rec = WENOJS53(eps=1e-6) if weights is None else LearnedWENOJS53(eps=1e-6, weights=weights)
    src = BoundOperator(Rusanov(rec=rec), grid, PeriodicBoundary())
    u0 = funcs.cell_averages(funcs.sine_offset(0.5), grid)
    umax = float(u0[3:-3].abs().max())
    t_end = steps_ref * 0.5 / n_ref / umax
    nsteps = int(math.ceil(t_end / (0.5 * grid.h / umax)))
    dt = t_end / nsteps
    every = every or nsteps
    target = CheckpointedSolve(src, u0, nsteps, dt, every).u_final
    ub = []
    for b in range(batch):
        rb = np.random.default_rng(2000 + b)
        ub.append(u0 + funcs.cell_averages(funcs.SineStep(
            c=0.0, amplitudes=tuple(0.001 * rb.uniform(-1, 1, 4)), wavenumbers=(1, 2, 3, 4),
            phases=tuple(rb.uniform(0, 2 * math.pi, 4))), grid))
    ub = torch.stack(ub)

    def gradients() -> torch.Tensor:
        solve = BatchedCheckpointedSolve(src, ub, nsteps, dt, every)
        return solve.gradient((solve.u_final - target[None, :]) * grid.h)

    gradients()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = gradients()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if not bool(torch.isfinite(g).all()):
        raise RuntimeError("non-finite gradient")
    return {"n": n, "nsteps": nsteps, "batch": batch, "seconds": elapsed,
            "gradients_per_s": batch / elapsed}


def main(argv: list[str] | None = None) -> int:
    args = parse_args(argv)
    if args.copy_calibration:
        return copy_calibration(args)
    if args.workload == "batched":
        b = batched_leg(args)
        cpu = batched_cpu_leg(args, b.pop("ns"), b.pop("states"))
        print(json.dumps({"metric": METRIC, "value": b["value"], "unit": "cell-updates/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
                          "data": "synthetic",
                          "config": {"workload": f"{b['ninst']} periodic Burgers Rusanov "
                                                 f"{REC_NAMES[args.rec]} instances, N "
                                                 "log-uniform [128, 65536], f32, each to "
                                                 "t = 0.3, one graph launch",
                                     "reconstruction": (
                                         f"{args.rec}-{args.limiter}"
                                         if args.rec in ("muscl", "muscls") else args.rec)},
                          "batched": b, "cpu_baseline": cpu}), flush=True)
        return 0
    if args.workload == "learned":
        lr = learned_leg(args)
        print(json.dumps({"metric": METRIC, "value": lr["effective_cell_updates_per_s"],
                          "unit": "cell-updates/s (plain FP64 cell-updates at equal L1 "
                                  "error per learned second)",
                          "n_gpus": 1, "higher_is_better": True, "dtype": lr["best"],
                          "data": "synthetic", "learned": lr}), flush=True)
        return 0
    if args.workload == "adjoint":
        a = adjoint_leg(args)
        print(json.dumps({"metric": METRIC, "value": a["gradients_per_s"],
                          "unit": "gradients/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f64", "data": "synthetic", "adjoint": a}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the workload is one GPU's instance of the periodic grid)",
              file=sys.stderr)
        return 2

    # the small workload and the adjoint first: the timed steps of the large
    # workload then start on a GPU at its running clocks
    c1 = None if args.no_c1 else c1_leg(args)
    adjoint = None if args.no_adjoint else adjoint_leg(args)
    gpu = gpu_leg(args)
    batched = batched_cpu = None
    if not args.no_batched:
        b = batched_leg(args)
        batched_cpu = batched_cpu_leg(args, b.pop("ns"), b.pop("states"))
        batched = b
    learned = None if args.no_learned else learned_leg(args)
    cpu = cpu_leg(args, gpu)

    kname = "step" if gpu["whole_step"] else f"stage{gpu['stage_dominant']}"
    recs = "" if args.rec == "wenojs53" else f"_{args.rec}"
    if args.rec in ("muscl", "muscls"):
        recs = f"_{args.rec}_{args.limiter}"
    key = f"{args.scheme}{recs}_{args.dtype}_{args.n}_{kname}"
    traffic = _profile_entry("traffic.json", key)
    valu = _profile_entry("valu.json", key)
    roofline = roofline_record(args, gpu, traffic, valu)

    dtype_name = "f64" if args.dtype == "f64" else "f32"
    line = {
        "metric": METRIC,
        "value": gpu["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_name,
        "data": "synthetic",
        "config": {
            "workload": f"periodic {'burgers' if args.scheme != 'advection' else 'advection'} "
                        f"{args.scheme} {REC_NAMES[args.rec]} SSP-RK3, N={args.n}, {dtype_name}, "
                        f"Courant 0.5, sine-sum+step u0 seed {args.seed}",
            "n_cells": args.n,
            "scheme": args.scheme,
            "reconstruction": (f"{args.rec}-{args.limiter}" if args.rec in ("muscl", "muscls")
                               else args.rec),
            "step": ("one SSP-RK3 time step (one whole-step kernel: the three stages of "
                     "each tile in LDS, device CFL dt)" if gpu["whole_step"] else
                     "one SSP-RK3 time step (3 fused stage kernels, device CFL dt)"),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "c1": c1,
        "batched": batched,
        "batched_cpu_baseline": batched_cpu,
        "adjoint": adjoint,
        "learned": learned,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
