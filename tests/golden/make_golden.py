"""Generate the golden fixtures of ``tests/golden`` from the CPU oracle.

The reference ships no golden vectors for this path, so these fixtures are
produced by the oracle (``oracle/weno_np.py``), whose own pins are in
``tests/test_oracle.py``. Run ``python tests/golden/make_golden.py`` from the
repository root; the output is deterministic for a given NumPy version.
"""

from __future__ import annotations

import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import exact as E  # noqa: E402
from oracle import weno_np as W  # noqa: E402


def weno_fields() -> dict[str, np.ndarray]:
    u = np.sin(np.linspace(0.0, 2.0 * np.pi, 64 + 6))
    rng = np.random.default_rng(3)
    r = rng.uniform(-1, 1, size=70)
    out = {"u_sine": u, "u_random": r}
    for name, v in (("sine", u), ("random", r)):
        out[f"beta_{name}"] = W.weno_smoothness(W.WENOJS53, v)
        out[f"uhat_{name}"] = W.weno_interp(W.WENOJS53, v)
        for eps in (1e-6, 1e-12):
            tag = f"{name}_eps{int(-np.log10(eps))}"
            out[f"omega_{tag}"] = W.weno_js_weights(W.WENOJS53, v, eps=eps)
            ul, ur = W.reconstruct_wenojs(v, eps)
            out[f"ul_{tag}"] = ul
            out[f"ur_{tag}"] = ur
    return out


def operators() -> dict[str, np.ndarray]:
    n = 32
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, size=grid.n)
    vel = rng.uniform(-1, 1, size=grid.n)
    out = {"u": u, "velocity": vel}
    for kind in ("rusanov", "godunov", "advection"):
        sch = W.Scheme(kind, eps=1e-6, velocity=vel)
        out[f"flux_{kind}"] = W.numerical_flux(sch, grid, u)
        out[f"op_{kind}"] = W.apply_operator(sch, grid, 0.0, u)
    out["boundary"] = W.apply_boundary_periodic(grid, u)
    return out


def burgers_c1(kind: str, n: int) -> dict[str, np.ndarray]:
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u0 = E.exact_cell_averages(E.c1_initial_condition(), n)
    sch = W.Scheme(kind, eps=1e-6)

    def predict(t, u):
        return 1.0 * W.predict_timestep(sch, grid, t, u)

    def source(t, u):
        return W.apply_operator(sch, grid, t, u)

    out = {"u0": u0}
    dts = []
    event = None
    for event in W.step(predict, source, u0, tfinal=0.5):
        if event.iteration in (1, 10):
            out[f"u{event.iteration}"] = event.u
        if event.iteration > 0:
            dts.append(float(event.dt))
    assert event is not None
    out["ufinal"] = event.u
    out["dts"] = np.array(dts)
    out["tfinal"] = np.array(float(event.t))
    return out


def advection(n: int) -> dict[str, np.ndarray]:
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    ic = E.large_initial_condition(seed=1)
    u0 = E.exact_cell_averages(ic, n)
    vel = np.ones(grid.n)
    sch = W.Scheme("advection", eps=1e-6, velocity=vel)
    u, dts, t = W.solve(sch, grid, u0, theta=0.5, tfinal=0.25)
    return {"u0": u0, "ufinal": u, "dts": np.array(dts), "tfinal": np.array(t)}


def main() -> None:
    np.savez_compressed(HERE / "weno_fields.npz", **weno_fields())
    np.savez_compressed(HERE / "operators_n32.npz", **operators())
    np.savez_compressed(HERE / "burgers_rusanov_c1.npz", **burgers_c1("rusanov", 256))
    np.savez_compressed(HERE / "burgers_godunov_n128.npz", **burgers_c1("godunov", 128))
    np.savez_compressed(HERE / "advection_n128.npz", **advection(128))
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size)


if __name__ == "__main__":
    main()
