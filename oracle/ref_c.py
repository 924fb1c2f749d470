"""ctypes wrapper of the C oracle ``libweno_ref.so``.

TEST INFRASTRUCTURE ONLY (tests and the CPU-baseline leg of ``bench.py``).
"""

from __future__ import annotations

import ctypes
import pathlib

import numpy as np

LIB = pathlib.Path(__file__).resolve().parent / "libweno_ref.so"

FLUX = {"rusanov": 0, "godunov": 1, "advection": 2, "eo": 3, "lf": 4}
#: reconstruction codes, passed in bits 8 and up of the flux argument
REC = {"wenojs53": 0, "wenojs32": 1, "esweno32": 2}

_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB.name} is not built: run 'python -m oracle.build'")
        lib = ctypes.CDLL(str(LIB))
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_ssprk33_run.argtypes = [
            ctypes.c_int, dp, dp, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
            ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_longlong, ctypes.c_double, ctypes.c_double,
            dp, ctypes.POINTER(ctypes.c_longlong), dp, ctypes.c_longlong]
        lib.oracle_ssprk33_run.restype = ctypes.c_int
        lib.oracle_apply_operator.argtypes = [
            ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_int64, ctypes.c_int,
            ctypes.c_double, ctypes.c_double]
        lib.oracle_apply_operator.restype = None
        lib.oracle_operator_vjp.argtypes = [dp, dp, dp, ctypes.c_int64, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double]
        lib.oracle_operator_vjp.restype = None
        lib.oracle_ssprk33_step_vjp.argtypes = [dp, dp, dp, ctypes.c_int64, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double]
        lib.oracle_ssprk33_step_vjp.restype = None
        lib.oracle_adjoint_rusanov.argtypes = [
            dp, dp, dp, dp, dp, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_longlong, ctypes.c_longlong]
        lib.oracle_adjoint_rusanov.restype = ctypes.c_int
        lib.oracle_num_threads.restype = ctypes.c_int
        lib.oracle_set_num_threads.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_num_threads(n: int) -> None:
    load().oracle_set_num_threads(int(n))


def num_threads() -> int:
    return int(load().oracle_num_threads())


def apply_operator(kind: str, u: np.ndarray, *, g: int, dx: float, eps: float,
                   velocity: np.ndarray | None = None,
                   rec: str = "wenojs53") -> np.ndarray:
    lib = load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    m = u.size
    out = np.empty(m)
    uf = np.empty(m)
    F = np.empty(m + 1)
    vel = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
    lib.oracle_apply_operator(FLUX[kind] | (REC[rec] << 8), _p(u), _p(vel), _p(out),
                              _p(uf), _p(F), m, g, dx, eps)
    return out


def solve(kind: str, u0: np.ndarray, *, n: int, g: int, dx: float, eps: float,
          theta: float = 1.0, cfl_num: float = 0.0, fixed_dt: float | None = None,
          maxit: int | None = None, tstart: float = 0.0, tfinal: float | None = None,
          velocity: np.ndarray | None = None,
          rec: str = "wenojs53") -> tuple[np.ndarray, float, int, np.ndarray]:
    """Run the reference step loop; returns ``(u, t, nsteps, dts)``."""
    lib = load()
    u = np.array(u0, dtype=np.float64, copy=True)
    vel = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
    tf = np.finfo(np.float64).max if tfinal is None else float(tfinal)
    cap = 1 << 20
    dts = np.zeros(min(cap, maxit if maxit is not None else cap))
    t = ctypes.c_double(0.0)
    it = ctypes.c_longlong(0)
    status = lib.oracle_ssprk33_run(
        FLUX[kind] | (REC[rec] << 8), _p(u), _p(vel), n, g, dx, eps,
        0 if fixed_dt is None else 1,
        theta, cfl_num, 0.0 if fixed_dt is None else fixed_dt,
        -1 if maxit is None else maxit, tstart, tf, ctypes.byref(t),
        ctypes.byref(it), _p(dts), dts.size)
    if status != 0:
        raise ValueError("Time step is not finite")
    return u, t.value, it.value, dts[: min(it.value, dts.size)]


def operator_vjp(u: np.ndarray, mu: np.ndarray, *, g: int, dx: float,
                 eps: float) -> np.ndarray:
    """``J_L(u)^T mu`` of the Rusanov WENO-JS5 operator (periodic); ghost
    entries of the result are zero."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    out = np.empty_like(u)
    load().oracle_operator_vjp(_p(u), _p(mu), _p(out), u.size, g, dx, eps)
    return out


def step_vjp(u: np.ndarray, lam: np.ndarray, *, g: int, dx: float, eps: float,
             dt: float) -> np.ndarray:
    """The transpose of one fixed-dt SSP-RK3 step's Jacobian at ``u`` applied
    to ``lam`` (Rusanov, WENO-JS5, periodic)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.empty_like(u)
    load().oracle_ssprk33_step_vjp(_p(u), _p(lam), _p(out), u.size, g, dx, eps, dt)
    return out


def adjoint_rusanov(u0: np.ndarray, target: np.ndarray, *, n: int, g: int, dx: float,
                    eps: float, dt: float, nsteps: int,
                    every: int = 50) -> tuple[np.ndarray, np.ndarray, float]:
    """Gradient of ``J = 0.5 sum_interior (u(T) - target)^2 dx`` with respect
    to ``u0`` after ``nsteps`` fixed-dt steps, with checkpoints every
    ``every`` steps; returns ``(grad, u_final, J)``."""
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    grad = np.empty_like(u0)
    uf = np.empty_like(u0)
    J = ctypes.c_double(0.0)
    status = load().oracle_adjoint_rusanov(_p(u0), _p(target), _p(grad), _p(uf),
                                           ctypes.byref(J), n, g, dx, eps, dt, nsteps,
                                           every)
    if status != 0:
        raise MemoryError("oracle_adjoint_rusanov: out of memory")
    return grad, uf, J.value
