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
