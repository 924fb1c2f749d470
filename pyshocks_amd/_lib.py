"""ctypes binding of ``libpyshocks_amd.so`` (declared in
``include/pyshocks_amd.h``).

The library binds to the HIP runtime that torch has already loaded (both use
the SONAME ``libamdhip64.so.7``), so torch's streams and device pointers are
valid arguments. There is no CPU fallback: every entry point raises if the
library is missing or a tensor is not on the GPU.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
from typing import Any

import torch

LIBPATH = pathlib.Path(__file__).resolve().parent / "libpyshocks_amd.so"

# status codes
PSA_OK = 0
PSA_ERR_INVALID_ARGUMENT = -1
PSA_ERR_NOT_IMPLEMENTED = -2
PSA_ERR_HIP = -3
PSA_ERR_NONFINITE_DT = -4
PSA_ERR_WORKSPACE = -5

PSA_FLOAT64 = 0
PSA_FLOAT32 = 1

PSA_FLUX_RUSANOV = 0
PSA_FLUX_BURGERS_UPWIND = 1
PSA_FLUX_ADVECTION_UPWIND = 2
PSA_FLUX_ENGQUIST_OSHER = 3
PSA_FLUX_LAX_FRIEDRICHS = 4
PSA_FLUX_ESWENO = 5

PSA_REC_WENOJS53 = 0
PSA_REC_WENOJS32 = 1
PSA_REC_ESWENO32 = 2
PSA_REC_MUSCL = 3
PSA_REC_MUSCLS = 4

PSA_BC_PERIODIC = 0
PSA_BC_GIVEN = 1
PSA_BC_NEUMANN = 2
PSA_BC_DIRICHLET = 3

(PSA_LIMITER_MINMOD, PSA_LIMITER_MC, PSA_LIMITER_SUPERBEE, PSA_LIMITER_VANALBADA1,
 PSA_LIMITER_VANALBADA2, PSA_LIMITER_VANLEER, PSA_LIMITER_KOREN,
 PSA_LIMITER_NONE) = range(8)

#: flux codes of the batched solver's Burgers scheme names
BATCH_FLUX = {"rusanov": PSA_FLUX_RUSANOV, "godunov": PSA_FLUX_BURGERS_UPWIND,
              "eo": PSA_FLUX_ENGQUIST_OSHER, "lf": PSA_FLUX_LAX_FRIEDRICHS}

PSA_DT_CFL = 0
PSA_DT_FIXED = 1


class NonFiniteTimestepError(ValueError):
    """Raised for the ``Time step is not finite`` condition of
    ``timestepping.py:144-145``."""


class PsaControl(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("tfinal", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("cfl_num", ctypes.c_double),
        ("fixed_dt", ctypes.c_double),
        ("m", ctypes.c_longlong),
        ("maxit", ctypes.c_longlong),
        ("umax_bits", ctypes.c_ulonglong),
        ("active", ctypes.c_int),
        ("status", ctypes.c_int),
        ("dt_mode", ctypes.c_int),
        ("speed_from_state", ctypes.c_int),
        ("seq_t", ctypes.c_double * 2),
        ("seq_m", ctypes.c_longlong * 2),
        ("seq_status", ctypes.c_int * 2),
        ("umax_slot", (ctypes.c_ulonglong * 16) * 64),
        ("speed_slot", ((ctypes.c_ulonglong * 16) * 64) * 2),
        ("step_slot", ((ctypes.c_ulonglong * 4) * 64) * 3),
    ]


class PsaStaticResult(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("m", ctypes.c_longlong),
        ("status", ctypes.c_int),
        ("pad_", ctypes.c_int),
    ]


#: m limit of psa_ssprk33_solve_static (cells of either dtype)
PSA_STATIC_MAX_CELLS = 1 << 20


class PsaSolverDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("n", ctypes.c_int64),
        ("g", ctypes.c_int),
        ("velocity_is_const", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("velocity_const", ctypes.c_double),
        ("velocity", ctypes.c_void_p),
        ("nu", ctypes.c_void_p),
        ("block", ctypes.c_int),
        ("cells_per_thread", ctypes.c_int),
        ("mlp", ctypes.c_void_p),
        ("mlp_tau", ctypes.c_double),
        ("mlp_sigma", ctypes.c_double),
        ("schedule", ctypes.c_int),
        ("rec", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("pad2_", ctypes.c_int),
        ("limiter_theta", ctypes.c_double),
        ("muscl_atol", ctypes.c_double),
        ("bc", ctypes.c_int),
        ("pad3_", ctypes.c_int),
        ("es_delta", ctypes.c_double),
        ("bc_left", ctypes.c_double),
        ("bc_right", ctypes.c_double),
        ("bc_values", ctypes.c_void_p),
    ]


class PsaBatchDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("g", ctypes.c_int),
        ("grid", ctypes.c_int),
        ("ninst", ctypes.c_longlong),
        ("offsets", ctypes.c_void_p),
        ("ncells", ctypes.c_void_p),
        ("dx", ctypes.c_void_p),
        ("tfinal", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("eps", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("max_steps", ctypes.c_longlong),
        ("block", ctypes.c_int),
        ("cells_per_thread", ctypes.c_int),
        ("mlp", ctypes.c_void_p),
        ("mlp_tau", ctypes.c_double),
        ("mlp_sigma", ctypes.c_double),
        ("time_f64", ctypes.c_int),
        ("max_team", ctypes.c_int),
        ("team", ctypes.c_void_p),
        ("slot_inst", ctypes.c_void_p),
        ("slot_rank", ctypes.c_void_p),
        ("nslots", ctypes.c_longlong),
        ("team_state", ctypes.c_void_p),
        ("team_wait_s", ctypes.c_double),
        ("schedule", ctypes.c_int),
        ("rec", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("pad2_", ctypes.c_int),
        ("limiter_theta", ctypes.c_double),
        ("muscl_atol", ctypes.c_double),
    ]


_VP = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_LL = ctypes.c_longlong
_D = ctypes.c_double

# name -> argtypes (restype is always int)
SIGNATURES: dict[str, list[Any]] = {
    "psa_version": [],
    "psa_control_bytes": [],
    "psa_get_last_error": [ctypes.c_char_p, ctypes.c_size_t],
    "psa_device_count": [ctypes.POINTER(ctypes.c_int)],
    "psa_weno_smoothness": [_I, _VP, _VP, _I64, _VP],
    "psa_weno_interp": [_I, _VP, _VP, _I64, _VP],
    "psa_weno_js_weights": [_I, _VP, _VP, _I64, _D, _VP],
    "psa_weno_reconstruct": [_I, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_reconstruct": [_I, _I, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_muscl_reconstruct": [_I, _I, _D, _D, _VP, _VP, _VP, _I64, _VP],
    "psa_muscls_reconstruct": [_I, _I, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_esweno_dissipation": [_I, _VP, _VP, _I64, _D, _D, _I, _VP],
    "psa_flux_divergence": [_I, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_esweno_advection_flux": [_I, _VP, _VP, _VP, _VP, _I64, _D, _D, _VP],
    "psa_apply_boundary_periodic": [_I, _VP, _VP, _I64, _I, _VP],
    "psa_apply_boundary_dirichlet": [_I, _VP, _VP, _I64, _I, _VP, _VP, _VP],
    "psa_apply_boundary_neumann": [_I, _VP, _VP, _I64, _I, _D, _D, _D, _I, _VP],
    "psa_numerical_flux": [_I, _I, _VP, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_apply_operator": [_I, _I, _VP, _VP, _VP, _VP, _I64, _I, _D, _D, _VP],
    "psa_numerical_flux_rec": [_I, _I, _I, _VP, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_apply_operator_rec": [_I, _I, _I, _VP, _VP, _VP, _VP, _I64, _I, _D, _D,
                               _VP],
    "psa_max_abs_interior": [_I, _VP, _VP, _I64, _I, _VP],
    "psa_norm_workspace_bytes": [],
    "psa_interior_norm": [_I, _I, _VP, _I64, _I64, _D, _VP, _VP, ctypes.c_size_t,
                          _VP],
    "psa_control_init": [ctypes.POINTER(PsaSolverDesc), _VP, _VP, _D, _D, _LL,
                         _I, _D, _D, _D, _VP],
    "psa_numerical_flux_desc": [ctypes.POINTER(PsaSolverDesc), _VP, _VP, _VP, _VP],
    "psa_ssprk33_steps": [ctypes.POINTER(PsaSolverDesc), _VP, _VP, _VP, _VP,
                          _VP, _LL, _VP],
    "psa_ssprk33_advance": [ctypes.POINTER(PsaSolverDesc), _VP, _VP, _VP, _VP,
                            _VP, _VP],
    "psa_ssprk33_stage": [ctypes.POINTER(PsaSolverDesc), _I, _VP, _VP, _VP, _VP,
                          _VP, _VP],
    "psa_batch_solve": [ctypes.POINTER(PsaBatchDesc), _VP, _VP, _VP, _VP, _VP, _VP,
                        _VP, _VP],
    "psa_batch_kernel_info": [ctypes.POINTER(PsaBatchDesc), ctypes.POINTER(_I),
                              ctypes.POINTER(_LL), ctypes.POINTER(_LL)],
    "psa_batch_resident_workgroups": [ctypes.POINTER(PsaBatchDesc), ctypes.POINTER(_I)],
    "psa_ssprk33_stage_vjp": [ctypes.POINTER(PsaSolverDesc), _I, _VP, _VP, _VP,
                              _VP, _VP, _VP, _VP],
    "psa_ssprk33_step_vjp": [ctypes.POINTER(PsaSolverDesc), _VP, _VP, _VP, _VP,
                             ctypes.c_size_t, _VP, _VP],
    "psa_adjoint_workspace_bytes": [ctypes.POINTER(PsaSolverDesc), _LL, _I,
                                    ctypes.POINTER(ctypes.c_size_t)],
    "psa_ssprk33_forward_checkpointed": [ctypes.POINTER(PsaSolverDesc), _VP, _VP,
                                         _LL, _VP, _I, _VP, ctypes.c_size_t, _VP],
    "psa_ssprk33_adjoint": [ctypes.POINTER(PsaSolverDesc), _LL, _VP, _I, _VP, _VP,
                            _VP, ctypes.c_size_t, _VP],
    "psa_adjoint_workspace_bytes_batch": [ctypes.POINTER(PsaSolverDesc), _LL, _LL, _LL,
                                          _I, ctypes.POINTER(ctypes.c_size_t)],
    "psa_ssprk33_forward_checkpointed_batch": [ctypes.POINTER(PsaSolverDesc), _LL, _LL,
                                               _VP, _VP, _LL, _VP, _I, _VP,
                                               ctypes.c_size_t, _VP],
    "psa_ssprk33_adjoint_batch": [ctypes.POINTER(PsaSolverDesc), _LL, _LL, _LL, _VP, _I,
                                  _VP, _VP, _VP, ctypes.c_size_t, _VP],
    "psa_mlp_weights": [_VP, _VP, _LL, _VP, _VP, _VP],
    "psa_mlp_grad_workspace_bytes": [_LL, ctypes.POINTER(ctypes.c_size_t)],
    "psa_mlp_grad": [_VP, _VP, _LL, _D, _D, _VP, _VP, ctypes.c_size_t, _VP],
    "psa_mlp_adam": [_VP, _VP, _VP, _VP, _LL, _D, _D, _D, _D, _VP],
    "psa_learned_reconstruct": [_I, _VP, _D, _D, _VP, _VP, _VP, _I64, _D, _VP],
    "psa_ssprk33_solve_static": [ctypes.POINTER(PsaSolverDesc), _VP, _VP, _D, _D,
                                 _LL, _I, _D, _D, _D, _LL, _VP, _VP],
}

TEAM_STATE_BYTES = 64
PSA_SCHEDULE_DEFAULT, PSA_SCHEDULE_TILE, PSA_SCHEDULE_PIPELINE, PSA_SCHEDULE_STEP = 0, 1, 2, 3
PSA_SCHEDULE_RING = 4
PSA_RING_MAX_CELLS = 1024
MLP_NPARAM = 1347
MLP_RECORD = 12

_LIB: ctypes.CDLL | None = None


def load(path: str | os.PathLike[str] | None = None) -> ctypes.CDLL:
    """Load the library (once) and declare its signatures. Raises
    :class:`RuntimeError` if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB

    libpath = pathlib.Path(path) if path is not None else LIBPATH
    if not libpath.exists():
        raise RuntimeError(
            f"{libpath.name} is not built: run 'python -m pyshocks_amd.build' "
            "(the HIP library has no CPU fallback)")

    lib = ctypes.CDLL(str(libpath))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int

    if lib.psa_control_bytes() != ctypes.sizeof(PsaControl):
        raise RuntimeError("psa_control layout mismatch between header and binding")

    _LIB = lib
    return lib


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.psa_get_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int, what: str = "") -> None:
    """Map a status code onto the reference's exception classes."""
    if status == PSA_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if status == PSA_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == PSA_ERR_NOT_IMPLEMENTED:
        raise NotImplementedError(msg)
    if status == PSA_ERR_NONFINITE_DT:
        raise NonFiniteTimestepError(msg)
    raise RuntimeError(msg)


def call(name: str, *args: Any) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)


# {{{ tensor helpers


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return PSA_FLOAT64
    if dtype == torch.float32:
        return PSA_FLOAT32
    raise NotImplementedError(f"unsupported dtype: {dtype}")


def require_device(*tensors: torch.Tensor | None) -> None:
    """Every array handed to the library must be a contiguous GPU tensor."""
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"expected a torch.Tensor, got {type(t).__name__}")
        if not t.is_cuda:
            raise ValueError(
                "pyshocks_amd computes on the GPU only; got a tensor on "
                f"{t.device} (move it to 'cuda')")
        if not t.is_contiguous():
            raise ValueError("arrays must be contiguous")


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


# }}}
