"""Device reductions of the library: interior norms (``grid.py:445-506``)
and ``max |u_int|`` (the ``smax`` / ``amax`` of ``predict_timestep``)."""

from __future__ import annotations

import math
from typing import Any

import torch

from pyshocks_amd import _lib

_NORM_CODES = {1: 1, 2: 2, "inf": 3, math.inf: 3, "tvd": 4}

_WORKSPACES: dict[torch.device, torch.Tensor] = {}


def _workspace(device: torch.device) -> torch.Tensor:
    ws = _WORKSPACES.get(device)
    if ws is None:
        nbytes = _lib.load().psa_norm_workspace_bytes()
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _WORKSPACES[device] = ws
    return ws


def interior_norm(u: torch.Tensor, lo: int, hi: int, *, p: Any = 1,
                  weight: float = 1.0) -> torch.Tensor:
    """``_norm(u[lo:hi], weight, p)`` as a 0-dim tensor of ``u.dtype``."""
    code = _NORM_CODES.get(p)
    if code is None:
        if isinstance(p, (int, float)) and p <= 0:
            raise ValueError(f"'p' must be a positive float: {p!r} <= 0.")
        raise NotImplementedError(f"norm order {p!r} is not supported")

    _lib.require_device(u)
    out = torch.empty((), dtype=u.dtype, device=u.device)
    ws = _workspace(u.device)
    _lib.call("psa_interior_norm", _lib.dtype_code(u.dtype), code, _lib.ptr(u),
              lo, hi, float(weight), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
              _lib.stream_of(u))
    return out


def max_abs_interior(u: torch.Tensor, nghosts: int) -> torch.Tensor:
    """``max |u[g:m-g]|`` as a 0-dim tensor of ``u.dtype``."""
    _lib.require_device(u)
    out = torch.empty((), dtype=u.dtype, device=u.device)
    _lib.call("psa_max_abs_interior", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(out), u.numel(), nghosts, _lib.stream_of(u))
    return out
