"""Checkpoints keyed by step index.

Mirrors ``checkpointing.py`` of pyshocks: ``Checkpoint`` (``:36-72``),
``save`` / ``load`` (``:75-95``), ``InMemoryCheckpoint`` (``:101-140``),
including its ``KeyError`` on overwrite and on a missing index, and the
on-disk ``PickleCheckpoint`` (``:149-203``) and ``NumpyCheckpoint``
(``:209-254``) with the reference's file names
(``<dirname>/<basename>_<index:09d>.npz``) and contents (the saved dict,
e.g. ``{m, t, u}``). On disk, ``save`` raises ``KeyError`` when the
checkpoint already exists (the reference's ``save`` tests the opposite
condition and so refuses every new file). Device tensors are written as
host arrays and loaded back onto ``device`` (0-d values stay 0-d arrays, as
``jnp.load`` returns them), so a forward run in one process can feed the
adjoint in another.
"""

from __future__ import annotations

import pathlib

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from functools import singledispatch
from typing import Any, Hashable

import numpy as np
import torch


@dataclass(frozen=True)
class Checkpoint(ABC):
    basename: str

    @abstractmethod
    def index_to_key(self, i: int) -> Hashable:
        """Key of checkpoint ``i``."""

    @abstractmethod
    def __contains__(self, i: int) -> bool:
        pass


@singledispatch
def save(chk: Checkpoint, idx: int, values: dict[str, Any]) -> None:
    raise NotImplementedError(type(chk).__name__)


@singledispatch
def load(chk: Checkpoint, idx: int, *,
         include: tuple[str, ...] | None = None) -> dict[str, Any]:
    raise NotImplementedError(type(chk).__name__)


@dataclass(frozen=True)
class InMemoryCheckpoint(Checkpoint):
    """All checkpoints in a dictionary (device tensors stay on the device)."""

    storage: dict[Hashable, dict[str, Any]] = field(default_factory=dict)

    def index_to_key(self, i: int) -> Hashable:
        return (self.basename, i)

    def __contains__(self, i: int) -> bool:
        return self.index_to_key(i) in self.storage


@save.register(InMemoryCheckpoint)
def save_in_memory(chk: InMemoryCheckpoint, idx: int, values: dict[str, Any]) -> None:
    key = chk.index_to_key(idx)
    if key in chk.storage:
        raise KeyError(f"Cannot set existing checkpoint at {idx!r}.")

    chk.storage[key] = values


@load.register(InMemoryCheckpoint)
def load_in_memory(chk: InMemoryCheckpoint, idx: int, *,
                   include: tuple[str, ...] | None = None) -> dict[str, Any]:
    key = chk.index_to_key(idx)
    if key not in chk.storage:
        raise KeyError(f"Cannot find checkpoint at index {idx!r}.")

    if include is None:
        return chk.storage[key]

    return {k: v for k, v in chk.storage[key].items() if k in include}


# {{{ on disk


def _to_host(values: dict[str, Any]) -> dict[str, np.ndarray]:
    out = {}
    for k, v in values.items():
        if isinstance(v, torch.Tensor):
            v = v.detach().cpu().numpy()
        out[k] = np.asarray(v)
    return out


def _from_host(values: dict[str, np.ndarray], device: Any) -> dict[str, Any]:
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    out: dict[str, Any] = {}
    for k, v in values.items():
        out[k] = v if v.ndim == 0 else torch.from_numpy(np.array(v)).to(device)
    return out


def _select(values: dict[str, Any], include: tuple[str, ...] | None) -> dict[str, Any]:
    if include is None:
        return values
    return {k: v for k, v in values.items() if k in include}


@dataclass(frozen=True)
class NumpyCheckpoint(Checkpoint):
    """One uncompressed ``npz`` file per checkpoint (:func:`numpy.savez`)."""

    dirname: pathlib.Path
    device: Any = None
    """Device of the loaded arrays (``None``: the GPU if there is one)."""

    def index_to_key(self, i: int) -> pathlib.Path:
        return pathlib.Path(self.dirname) / f"{self.basename}_{i:09d}.npz"

    def __contains__(self, i: int) -> bool:
        return self.index_to_key(i).exists()


@save.register(NumpyCheckpoint)
def save_npz(chk: NumpyCheckpoint, idx: int, values: dict[str, Any]) -> None:
    filename = chk.index_to_key(idx)
    if filename.exists():
        raise KeyError(f"Cannot set existing checkpoint at {idx!r}.")

    np.savez(filename, **_to_host(values))


@load.register(NumpyCheckpoint)
def load_npz(chk: NumpyCheckpoint, idx: int, *,
             include: tuple[str, ...] | None = None) -> dict[str, Any]:
    filename = chk.index_to_key(idx)
    if not filename.exists():
        raise KeyError(f"Cannot find checkpoint at index {idx!r}.")

    with np.load(filename, allow_pickle=False) as f:
        values = {k: f[k] for k in f.files}
    return _select(_from_host(values, chk.device), include)


@dataclass(frozen=True)
class PickleCheckpoint(Checkpoint):
    """One :mod:`lzma`-compressed :mod:`pickle` per checkpoint, as the
    reference writes them. Loading unpickles: use it only for files this
    class wrote."""

    dirname: pathlib.Path
    device: Any = None

    def index_to_key(self, i: int) -> pathlib.Path:
        return pathlib.Path(self.dirname) / f"{self.basename}_{i:09d}.npz"

    def __contains__(self, i: int) -> bool:
        return self.index_to_key(i).exists()


@save.register(PickleCheckpoint)
def save_pickle(chk: PickleCheckpoint, idx: int, values: dict[str, Any]) -> None:
    filename = chk.index_to_key(idx)
    if filename.exists():
        raise KeyError(f"Cannot set existing checkpoint at {idx!r}.")

    import lzma
    import pickle

    with lzma.open(filename, "wb") as outfile:
        pickle.dump(_to_host(values), outfile)


@load.register(PickleCheckpoint)
def load_pickle(chk: PickleCheckpoint, idx: int, *,
                include: tuple[str, ...] | None = None) -> dict[str, Any]:
    filename = chk.index_to_key(idx)
    if not filename.exists():
        raise KeyError(f"Cannot find checkpoint at index {idx!r}.")

    import lzma
    import pickle

    with open(filename, "rb") as infile:
        values = pickle.loads(lzma.decompress(infile.read()))
    return _select(_from_host(dict(values), chk.device), include)


# }}}
