"""Build ``libpyshocks_amd.so`` (the HIP kernels and the C ABI) for gfx950.

Run as ``python -m pyshocks_amd.build``; ``__graft_entry__.build()`` calls
:func:`build`. Every ``csrc/*.hip`` file is compiled to an object of its own,
concurrently, and the objects are linked into one shared library written next
to this file, so that it is found in-tree (no install step).

``python -m pyshocks_amd.build --incremental`` recompiles only the objects
older than their source or any header of ``csrc/`` and ``include/``.
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
OBJDIR = ROOT.parent / "_build"
LIBNAME = "libpyshocks_amd.so"
LIBPATH = ROOT / LIBNAME

ARCH = os.environ.get("PYSHOCKS_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed FP32 instructions: hipcc would otherwise form them from
    # scalar float code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core results in VGPRs (no copies out of accumulation registers)
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    # multiply-adds fused within an expression only: a formula rounds the
    # same way in every kernel it is inlined into (the stage kernels and the
    # whole-step kernel agree bit for bit)
    "-ffp-contract=on",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def find_hipcc() -> str:
    hipcc = shutil.which("hipcc")
    if hipcc is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        candidate = pathlib.Path(rocm) / "bin" / "hipcc"
        if candidate.exists():
            hipcc = str(candidate)
    if hipcc is None:
        raise RuntimeError("hipcc not found (set ROCM_PATH)")
    return hipcc


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def _headers() -> list[pathlib.Path]:
    return sorted([*CSRC.glob("*.hpp"), *INCLUDE.glob("*.h")])


def _stale(obj: pathlib.Path, src: pathlib.Path, newest_header: float) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return t < src.stat().st_mtime or t < newest_header


def build(*, verbose: bool = True, extra_flags: list[str] | None = None,
          incremental: bool = False, jobs: int | None = None) -> pathlib.Path:
    """Compile every ``csrc/*.hip`` and link ``pyshocks_amd/libpyshocks_amd.so``."""
    hipcc = find_hipcc()
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no HIP sources in {CSRC}")
    OBJDIR.mkdir(exist_ok=True)
    flags = [*HIPCC_FLAGS, *(extra_flags or []), f"-I{INCLUDE}", f"-I{CSRC}"]
    newest_header = max((h.stat().st_mtime for h in _headers()), default=0.0)

    def compile_one(src: pathlib.Path) -> pathlib.Path:
        obj = OBJDIR / (src.stem + ".o")
        if incremental and not extra_flags and not _stale(obj, src, newest_header):
            return obj
        tmp = obj.with_suffix(".o.tmp")
        cmd = [hipcc, *flags, "-c", str(src), "-o", str(tmp)]
        if verbose:
            print("[pyshocks_amd.build]", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, obj)
        return obj

    workers = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        objs = list(pool.map(compile_one, srcs))

    tmp = LIBPATH.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs),
           "-o", str(tmp)]
    if verbose:
        print("[pyshocks_amd.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBPATH)
    return LIBPATH


if __name__ == "__main__":
    args = sys.argv[1:]
    inc = "--incremental" in args
    build(extra_flags=[a for a in args if a != "--incremental"], incremental=inc)
