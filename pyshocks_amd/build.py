"""Build ``libpyshocks_amd.so`` (the HIP kernels and the C ABI) for gfx950.

Run as ``python -m pyshocks_amd.build``; ``__graft_entry__.build()`` calls
:func:`build`. The library is written next to this file so that it is found
in-tree (no install step).
"""

from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
LIBNAME = "libpyshocks_amd.so"
LIBPATH = ROOT / LIBNAME

ARCH = os.environ.get("PYSHOCKS_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
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


def build(*, verbose: bool = True, extra_flags: list[str] | None = None) -> pathlib.Path:
    """Compile every ``csrc/*.hip`` into ``pyshocks_amd/libpyshocks_amd.so``."""
    hipcc = find_hipcc()
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no HIP sources in {CSRC}")

    tmp = LIBPATH.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, *(extra_flags or []), f"-I{INCLUDE}", f"-I{CSRC}",
           *map(str, srcs), "-o", str(tmp)]
    if verbose:
        print("[pyshocks_amd.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBPATH)
    return LIBPATH


if __name__ == "__main__":
    build(extra_flags=sys.argv[1:])
