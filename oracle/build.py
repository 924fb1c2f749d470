"""Build the C restatement of the oracle (``weno_ref.c``) into
``oracle/libweno_ref.so``.

TEST INFRASTRUCTURE ONLY: the library is the CPU checker and the timed CPU
baseline; the product never links it. ``-ffp-contract=off`` keeps the
reference's rounding (no fused multiply-add); ``x86-64-v3`` keeps the binary
portable to the host CPUs of the GPU machines.
"""

from __future__ import annotations

import os
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parent
SRC = ROOT / "weno_ref.c"
LIB = ROOT / "libweno_ref.so"

CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
          "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall"]


def build(*, verbose: bool = True) -> pathlib.Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, str(SRC), "-o", str(tmp), "-lm"]
    if verbose:
        print("[oracle.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
