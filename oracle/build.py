"""Builds the oracle's C restatements (bp_oracle.c: FBP back-projection;
fp_oracle.c: forward projection and IR back-projection) into
oracle/libbp_oracle.so.

TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline), never part of
the product.  The reference itself is Python + Taichi (not installed here)
and has no C/C++ sources, so there is no oracle/_ref build.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRCS = [HERE / "bp_oracle.c", HERE / "fp_oracle.c"]
LIB = HERE / "libbp_oracle.so"


def build() -> Path:
    cmd = ["gcc", "-O3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fPIC", "-shared", *map(str, SRCS), "-o", str(LIB), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{r.stderr}")
    return LIB


if __name__ == "__main__":
    print(build())
