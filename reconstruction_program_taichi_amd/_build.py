"""Builds libmct.so (the HIP kernels and their C ABI) for gfx950 with hipcc.

Run as ``python -m reconstruction_program_taichi_amd._build``; the package
loads the library from this directory.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmct.so"
PROBE = PKG / "issue_rate_probe"
ARCH = os.environ.get("MCT_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: products are fused into FMAs only where the source says
# fmaf(); the parity-critical index arithmetic then rounds like the
# reference's FP32 expressions.  Packed-FP32 instructions are disabled.
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def kernel_source_hash(source: str) -> str:
    """Digest of what one csrc file's kernels are compiled from: the compile
    flags, that file, every csrc header and include/mct.h.  A counter
    profile records it, so a reader can tell whether the profile belongs to
    the build at hand."""
    import hashlib

    h = hashlib.sha256(" ".join(CFLAGS).encode())
    files = [CSRC / source, *sorted(CSRC.glob("*.h")), PKG.parent / "include" / "mct.h"]
    for f in files:
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [hipcc(), *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    return obj


def build_issue_probe(verbose: bool = True) -> Path:
    """The VALU issue-rate probe (csrc/probe/issue_rate.hip), a standalone
    executable beside the library; it measures which instruction forms
    gfx950 can issue two at a time (the back-projector's roofline uses it)."""
    src = CSRC / "probe" / "issue_rate.hip"
    r = subprocess.run([hipcc(), *CFLAGS, str(src), "-o", str(PROBE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    if verbose:
        print(f"built {PROBE}")
    return PROBE


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    build_issue_probe(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
