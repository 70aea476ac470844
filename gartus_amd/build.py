"""Build step for the HIP path tracer library (gfx950) and the CPU oracle.

``python -m gartus_amd.build`` compiles ``gartus_amd/csrc`` into
``gartus_amd/libgartus_hip.so`` with hipcc, in-tree, so the library travels
with the source tree. Each translation unit is compiled separately and
linked once; nothing is cached between calls.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libgartus_hip.so"
ARCH = os.environ.get("GARTUS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops off: no v_pk_* FP32 forms are generated.
# -ffp-contract=on: a multiply and an add fuse only within one source
# expression, so a function's results do not depend on the code it is
# inlined into (a batched and a single render of a frame stay bit-identical).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=on",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_library(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    obj_dir = PKG / "_objs"
    obj_dir.mkdir(exist_ok=True)
    units = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = []
    for src in units:
        obj = obj_dir / (src.name + ".o")
        if src.suffix == ".hip":
            cmd = [hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
        else:
            # host-only units still include hip_runtime.h for the runtime API
            cmd = [hipcc, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
        jobs.append((cmd, obj))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for _, o in jobs]])
    os.replace(tmp, LIB)
    for _, o in jobs:
        o.unlink(missing_ok=True)
    try:
        obj_dir.rmdir()
    except OSError:
        pass
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    """Compiles the CPU oracle (test infrastructure) with its Makefile."""
    oracle = ROOT / "oracle"
    proc = subprocess.run(["make", "-s", "-B", "-C", str(oracle)], capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("oracle build failed")
    if verbose:
        print(f"built {oracle / 'libcpu_ref.so'}")
    return oracle / "libcpu_ref.so"


if __name__ == "__main__":
    build_library(verbose=True)
    build_oracle(verbose=True)
