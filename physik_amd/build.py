"""Builds physik_amd/lib/libphysik_amd.so for gfx950 with hipcc.

    python -m physik_amd.build           # or physik_amd.build.build()

The library holds the HIP kernels, the pk_* device ABI (include/physik_amd.h)
and the PHYSIK_* drop-in ABI (include/PhysiKAPI.h). It links only the HIP
runtime; PyTorch is not involved.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libphysik_amd.so")
SOURCES = ["csrc/pk_assembly.hip", "csrc/pk_cg.hip", "csrc/pk_system.cpp", "csrc/pk_topology.cpp",
           "csrc/physik_api.cpp", "csrc/pk_adjoint.hip", "csrc/pk_linear.hip"]
# the counter-calibration copy kernel of physik_amd/diag.py: a library of its own
DIAG_LIB = os.path.join(LIB_DIR, "libphysik_diag.so")
DIAG_SOURCES = ["csrc/pk_diag.hip"]
ARCH = os.environ.get("PHYSIK_AMD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def command(out: str = LIB, sources=None) -> list[str]:
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        # packed FP32 ops are kept off for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fvisibility=hidden",
        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
        *[os.path.join(HERE, s) for s in (SOURCES if sources is None else sources)],
        "-o", out,
    ]


def _compile(out: str, sources, verbose: bool) -> str:
    tmp = out + ".tmp"
    cmd = command(tmp, sources)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> str:
    """Builds the product library and the diagnostics library; returns the
    product library's path."""
    os.makedirs(LIB_DIR, exist_ok=True)
    _compile(DIAG_LIB, DIAG_SOURCES, verbose)
    return _compile(LIB, SOURCES, verbose)


if __name__ == "__main__":
    print(build(verbose=True))
