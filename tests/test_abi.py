"""The C-ABI library loads without a GPU and exports every declared entry point."""

import ctypes
import os
import re
import subprocess

import pytest

from physik_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:pk|PHYSIK)_[A-Za-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["physik_amd.h", "PhysiKAPI.h", "physik_amd_linear.h"])
def test_library_exports_every_declared_symbol(header):
    names = declared(header)
    assert len(names) >= 2
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_bindings_cover_the_headers():
    names = set(declared("physik_amd.h")) | set(declared("PhysiKAPI.h")) | set(declared("physik_amd_linear.h"))
    bound = set(native.PK_SIGNATURES) | set(native.PHYSIK_SIGNATURES)
    assert names == bound


def test_reference_entry_points_keep_their_names():
    """The drop-in subset of include/PhysiK/API/PhysiKAPI.h (reference)."""
    subset = ["PHYSIK_CreateWorld", "PHYSIK_DestroyWorld", "PHYSIK_Step", "PHYSIK_SetSubstepCount",
              "PHYSIK_SetSolverMode", "PHYSIK_SetConjugateGradientTolerance",
              "PHYSIK_GetConjugateGradientTolerance", "PHYSIK_SetConjugateGradientMaxIterations",
              "PHYSIK_GetConjugateGradientMaxIterations", "PHYSIK_GetLastConjugateGradientIterations",
              "PHYSIK_GetLastConjugateGradientResidualNorm", "PHYSIK_DidLastConjugateGradientSolveConverge",
              "PHYSIK_SetGravity", "PHYSIK_AddNode", "PHYSIK_SetNodeFixed", "PHYSIK_IsNodeFixed",
              "PHYSIK_GenerateTetMesh", "PHYSIK_IsGeneratedTetMeshHandleValid", "PHYSIK_DestroyGeneratedTetMesh",
              "PHYSIK_GetGeneratedTetMeshVertexCount", "PHYSIK_GetGeneratedTetMeshTetCount",
              "PHYSIK_GetGeneratedTetMeshTetIndexCount", "PHYSIK_GetGeneratedTetMeshVertex",
              "PHYSIK_GetGeneratedTetMeshTetNodeIndex", "PHYSIK_CreateTetMeshComponent",
              "PHYSIK_CreateTetMeshPhysicsComponent", "PHYSIK_GetTetMeshGlobalNodeBeginIndex",
              "PHYSIK_GetTetMeshGlobalNodeCount", "PHYSIK_GetTetMeshGlobalNodeIndex", "PHYSIK_SetTetMeshMaterial",
              "PHYSIK_DestroyComponent", "PHYSIK_IsComponentHandleValid", "PHYSIK_GetTetMeshTetCount",
              "PHYSIK_GetTetMeshNodeCount", "PHYSIK_IsTetActive", "PHYSIK_SetTetActive", "PHYSIK_DeactivateTet",
              "PHYSIK_GetActiveTetCount", "PHYSIK_GetNodePosition", "PHYSIK_GetNodeVelocity",
              "PHYSIK_SetNodePosition", "PHYSIK_SetNodeVelocity"]
    assert set(subset) <= set(declared("PhysiKAPI.h"))


def test_no_gpu_means_a_clear_error():
    if native.lib().pk_device_count() > 0:
        pytest.skip("a GPU is visible")
    from physik_amd.system import System
    import numpy as np
    rest = np.eye(4, 3)
    with pytest.raises(RuntimeError, match="no GPU"):
        System(rest, [[0, 1, 2, 3]], 1.0, 0.3, 0.0, 1.0)
