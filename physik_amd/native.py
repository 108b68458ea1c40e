"""ctypes declarations of libphysik_amd.so (include/physik_amd.h, include/PhysiKAPI.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libphysik_amd.so")

PK_OK, PK_ERR_ARG, PK_ERR_HIP, PK_ERR_MASS, PK_ERR_NO_DEVICE, PK_ERR_REBUILD, PK_ERR_STATE = 0, 1, 2, 3, 4, 5, 6
PK_EXPORT_POSITIONS, PK_EXPORT_VELOCITIES = 1, 2
PK_FP64, PK_FP32_STORAGE, PK_FP32_UPPER, PK_FP64_UPPER = 0, 1, 2, 3
GROUP_INTEGRATED, GROUP_SOLVE_FAILED, GROUP_NONFINITE, GROUP_NO_DYNAMIC = 0, 1, 2, 3
K_ELEMENT, K_ASSEMBLE, K_RHS, K_SPMV, K_UPDATE, K_DIRECTION, K_INTEGRATE, K_CONTACT = range(8)
KERNEL_NAMES = ["element", "assemble", "rhs", "spmv", "update", "direction", "integrate", "contact"]

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class MeshDesc(C.Structure):
    _fields_ = [
        ("node_count", C.c_int), ("rest", _vp), ("tet_count", C.c_int), ("tets", _vp),
        ("young", _vp), ("poisson", _vp), ("damping", _vp), ("density", _vp),
        ("tet_active", _vp), ("node_fixed", _vp), ("node_active", _vp), ("node_group", _vp),
        ("extra_pair_count", C.c_int), ("extra_pairs", _vp),
    ]


class ComponentHandle(C.Structure):
    _fields_ = [("index", C.c_uint32), ("generation", C.c_uint32)]

    def valid(self) -> bool:
        return self.index != 0xFFFFFFFF and self.generation != 0


class PhysikMaterialDesc(C.Structure):
    _fields_ = [("density", C.c_float), ("youngModulus", C.c_float), ("poissonRatio", C.c_float),
                ("damping", C.c_float)]


# PhysikOverlapGeometryType (the reference's PhysiKAPI.h:24-31)
PHYSIK_OverlapGeometry_Unknown = 0
PHYSIK_OverlapGeometry_Tetrahedron = 1
PHYSIK_OverlapGeometry_Triangle = 2
PHYSIK_OverlapGeometry_Sphere = 3
PHYSIK_OverlapGeometry_Node = 4


class PhysikCollisionSphereOverlap(C.Structure):
    _fields_ = [("geometryType", C.c_int), ("component", ComponentHandle), ("primitiveIndex", C.c_int),
                ("node0", C.c_int), ("node1", C.c_int), ("node2", C.c_int), ("node3", C.c_int),
                ("overlappedNodeMask", C.c_int), ("overlappedNodeCount", C.c_int),
                ("sphereCenterX", C.c_float), ("sphereCenterY", C.c_float), ("sphereCenterZ", C.c_float),
                ("sphereRadius", C.c_float), ("minDistance", C.c_float)]


class LinearResult(C.Structure):
    """pk_linear_result (include/physik_amd_linear.h)."""
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("solution_valid", C.c_int), ("pad", C.c_int),
                ("residual_norm", C.c_double)]


_lib = None

PK_SIGNATURES = {
    "pk_last_error": (C.c_char_p, []),
    "pk_device_count": (C.c_int, []),
    "pk_system_create": (C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(_vp)]),
    "pk_system_destroy": (None, [_vp]),
    "pk_system_sizes": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "pk_system_tune": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pk_system_get_tune": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "pk_system_set_spmv_staging": (C.c_int, [_vp, C.c_int]),
    "pk_system_get_spmv_staging": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "pk_system_get_spmv_stage_capacity": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "pk_system_get_slice_layout": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "pk_system_set_state": (C.c_int, [_vp, _vp, _vp]),
    "pk_system_get_state": (C.c_int, [_vp, _vp, _vp]),
    "pk_system_set_state_device": (C.c_int, [_vp, _vp, _vp]),
    "pk_system_get_state_device": (C.c_int, [_vp, _vp, _vp]),
    "pk_system_tape_id": (C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    "pk_system_export_begin": (C.c_int, [_vp, C.c_int]),
    "pk_system_export_view": (C.c_int, [_vp, _vp, _vp]),
    "pk_system_export_read": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    "pk_system_set_gravity": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "pk_system_set_solver": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "pk_system_set_mode": (C.c_int, [_vp, C.c_int]),
    "pk_system_set_young": (C.c_int, [_vp, _dp]),
    "pk_system_set_fem_models": (C.c_int, [_vp, _vp]),
    "pk_system_step": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "pk_system_group_results": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "pk_system_profile": (C.c_int, [_vp, C.c_int]),
    "pk_system_kernel_time": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]),
    "pk_system_set_graphs": (C.c_int, [_vp, C.c_int]),
    "pk_system_set_connections": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "pk_system_set_spheres": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "pk_system_set_tets_active": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "pk_system_set_matrix_store": (C.c_int, [_vp, C.c_longlong]),
    "pk_system_spmv_work": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "pk_system_assemble": (C.c_int, [_vp, C.c_double]),
    "pk_system_get_matrix": (C.c_int, [_vp, _vp, _vp, _vp]),
    "pk_system_get_elements": (C.c_int, [_vp, _vp, _vp, _vp]),
    "pk_system_get_rotations": (C.c_int, [_vp, _vp, _vp]),
    "pk_system_spmv": (C.c_int, [_vp, _dp, _dp]),
    "pk_system_get_solution": (C.c_int, [_vp, _dp]),
    "pk_system_record": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "pk_system_backward": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, _vp, _vp, _vp, _vp]),
    "pk_system_backward_device": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, _vp, _vp, _vp, _vp]),
    "pk_linear_solve": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_float, C.c_int,
                                  C.POINTER(LinearResult)]),
    "pk_bsr_pcg": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_int, _vp,
                             C.POINTER(LinearResult)]),
    "pk_bsr_pcg_noalloc": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_int, _vp,
                                     C.POINTER(LinearResult)]),
    "pk_pattern_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(_vp)]),
    "pk_pattern_destroy": (None, [_vp]),
    "pk_pattern_sizes": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_longlong)]),
    "pk_pattern_copy": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}

_W = _vp
_H = ComponentHandle
_fp = C.POINTER(C.c_float)
PHYSIK_SIGNATURES = {
    "PHYSIK_CreateWorld": (_W, []),
    "PHYSIK_DestroyWorld": (None, [_W]),
    "PHYSIK_Step": (None, [_W, C.c_float]),
    "PHYSIK_SetSubstepCount": (None, [_W, C.c_int]),
    "PHYSIK_SetSolverMode": (None, [_W, C.c_int]),
    "PHYSIK_SetConjugateGradientTolerance": (None, [_W, C.c_float]),
    "PHYSIK_GetConjugateGradientTolerance": (C.c_float, [_W]),
    "PHYSIK_SetConjugateGradientMaxIterations": (None, [_W, C.c_int]),
    "PHYSIK_GetConjugateGradientMaxIterations": (C.c_int, [_W]),
    "PHYSIK_GetLastConjugateGradientIterations": (C.c_int, [_W]),
    "PHYSIK_GetLastConjugateGradientResidualNorm": (C.c_float, [_W]),
    "PHYSIK_DidLastConjugateGradientSolveConverge": (C.c_int, [_W]),
    "PHYSIK_SetGravity": (None, [_W, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_AddNode": (C.c_int, [_W, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_SetNodeFixed": (None, [_W, C.c_int, C.c_int]),
    "PHYSIK_IsNodeFixed": (C.c_int, [_W, C.c_int]),
    "PHYSIK_GenerateTetMesh": (_vp, [_vp, C.c_int, _vp, C.c_int]),
    "PHYSIK_IsGeneratedTetMeshHandleValid": (C.c_int, [_vp]),
    "PHYSIK_DestroyGeneratedTetMesh": (None, [_vp]),
    "PHYSIK_GetGeneratedTetMeshVertexCount": (C.c_int, [_vp]),
    "PHYSIK_GetGeneratedTetMeshTetCount": (C.c_int, [_vp]),
    "PHYSIK_GetGeneratedTetMeshTetIndexCount": (C.c_int, [_vp]),
    "PHYSIK_GetGeneratedTetMeshVertex": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp]),
    "PHYSIK_GetGeneratedTetMeshTetNodeIndex": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "PHYSIK_CreateTetMeshComponent": (_H, [_W, _vp]),
    "PHYSIK_CreateTetMeshPhysicsComponent": (_H, [_W, _vp, C.POINTER(PhysikMaterialDesc)]),
    "PHYSIK_GetTetMeshGlobalNodeBeginIndex": (C.c_int, [_W, _H]),
    "PHYSIK_GetTetMeshGlobalNodeCount": (C.c_int, [_W, _H]),
    "PHYSIK_GetTetMeshGlobalNodeIndex": (C.c_int, [_W, _H, C.c_int]),
    "PHYSIK_SetTetMeshMaterial": (None, [_W, _H, C.POINTER(PhysikMaterialDesc)]),
    "PHYSIK_SetTetMeshLocalCurrentPosition": (C.c_int, [_W, _H, C.c_int, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_GetTetMeshLocalCurrentPosition": (C.c_int, [_W, _H, C.c_int, _fp, _fp, _fp]),
    "PHYSIK_DestroyComponent": (None, [_W, _H]),
    "PHYSIK_IsComponentHandleValid": (C.c_int, [_W, _H]),
    "PHYSIK_GetTetMeshTetCount": (C.c_int, [_W, _H]),
    "PHYSIK_GetTetMeshNodeCount": (C.c_int, [_W, _H]),
    "PHYSIK_IsTetActive": (C.c_int, [_W, _H, C.c_int]),
    "PHYSIK_SetTetActive": (None, [_W, _H, C.c_int, C.c_int]),
    "PHYSIK_DeactivateTet": (None, [_W, _H, C.c_int]),
    "PHYSIK_GetActiveTetCount": (C.c_int, [_W, _H]),
    "PHYSIK_CreateCollisionSphereComponent": (_H, [_W, C.c_float, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_SetCollisionSphereConnectionSettings": (None, [_W, _H, C.c_float, C.c_float]),
    "PHYSIK_GetCollisionSphereConnectionSettings": (None, [_W, _H, _fp, _fp]),
    "PHYSIK_GetCollisionSphereOverlapCount": (C.c_int, [_W, _H]),
    "PHYSIK_GetCollisionSphereOverlaps": (C.c_int, [_W, _H, C.POINTER(PhysikCollisionSphereOverlap), C.c_int]),
    "PHYSIK_SetCollisionComponentKinematicTarget": (None, [_W, _H, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_AMD_GetSystemBuildCount": (C.c_int, [_W]),
    "PHYSIK_GetPointConnectionCount": (C.c_int, [_W]),
    "PHYSIK_AddPointConnection": (None, [_W, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_GetNodePosition": (None, [_W, C.c_int, _fp, _fp, _fp]),
    "PHYSIK_GetNodeVelocity": (None, [_W, C.c_int, _fp, _fp, _fp]),
    "PHYSIK_SetNodePosition": (None, [_W, C.c_int, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_SetNodeVelocity": (None, [_W, C.c_int, C.c_float, C.c_float, C.c_float]),
    "PHYSIK_AMD_SetPrecision": (C.c_int, [_W, C.c_int]),
    "PHYSIK_AMD_TryStep": (C.c_int, [_W, C.c_float]),
    "PHYSIK_AMD_GetLastStepStatus": (C.c_int, [_W]),
    "PHYSIK_AMD_CopyNodeStateF64": (C.c_int, [_W, _vp, _vp, C.c_int]),
    "PHYSIK_AMD_SetNodeStateF64": (C.c_int, [_W, _vp, _vp, C.c_int]),
    "PHYSIK_AMD_GenerateTetMeshF64": (_vp, [_vp, C.c_int, _vp, C.c_int]),
    "PHYSIK_AMD_SetTetMeshFemModel": (C.c_int, [_W, _H, C.c_int]),
    "PHYSIK_AMD_SetTetMeshYoungModuli": (C.c_int, [_W, _H, _vp, C.c_int]),
    "PHYSIK_AMD_SetFrameExport": (C.c_int, [_W, C.c_int]),
    "PHYSIK_AMD_GetNodePositions": (C.c_int, [_W, C.c_int, C.c_int, _vp]),
    "PHYSIK_AMD_GetNodeVelocities": (C.c_int, [_W, C.c_int, C.c_int, _vp]),
    "PHYSIK_AMD_GetLastError": (C.c_char_p, []),
}


def lib():
    """The loaded library. Raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m physik_amd.build`")
        handle = C.CDLL(LIB_PATH)
        for table in (PK_SIGNATURES, PHYSIK_SIGNATURES):
            for name, (res, args) in table.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().pk_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str):
    if rc != PK_OK:
        raise RuntimeError(f"{what} failed with status {rc}: {last_error()}")


def ptr(a):
    return None if a is None else a.ctypes.data


def as_f64(a, shape=None):
    out = np.ascontiguousarray(a, dtype=np.float64)
    return out if shape is None else out.reshape(shape)
