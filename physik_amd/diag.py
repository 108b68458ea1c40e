"""Profiling harness: a calibration copy of known size, then CG work.

    python -m physik_amd.diag [--copy-mb 512] [--workload twist64] [--steps 1]
                              [--precision fp64|fp32] [--shape L,R,V,U]
                              [--spheres S]

Run under ``rocprofv3 --pmc FETCH_SIZE`` (and again with WRITE_SIZE): the copy
kernel moves exactly copy-mb MiB each way, which calibrates the counters; the
implicit steps that follow give the per-dispatch counts of every step kernel.
The JSON line it prints carries the work the SpMV launches did (blocks and
rows of the bodies still iterating), to set the counters against. With
--spheres S, S collision spheres sit along the diagonal of the mesh's bounding
box (radius a quarter of its smallest side) and the line adds the contact
kernels' time and the share of tets whose starting centroid is in contact.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os

import numpy as np

from . import native as N
from . import workloads as W


def diag_lib():
    """The diagnostics library (pk_diag.h), built beside the product library."""
    from .build import DIAG_LIB
    if not os.path.exists(DIAG_LIB):
        raise RuntimeError(f"{DIAG_LIB} is missing: run python -m physik_amd.build")
    lib = C.CDLL(DIAG_LIB)
    lib.pk_diag_copy.restype = C.c_int
    lib.pk_diag_copy.argtypes = [C.c_longlong, C.c_int, C.POINTER(C.c_double)]
    lib.pk_diag_last_error.restype = C.c_char_p
    lib.pk_diag_last_error.argtypes = []
    return lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-mb", type=int, default=512)
    ap.add_argument("--workload", default="twist64", choices=["twist64", "beam10", "batch256", "diff32", "batch128"])
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "fp32u", "fp64u"])
    ap.add_argument("--shape", default=None,
                    help="CG launch shape lanes,spmv_rows,vector_rows,unroll (default: the system's own)")
    ap.add_argument("--spheres", type=int, default=0)
    ap.add_argument("--staging", type=int, default=1, choices=[0, 1], help="SpMV staging of d: 0 off, 1 on")
    args = ap.parse_args()
    diag = diag_lib()
    ms = C.c_double()
    nbytes = args.copy_mb * (1 << 20)
    if diag.pk_diag_copy(nbytes, 3, C.byref(ms)) != 0:
        raise RuntimeError("pk_diag_copy failed: " + diag.pk_diag_last_error().decode())
    w = getattr(W, args.workload)()
    s = W.make_system(w, {"fp64": N.PK_FP64, "fp32": N.PK_FP32_STORAGE, "fp32u": N.PK_FP32_UPPER,
                          "fp64u": N.PK_FP64_UPPER}[args.precision])
    if args.shape:
        s.tune(*[int(a) for a in args.shape.split(",")])
    s.set_staging(args.staging)
    extra = {}
    if args.spheres > 0:
        lo, hi = w.x0.min(0), w.x0.max(0)
        centers = lo + (np.arange(args.spheres)[:, None] + 0.5) / args.spheres * (hi - lo)
        radius = 0.25 * float((hi - lo).min())
        s.set_spheres(centers, np.full(args.spheres, radius), 1e4, 10.0)
        cent = w.x0[w.tets].mean(1)
        inside = (np.linalg.norm(cent[:, None, :] - centers[None], axis=2) < radius).any(1)
        extra = {"spheres": args.spheres, "sphere_radius": radius, "contact_tet_share": float(inside.mean())}
    s.profile(True)
    s.step(w.dt, 1, args.steps)
    if args.spheres > 0:
        extra.update({k: s.kernel_time(i) for k, i in (("contact", N.K_CONTACT), ("element", N.K_ELEMENT),
                                                       ("assemble", N.K_ASSEMBLE), ("rhs", N.K_RHS))})
    it = s.results()[0]
    blocks, ublocks, rows, body_launches = s.spmv_work()
    n_spmv, ms_spmv = s.kernel_time(N.K_SPMV)
    n_upd, ms_upd = s.kernel_time(N.K_UPDATE)
    n_dir, ms_dir = s.kernel_time(N.K_DIRECTION)
    n_as, ms_as = s.kernel_time(N.K_ASSEMBLE)
    n_el, ms_el = s.kernel_time(N.K_ELEMENT)
    n_rhs, ms_rhs = s.kernel_time(N.K_RHS)
    print(json.dumps({"copy_bytes": nbytes, "copy_ms": ms.value, "copy_GBps": 2 * nbytes / ms.value / 1e6,
                      "workload": args.workload, "steps": args.steps, "cg_iterations": int(it.max()),
                      "nnzb_full": s.nnzb_full, "dynamic_rows": s.nd, "tets": s.t, "nodes": s.n, "bodies": s.groups,
                      "spmv_launches": s.kernel_time(N.K_SPMV)[0], "spmv_blocks": blocks, "spmv_upper_blocks": ublocks, "spmv_rows": rows, "nnzb_upper": s.nnzb_upper,
                      "spmv_body_launches": body_launches, "shape": s.get_tune(),
                      "staging": s.get_staging(), "spmv_us": 1e3 * ms_spmv / max(n_spmv, 1), "update_us": 1e3 * ms_upd / max(n_upd, 1),
                      "direction_us": 1e3 * ms_dir / max(n_dir, 1), "assemble_us": 1e3 * ms_as / max(n_as, 1),
                      "element_us": 1e3 * ms_el / max(n_el, 1), "rhs_us": 1e3 * ms_rhs / max(n_rhs, 1), **extra}),
          flush=True)


if __name__ == "__main__":
    main()
