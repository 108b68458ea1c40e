"""Benchmark of the implicit-Euler soft-body step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload twist64|beam10|batch256]
                    [--precision fp64|fp32] [--no-cpu-baseline] [--no-autotune | --shape L,R,V,U]

A step is one implicit-Euler frame (element corotation, assembly, block-Jacobi
PCG to rtol 1e-6, integration) over the whole workload, with every input
resident in GPU memory. Prints one JSON line: tet-steps/s over exactly K timed
steps after W untimed ones, the roofline of the dominant kernel (the CG SpMV)
from the kernel's own wall-clock span over the timed region, and a CPU
baseline: the reference's own FP32 step (its sources compiled in place) timed
on one host core over a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "tet·steps/s fwd+adjoint batched vs plain step, 1 GPU, rtol 1e-6; SpMV % HBM"


def workload(name):
    from physik_amd import workloads as W
    if name == "twist64":
        return W.twist64()
    if name == "beam10":
        return W.beam10()
    if name == "batch256":
        return W.batch256()
    raise SystemExit(f"unknown workload {name}")


DESCRIPTION = {
    "twist64": "64^3-cell Kuhn beam, 1,572,864 tets, 274,625 nodes, rest nodes jittered +-10% (seed 1), "
               "90-degree initial twist, x=0 face fixed, E=1e6 Pa, nu=0.45, rho=1000, g=-9.81 z, dt=1/120 s, "
               "rtol 1e-6, max 1024 CG iterations",
    "beam10": "10^3-cell 5-tet beam, 5,000 tets, 1,331 nodes, x=0 face fixed, E=1e5, nu=0.3, rho=1000, "
              "g=-9.81 z, dt=1/60, rtol 1e-6, max 200 CG iterations",
    "batch256": "256 bodies of U{4..32}^3 cells, 5-tet split (seed 2), E log-uniform [1e4,1e6], nu U[0.3,0.45], "
                "one random face fixed, v0 ~ N(0, 0.5 m/s), dt=1/60, rtol 1e-6, one CG per body",
}


def spmv_bytes(nnzf, nd, s):
    """Algorithmic bytes of one fused SpMV launch: the full BSR once (9 values
    + int32 column per block, row pointers), the direction update's reads of z
    and d_{k-1} and its write of d_k, and the write of q = A d_k."""
    return nnzf * (9 * s + 4) + (nd + 1) * 4 + 4 * nd * 3 * s


def spmv_bytes_plain(nnzf, nd, s):
    """The same for a plain SpMV y = A x (x read once, y written once)."""
    return nnzf * (9 * s + 4) + (nd + 1) * 4 + 2 * nd * 3 * s


def measured_traffic(workload_name, precision):
    """HBM bytes per SpMV launch from the committed counter passes
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known size), or
    None when no pass covers this workload and precision."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(f"{workload_name}/{precision}")
        return None if entry is None else float(entry["bytes"])
    except (OSError, ValueError, KeyError, TypeError):
        return None


def cpu_baseline(w, name, budget_s=30.0):
    """The reference step on one core: oracle/_ref (reference sources) when
    built, else the oracle's FP32 restatement. Times whole PHYSIK_Step calls
    after setup, as many as fit the budget (at least one)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    ncores = 1
    if checkers.reference_available():
        kind = "reference"
        sim = checkers.Reference(w.rest, w.tets, float(w.density[0]), float(w.young[0]), float(w.poisson[0]),
                                 float(w.damping[0]))
        if not (w.young == w.young[0]).all():
            sim.set_young(w.young)
        sim.set_params(w.gravity, 1, w.tol, w.max_iter)
    else:
        kind = "port"
        sim = checkers.Oracle(32, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
        sim.set_params(w.gravity, 1, w.tol, w.max_iter)
    sim.set_fixed(w.fixed)
    sim.set_state(w.x0, w.v0)
    steps, t0 = 0, time.perf_counter()
    while True:
        sim.step(w.dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= w.steps or el * (steps + 1) / steps > budget_s * 1.5:
            break
    it = sim.cg()[0]
    sim.close()
    return {
        "value": w.n_tets * steps / el,
        "unit": "tet·steps/s",
        "cores": ncores,
        "kind": kind,
        "sample": f"{steps} step(s) of {name} in FP32 from its initial state on one host core "
                  f"({el:.1f} s, last CG {it} iterations; setup excluded)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="twist64", choices=["twist64", "beam10", "batch256"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default CG launch shape")
    ap.add_argument("--shape", default=None, help="CG launch shape lanes,spmv_rows,vector_rows,unroll (no autotune)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the batched and single-body "
              f"paths shard across GPUs by body or by slab, which this build does not launch)", file=sys.stderr)
        sys.exit(2)

    from physik_amd import native as N
    from physik_amd import workloads as W

    w = workload(args.workload)
    prec = N.PK_FP64 if args.precision == "fp64" else N.PK_FP32_STORAGE
    s = W.make_system(w, prec)
    for _ in range(args.warmup):
        s.step(w.dt)
    s.set_state(w.x0, w.v0)
    if args.shape:
        s.tune(*[int(a) for a in args.shape.split(",")])
        shape = s.get_tune()
    else:
        shape = s.get_tune() if args.no_autotune else s.autotune(w.dt)
    s.set_state(w.x0, w.v0)  # the timed run starts from the workload's initial state
    s.profile(True)
    t0 = time.perf_counter()
    s.step(w.dt, 1, args.steps)  # returns after the stream has drained
    el = time.perf_counter() - t0
    its = s.results()[0]
    n_spmv, ms_spmv = s.kernel_time(N.K_SPMV)
    kernels = {N.KERNEL_NAMES[k]: dict(zip(("launches", "ms"), s.kernel_time(k))) for k in range(7)}
    s.profile(False)

    esize = 8 if prec == N.PK_FP64 else 4
    per_launch = spmv_bytes(s.nnzb_full, s.nd, esize)
    avg_s = ms_spmv / max(n_spmv, 1) * 1e-3
    achieved = per_launch / avg_s / 1e9 if n_spmv else 0.0
    traffic = measured_traffic(args.workload, args.precision)
    value = w.n_tets * args.steps / el
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "tet·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == N.PK_FP64 else "f32 storage, f64 dots",
        "data": "synthetic (seeded meshes, no checkpoint)",
        "config": {
            "workload": f"{args.workload}: {DESCRIPTION[args.workload]}",
            "tets": w.n_tets, "nodes": w.n_nodes, "bodies": w.n_bodies, "dynamic_rows": s.nd,
            "nnzb_full": s.nnzb_full, "nnzb_upper": s.nnzb_upper,
            "cg_iterations_last": int(its.max()), "precision": args.precision,
            "cg_launch_shape": dict(zip(("lanes_per_row", "spmv_rows", "vector_rows", "unroll"), shape)),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv (q = A d_k with d_k = z_k + beta d_{k-1} fused, d.q partial)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": per_launch,
            "plain_spmv_bytes_per_launch": spmv_bytes_plain(s.nnzb_full, s.nd, esize),
            "avg_launch_us": avg_s * 1e6,
            "launches": n_spmv,
        },
        "kernels_ms": {k: round(v["ms"], 3) for k, v in kernels.items()},
    }
    s.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, args.workload, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
