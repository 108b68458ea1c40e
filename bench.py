"""Benchmark of the implicit-Euler soft-body step on one MI355X.

    python bench.py [--steps K] [--warmup W]
                    [--workload batch128|diff32|twist64|beam10|batch256]
                    [--precision fp64|fp32] [--no-plain] [--no-cpu-baseline]
                    [--no-autotune | --shape L,R,V,U]

Default: the batched differentiable workload (128 bodies, 3.2 M tets, FP32
storage of the matrix's upper triangle with FP64 dots and residual), forward
+ adjoint; beside it the same run with the full matrix stored (identical
results, the full-matrix SpMV). FP64 workloads also store the upper triangle
by default (--precision fp64 for full storage). A step is one
implicit substep forward (element corotation, assembly, block-Jacobi PCG to
rtol 1e-6, integration) and its reverse (checkpoint-segment replay, adjoint
PCG to rtol 1e-7, corotational adjoint with dR/dF) over every body; K steps
are recorded with checkpoints every 8 and back-propagated from the per-body
losses, all inputs resident in GPU memory. Beside it, on the same line: the
plain single-body step (64^3 twisted Kuhn beam, FP64, 1.57 M tets), and the
CPU baselines (the oracle's FP64 forward + adjoint and the reference's own
FP32 forward, one host core, a bounded sample of the same workload).

`value` is tet-steps/s of the default workload; the SpMV roofline divides
the algorithmic bytes of the work the timed launches did (blocks and rows of
the bodies still iterating in each launch) by the kernel's summed in-kernel
wall-clock spans (the span a kernel trace reports).
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

DESCRIPTION = {
    "batch128": "128 bodies, n_x,n_y,n_z ~ U{8..24} (seed 4), Kuhn split, h 1/32 m, 3,203,808 tets, 639,897 nodes, "
                "per-tet log E ~ N(log 1e5, 0.2), nu 0.3, rho 1000, x=0 face fixed per body, per-body loss on the +x "
                "face vs a U[-0.05,0.05]^3 m offset target, dt 1/60 s, forward rtol 1e-6, adjoint rtol 1e-7, "
                "checkpoint every 8",
    "diff32": "32^3-cell Kuhn beam, h 1/64 m, 196,608 tets, 35,937 nodes, per-tet log E ~ N(log 1e5, 0.2) (seed 3), "
              "nu 0.3, "
              "rho 1000, x=0 face fixed, loss on the +x face vs 0.05 m down, dt 1/60 s, rtol 1e-6, checkpoint every "
              "10",
    "twist64": "64^3-cell Kuhn beam, 1,572,864 tets, 274,625 nodes, rest nodes jittered +-10% (seed 1), 90-degree "
               "initial twist, x=0 face fixed, E=1e6 Pa, nu=0.45, rho=1000, g=-9.81 z, dt=1/120 s, rtol 1e-6, max "
               "1024 CG iterations",
    "beam10": "10^3-cell 5-tet beam, 5,000 tets, 1,331 nodes, x=0 face fixed, E=1e5, nu=0.3, rho=1000, g=-9.81 z, "
              "dt=1/60, rtol 1e-6, max 200 CG iterations",
    "batch256": "256 bodies of U{4..32}^3 cells, 5-tet split (seed 2), E log-uniform [1e4,1e6], nu U[0.3,0.45], "
                "one random face fixed, v0 ~ N(0, 0.5 m/s), dt=1/60, rtol 1e-6, one CG per body",
}
DIFFERENTIABLE = ("batch128", "diff32")
# FP32 storage where the workload names it, FP64 elsewhere; both with the
# upper triangle stored (results identical to full storage, fewer bytes)
DEFAULT_PRECISION = {"batch128": "fp32u", "twist64": "fp64u", "batch256": "fp64u", "diff32": "fp64u",
                     "beam10": "fp64u"}
DEFAULT_STEPS = {"batch128": 40, "diff32": 50, "twist64": 50, "beam10": 100, "batch256": 100}


def workload(name):
    from physik_amd import workloads as W
    return {"twist64": W.twist64, "beam10": W.beam10, "batch256": W.batch256, "diff32": W.diff32,
            "batch128": W.batch128}[name]()


def spmv_bytes(blocks, rows, body_launches, s, upper_blocks=None, fused=False):
    """Algorithmic bytes of SpMV work q = A d.

    Full storage: every multiplied 3x3 block once (9 values of s bytes +
    int32 column), per row its int32 row pointer, one d entry read and one q
    entry written (3 values each), and the closing row pointer of each body in
    each launch. Upper-only storage (upper_blocks = the stored blocks among
    them): every stored record once (10 words of s bytes: 9 values and the
    column; it also serves the transposed product of its mirror row), an
    8-byte (slot, column) entry per lower block, per row an 8-byte pair of row
    pointers and d, q. With the direction update fused into the SpMV, a row
    reads z and d_{k-1} and writes d_k instead of reading d_k (two more vector
    passes). One launch over a whole system is
    spmv_bytes(nnzb_full, nd, bodies, s[, nnzb_upper])."""
    vec = (4 if fused else 2) * 3 * s
    if upper_blocks is None:
        return blocks * (9 * s + 4) + rows * (4 + vec) + 4 * body_launches
    return upper_blocks * 10 * s + (blocks - upper_blocks) * 8 + rows * (8 + vec) + 8 * body_launches


def measured_traffic(workload_name, precision):
    """HBM bytes per SpMV launch from the committed counter passes
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known size) and
    the algorithmic bytes per launch of that same profiled run, or None when
    no pass covers this workload and precision."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(f"{workload_name}/{precision}")
        return None if entry is None else (float(entry["bytes"]), float(entry["algorithmic"]), entry.get("shape"))
    except (OSError, ValueError, KeyError, TypeError):
        return None


def roofline(s, name, prec_name):
    """SpMV roofline over the timed region: algorithmic bytes of the work the
    counted launches did (a batched launch skips bodies whose CG loop has
    ended, so late launches move less) over their summed in-kernel spans."""
    from physik_amd import native as N
    n_spmv, ms_spmv = s.kernel_time(N.K_SPMV)
    esize = 8 if s.precision in (N.PK_FP64, N.PK_FP64_UPPER) else 4
    upper = s.precision in (N.PK_FP32_UPPER, N.PK_FP64_UPPER)
    blocks, ublocks, rows, body_launches = s.spmv_work()
    _, staged_rows, segmented, fused = s.get_staging()
    total = spmv_bytes(blocks, rows, body_launches, esize, ublocks if upper else None, fused)
    full = spmv_bytes(s.nnzb_full, s.nd, s.groups, esize, s.nnzb_upper if upper else None, fused)
    # the same work in full storage, for comparison across formats
    full_format = spmv_bytes(blocks, rows, body_launches, esize, None, fused)
    achieved = total / (ms_spmv * 1e-3) / 1e9 if n_spmv and ms_spmv > 0 else 0.0
    measured = measured_traffic(name, prec_name)
    return {
        "bound": "hbm",
        "kernel": "k_spmv (q = A d_k, d.q partial per tile; d_k = z_k + beta d_k-1 formed in it when fused)",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None if measured is None else measured[0],
        "traffic_vs_algorithmic": None if measured is None else measured[0] / measured[1],
        "traffic_shape": None if measured is None else measured[2],  # CG launch shape of the counter passes
        # the counters' bytes per launch over this run's average launch time:
        # the rate the kernel actually moves data at (a plain copy reaches ~5.3 TB/s)
        "traffic_rate_GBps": (None if measured is None or not n_spmv or ms_spmv <= 0
                              else measured[0] / (ms_spmv / n_spmv * 1e-3) / 1e9),
        "storage": "upper triangle" if upper else "full",
        "staging": {"rows": staged_rows, "segments": segmented, "direction_fused": fused},
        "bytes_per_launch": total / max(n_spmv, 1),
        "full_storage_bytes_per_launch": full_format / max(n_spmv, 1),
        "bytes_per_full_launch": full,
        "active_fraction": total / (full * max(n_spmv, 1)),
        "avg_launch_us": ms_spmv / max(n_spmv, 1) * 1e3,
        "launches": n_spmv,
    }


def kernels_ms(s):
    from physik_amd import native as N
    return {N.KERNEL_NAMES[k]: round(s.kernel_time(k)[1], 3) for k in range(len(N.KERNEL_NAMES))
            if s.kernel_time(k)[0] > 0 or k != N.K_CONTACT}


def prepare(name, prec_name, args):
    from physik_amd import native as N
    from physik_amd import workloads as W
    w = workload(name)
    prec = {"fp64": N.PK_FP64, "fp32": N.PK_FP32_STORAGE, "fp32u": N.PK_FP32_UPPER, "fp64u": N.PK_FP64_UPPER}[prec_name]
    s = W.make_system(w, prec)
    if args.shape:
        s.tune(*[int(a) for a in args.shape.split(",")])
        shape = s.get_tune()
    else:
        s.step(w.dt)
        s.set_state(w.x0, w.v0)
        shape = s.get_tune() if args.no_autotune else s.autotune(w.dt)
    s.set_state(w.x0, w.v0)
    s.set_graphs(args.cg_launch == "graphs")
    return w, s, dict(zip(("lanes_per_row", "spmv_rows", "vector_rows", "unroll"), shape))


def frame_export(s, reps=5):
    """The per-frame readback after a step, measured three ways (median of
    `reps`): the FP64 state downloaded into new host arrays
    (pk_system_get_state, 48 B per node), the float frame export copied out
    (pk_system_export_begin + export_read, 24 B per node through pinned
    memory), and the export alone up to the moment the host may read it."""
    def med(f):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return 1e3 * sorted(ts)[len(ts) // 2]

    s.state()
    s.export_begin()
    s.export_read()
    fp64 = med(s.state)
    exp = med(lambda: (s.export_begin(), s.export_read()))
    wait = med(lambda: (s.export_begin(), s.export_read(0, 0)))
    return {"nodes": s.n, "fp64_download_ms": fp64, "float_export_ms": exp, "float_export_wait_ms": wait,
            "float_export_GBps": 24 * s.n / (wait * 1e6)}


def run_plain(name, prec_name, steps, warmup, args):
    """K plain implicit steps from the workload's initial state."""
    w, s, shape = prepare(name, prec_name, args)
    for _ in range(warmup):
        s.step(w.dt)
    s.set_state(w.x0, w.v0)
    s.profile(True)
    t0 = time.perf_counter()
    s.step(w.dt, 1, steps)  # returns after the stream has drained
    el = time.perf_counter() - t0
    its = s.results()[0]
    out = {
        "workload": f"{name}: {DESCRIPTION[name]}",
        "value": w.n_tets * steps / el,
        "unit": "tet·steps/s",
        "steps": steps,
        "ms_per_step": el / steps * 1e3,
        "precision": prec_name,
        "tets": w.n_tets, "nodes": w.n_nodes, "bodies": w.n_bodies, "dynamic_rows": s.nd,
        "nnzb_full": s.nnzb_full, "nnzb_upper": s.nnzb_upper,
        "cg_iterations_last": int(its.max()),
        "cg_launch_shape": shape,
        "cg_launch": args.cg_launch,
        "roofline": roofline(s, name, prec_name),
        "kernels_ms": kernels_ms(s),
    }
    s.profile(False)
    out["frame_export"] = frame_export(s)
    s.close()
    return out, w


def run_fwd_adj(name, prec_name, steps, warmup, args):
    """Record K substeps with checkpoints, then back-propagate the loss."""
    from physik_amd import workloads as W
    w, s, shape = prepare(name, prec_name, args)
    every = w.checkpoint or steps
    for _ in range(warmup):
        s.set_state(w.x0, w.v0)
        s.record(w.dt, steps, every)
        x, _ = s.state()
        s.backward(W.loss_and_seed(w, x)[1], None, tol=w.adj_tol, max_iter=w.max_iter)
    s.set_state(w.x0, w.v0)
    s.profile(True)
    t0 = time.perf_counter()
    s.record(w.dt, steps, every)
    t1 = time.perf_counter()
    x, _ = s.state()
    loss, xbar = W.loss_and_seed(w, x)
    t2 = time.perf_counter()
    gE, _, gv, adj_its = s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter)
    t3 = time.perf_counter()
    el = (t1 - t0) + (t3 - t2)
    out = {
        "workload": f"{name}: {DESCRIPTION[name]}",
        "value": w.n_tets * steps / el,
        "unit": "tet·steps/s",
        "steps": steps,
        "ms_per_step": el / steps * 1e3,
        "forward_ms": (t1 - t0) * 1e3,
        "backward_ms": (t3 - t2) * 1e3,
        "precision": prec_name,
        "tets": w.n_tets, "nodes": w.n_nodes, "bodies": w.n_bodies, "dynamic_rows": s.nd,
        "nnzb_full": s.nnzb_full, "nnzb_upper": s.nnzb_upper, "checkpoint_every": every,
        "loss": loss, "grad_log_young_norm": float(np.linalg.norm(gE)), "grad_v0_norm": float(np.linalg.norm(gv)),
        "cg_iterations_last_forward": int(s.results()[0].max()),
        "adjoint_cg_iterations": {"mean": float(np.mean(adj_its)), "max": int(np.max(adj_its))},
        "cg_launch_shape": shape,
        "cg_launch": args.cg_launch,
        "roofline": roofline(s, name, prec_name),
        "kernels_ms": kernels_ms(s),
    }
    s.profile(False)
    s.close()
    return out, w


def _oracle(checkers, wb, precision=64):
    o = checkers.Oracle(precision, wb.rest, wb.tets, wb.young, wb.poisson, wb.damping, wb.density)
    o.set_params(wb.gravity, 1, wb.tol, wb.max_iter)
    o.set_fixed(wb.fixed)
    o.set_state(wb.x0, wb.v0)
    return o


def cpu_fwd_adj(w, name, steps, budget_s):
    """The oracle's FP64 forward + adjoint on one host core, body by body of
    the same workload, as many bodies as fit the budget (at least one).  A
    body too large for the budget is sampled over its first substeps: the
    untimed forward that produces the loss seed stops once it has used a
    third of the budget (the recorded forward plus the reverse pass cost
    about three forwards)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    from physik_amd import workloads as W
    tet_steps, bodies, el, sub = 0, 0, 0.0, steps
    for b in range(w.n_bodies):
        wb = w.body(b) if w.n_bodies > 1 else w
        # the loss seed needs the final state: an untimed forward run first
        p = _oracle(checkers, wb)
        t0 = time.perf_counter()
        sub = 0
        while sub < steps:
            p.step(wb.dt)
            sub += 1
            if bodies == 0 and time.perf_counter() - t0 >= budget_s / 3:
                break
        xbar = W.loss_and_seed(wb, p.state()[0])[1]
        p.close()
        o = _oracle(checkers, wb)
        t0 = time.perf_counter()
        o.adjoint(wb.dt, sub, xbar, None, tol=wb.adj_tol, max_iter=wb.max_iter)
        el += time.perf_counter() - t0
        o.close()
        bodies += 1
        tet_steps += wb.n_tets * sub
        if el >= budget_s or sub < steps:
            break
    what = f"{steps} substeps" if sub == steps else f"the first {sub} of {steps} substeps"
    return {
        "value": tet_steps / max(el, 1e-9),
        "unit": "tet·steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{bodies} of {w.n_bodies} bodies of {name} ({tet_steps // sub:,} tets), {what} forward + adjoint "
                  f"each (record, then reverse), FP64 oracle restatement (the reference has no adjoint), one host "
                  f"core, {el:.1f} s timed",
    }


def cpu_reference_forward(w, name, steps, budget_s):
    """The reference's own FP32 forward step (its sources compiled in place,
    else the oracle's FP32 restatement) on one host core, body by body."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    kind = "reference" if checkers.reference_available() else "port"
    tet_steps, el, bodies, last_it = 0, 0.0, 0, 0
    for b in range(w.n_bodies):
        wb = w.body(b) if w.n_bodies > 1 else w
        if kind == "reference":
            sim = checkers.Reference(wb.rest, wb.tets, float(wb.density[0]), float(wb.young[0]),
                                     float(wb.poisson[0]), float(wb.damping[0]))
            if not (wb.young == wb.young[0]).all():
                sim.set_young(wb.young)
            sim.set_params(wb.gravity, 1, wb.tol, wb.max_iter)
            sim.set_fixed(wb.fixed)
            sim.set_state(wb.x0, wb.v0)
        else:
            sim = _oracle(checkers, wb, 32)
        bodies += 1
        for _ in range(steps):
            t0 = time.perf_counter()
            sim.step(wb.dt)
            el += time.perf_counter() - t0
            tet_steps += wb.n_tets
            if el >= budget_s:
                break
        last_it = sim.cg()[0]
        sim.close()
        if el >= budget_s:
            break
    return {
        "value": tet_steps / max(el, 1e-9),
        "unit": "tet·steps/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{tet_steps:,} tet-steps of {name} ({bodies} bod{'y' if bodies == 1 else 'ies'}, forward only, "
                  f"FP32, from the initial state) on one host core, {el:.1f} s timed, last CG {last_it} iterations; "
                  "setup excluded",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch128", choices=list(DESCRIPTION))
    ap.add_argument("--precision", default=None, choices=["fp64", "fp32", "fp32u", "fp64u"],
                    help="fp32: FP32 storage of the full matrix; fp32u, fp64u: of its upper triangle only")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain single-body step beside the headline")
    ap.add_argument("--no-full-storage", action="store_true",
                    help="skip the full-matrix FP32 run beside an upper-triangle (fp32u) headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default CG launch shape")
    ap.add_argument("--shape", default=None, help="CG launch shape lanes,spmv_rows,vector_rows,unroll (no autotune)")
    ap.add_argument("--cg-launch", default="kernels", choices=["kernels", "graphs"],
                    help="issue each CG launch from the host, or replay captured 4-iteration HIP graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the batched and single-body "
              f"paths shard across GPUs by body or by slab, which this build does not launch)", file=sys.stderr)
        sys.exit(2)

    name = args.workload
    prec_name = args.precision or DEFAULT_PRECISION.get(name, "fp64")
    steps = args.steps or DEFAULT_STEPS[name]
    diff = name in DIFFERENTIABLE
    head, w = (run_fwd_adj if diff else run_plain)(name, prec_name, steps, args.warmup, args)
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "tet·steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp64u": "f64 (upper triangle stored)", "fp32": "f32 storage, f64 dots",
                  "fp32u": "f32 storage (upper triangle), f64 dots"}[prec_name],
        "data": "synthetic (seeded meshes, no checkpoint files)",
        "config": {"workload": head["workload"], "mode": "forward+adjoint" if diff else "forward",
                   "parallelism": "single GPU",
                   **{k: v for k, v in head.items() if k not in ("workload", "value", "unit", "roofline",
                                                                 "kernels_ms", "ms_per_step", "steps")}},
        "roofline": head["roofline"],
        "kernels_ms": head["kernels_ms"],
    }
    if diff and prec_name == "fp32u" and not args.no_full_storage:
        # the same run with the full matrix stored (identical results): its
        # SpMV moves the whole matrix, the headline's reads the upper half
        full, _ = run_fwd_adj(name, "fp32", steps, args.warmup, args)
        line["full_storage_step"] = {k: full[k] for k in ("value", "ms_per_step", "precision", "cg_launch_shape",
                                                          "roofline", "kernels_ms")}
    if not args.no_plain and name != "twist64":
        plain, _ = run_plain("twist64", DEFAULT_PRECISION["twist64"], DEFAULT_STEPS["twist64"], args.warmup, args)
        line["plain_step"] = plain
        line["vs_plain_step"] = head["value"] / plain["value"]
    if not args.no_cpu_baseline:
        if diff:
            line["cpu_baseline"] = cpu_fwd_adj(w, name, steps, args.cpu_budget)
            line["cpu_reference_forward"] = cpu_reference_forward(w, name, steps, args.cpu_budget / 2)
        else:
            line["cpu_baseline"] = cpu_reference_forward(w, name, steps, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
