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

from physik_amd.roofline import assembly_bytes as _assembly_bytes  # noqa: E402
from physik_amd.roofline import precision_sizes, spmv_bytes, spmv_format_bytes  # noqa: E402

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


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    from physik_amd import workloads as W
    return {"twist64": W.twist64, "beam10": W.beam10, "batch256": W.batch256, "diff32": W.diff32,
            "batch128": W.batch128}[name]()


def measured_traffic(workload_name, precision):
    """HBM bytes per launch from the committed counter passes
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known size): the
    SpMV's bytes, the algorithmic bytes of that same profiled run and its
    launch shape, and the assembly kernels' bytes per substep, or None when
    no pass covers this workload and precision."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(f"{workload_name}/{precision}")
    except (OSError, ValueError):
        return None
    if entry is None:
        return None
    try:
        return {"bytes": float(entry["bytes"]), "algorithmic": float(entry["algorithmic"]),
                "shape": entry.get("shape"), "assembly": entry.get("assembly"), "source": entry.get("source")}
    except (KeyError, TypeError, ValueError):
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
    _, staged_rows, segmented = s.get_staging()
    ub = ublocks if upper else None
    total = spmv_bytes(blocks, rows, body_launches, esize, ub)
    fmt = spmv_format_bytes(blocks, rows, body_launches, esize, ub)
    full = spmv_bytes(s.nnzb_full, s.nd, s.groups, esize, s.nnzb_upper if upper else None)
    # the same work in full storage, for comparison across formats
    full_format = spmv_bytes(blocks, rows, body_launches, esize, None)
    sec = ms_spmv * 1e-3
    achieved = total / sec / 1e9 if n_spmv and ms_spmv > 0 else 0.0
    measured = measured_traffic(name, prec_name)
    per = max(n_spmv, 1)
    return {
        "bound": "hbm",
        "kernel": ("k_spmv_upper (sliced upper-only storage, one lane per row)" if upper else "k_spmv")
                  + " (q = A d_k and the d.q partial per tile)",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None if measured is None else measured["bytes"],
        "traffic_vs_algorithmic": None if measured is None else measured["bytes"] / measured["algorithmic"],
        "traffic_shape": None if measured is None else measured["shape"],  # CG launch shape of the counter passes
        # the counters' bytes per launch over this run's average launch time:
        # the rate the kernel actually moves data at (a plain copy reaches ~5.3 TB/s)
        "traffic_rate_GBps": (None if measured is None or not n_spmv or ms_spmv <= 0
                              else measured["bytes"] / (ms_spmv / per * 1e-3) / 1e9),
        "storage": "upper triangle, sliced" if upper else "full",
        "staging": {"rows": staged_rows, "segments": segmented},
        "bytes_per_launch": total / per,
        # what the storage format adds to the algorithmic bytes (upper
        # storage: the mirror list and FP64 column words, less the row pointers)
        "overhead_bytes_per_launch": (fmt - total) / per,
        "format_rate_GBps": fmt / sec / 1e9 if n_spmv and ms_spmv > 0 else 0.0,
        "full_storage_bytes_per_launch": full_format / per,
        "bytes_per_full_launch": full,
        "active_fraction": total / (full * per),
        "avg_launch_us": ms_spmv / per * 1e3,
        "launches": n_spmv,
    }


def assembly_bytes(s, w):
    """Algorithmic bytes of one assembly of system s (physik_amd.roofline)."""
    sA, upper = precision_sizes(s.precision)
    return _assembly_bytes(w.n_tets, w.n_nodes, s.nnzb_upper if upper else s.nnzb_full, s.nd, sA)


def assembly_roofline(s, w, name, prec_name):
    """Assembly roofline over the timed region: assembly_bytes per assembly
    over the summed times of the element, slot-gather and right-hand-side
    launches (HIP events on the system's stream)."""
    from physik_amd import native as N
    n_el, ms_el = s.kernel_time(N.K_ELEMENT)
    _, ms_as = s.kernel_time(N.K_ASSEMBLE)
    _, ms_rhs = s.kernel_time(N.K_RHS)
    ms = ms_el + ms_as + ms_rhs
    b = assembly_bytes(s, w)
    achieved = b * n_el / (ms * 1e-3) / 1e9 if n_el and ms > 0 else 0.0
    measured = measured_traffic(name, prec_name)
    asm = None if measured is None else measured.get("assembly")
    traffic = None if not asm else sum(float(asm[k]) for k in ("element", "assemble", "rhs"))
    return {
        "bound": "hbm",
        "kernel": "k_element + k_assemble + k_rhs (one assembly: corotation, slot gather, right-hand side)",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "bytes_per_assembly": b,
        "assemblies": n_el,
        "ms_per_assembly": ms / max(n_el, 1),
        "ms_per_assembly_by_kernel": {"element": ms_el / max(n_el, 1), "assemble": ms_as / max(n_el, 1),
                                      "rhs": ms_rhs / max(n_el, 1)},
        "traffic": traffic,
        "traffic_by_kernel": asm,
        "traffic_vs_algorithmic": None if traffic is None else traffic / b,
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
    s.set_graphs({"kernels": 0, "graphs": 1, "steps": 2}[args.cg_launch])
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


def run_plain(name, prec_name, steps, warmup, args, export=True, profile=True):
    """K plain implicit steps from the workload's initial state (profiled
    unless profile is False or the substeps are captured whole)."""
    w, s, shape = prepare(name, prec_name, args)
    for _ in range(warmup):
        s.step(w.dt)
    s.set_state(w.x0, w.v0)
    # a profiled step launches its kernels: substeps captured whole run unprofiled
    profiled = profile and args.cg_launch != "steps"
    s.profile(profiled)
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
        "roofline": roofline(s, name, prec_name) if profiled else None,
        "assembly_roofline": assembly_roofline(s, w, name, prec_name) if profiled else None,
        "kernels_ms": kernels_ms(s) if profiled else None,
    }
    s.profile(False)
    if export:
        out["frame_export"] = frame_export(s)
    out["final_state"] = s.state()
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
        "assembly_roofline": assembly_roofline(s, w, name, prec_name),
        "kernels_ms": kernels_ms(s),
    }
    s.profile(False)
    out["final_state"] = (x, s.state()[1])
    out["gradients"] = (gE, gv)
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


def cpu_info():
    """Host CPU model, logical CPUs this process may use, and nproc."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "nproc": usable, "os_cpu_count": os.cpu_count()}


def oracle_error(w, head, budget_s):
    """FP64-oracle error of the GPU run per body, on the smallest bodies
    that fit the budget (at least one): relative L2 of the final x and v
    after the run's K substeps, and, for a differentiable workload, of
    dL/dlog E and dL/dv0 against the oracle adjoint seeded from the oracle's
    own final state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    from physik_amd import workloads as W

    def rel(a, b, floor=0.0):
        nb = max(np.linalg.norm(b), floor)
        return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))

    x, v = head["final_state"]
    grads = head.get("gradients")
    steps = head["steps"]
    order = sorted(range(w.n_bodies), key=lambda b: w.body_tets[b][1] - w.body_tets[b][0]) if w.n_bodies > 1 \
        else [0]
    rows, t0 = [], time.perf_counter()
    for b in order:
        wb = w.body(b) if w.n_bodies > 1 else w
        n0, n1 = w.body_nodes[b]
        t_0, t_1 = w.body_tets[b]
        o = _oracle(checkers, wb)
        for _ in range(steps):
            o.step(wb.dt)
        xo, vo = o.state()
        o.close()
        # v relative to the larger of the final and initial velocity norms: a
        # body that has come to rest keeps v near 0, where a plain relative
        # error measures the CG tolerance only
        row = {"body": b, "tets": t_1 - t_0, "x": rel(x[n0:n1], xo),
               "v": rel(v[n0:n1], vo, np.linalg.norm(wb.v0))}
        if grads is not None:
            a = _oracle(checkers, wb)
            oE, _, ov, _ = a.adjoint(wb.dt, steps, W.loss_and_seed(wb, xo)[1], None, tol=wb.adj_tol,
                                     max_iter=wb.max_iter)
            a.close()
            row["grad_log_young"] = rel(grads[0][t_0:t_1], oE)
            row["grad_v0"] = rel(grads[1][n0:n1], ov)
        rows.append(row)
        if time.perf_counter() - t0 >= budget_s:
            break
    keys = [k for k in ("x", "v", "grad_log_young", "grad_v0") if k in rows[0]]
    return {
        "reference": "FP64 oracle restatement of the reference step (and its adjoint), same inputs, "
                     f"{steps} substeps, forward rtol {w.tol:g}" + (f", adjoint rtol {w.adj_tol:g}" if grads else ""),
        "sample": f"the {len(rows)} smallest of {w.n_bodies} bodies" if w.n_bodies > 1 else "the whole body",
        "metric": "relative L2 per body (v: over max(|v_S|, |v_0|) of the oracle)",
        **{k: {"max": max(r[k] for r in rows), "median": float(np.median([r[k] for r in rows]))} for k in keys},
        "per_body": rows,
        "seconds": time.perf_counter() - t0,
    }


def _ref_worker(job):
    """One worker of the process-parallel CPU figure: the reference's FP32
    forward over its share of the bodies, `steps` substeps each, timing only
    the steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    el, tet_steps = 0.0, 0
    for body in job["bodies"]:
        rest, tets, young, poisson, damping, density, fixed, x0, v0, dt, tol, max_iter, gravity = body
        if checkers.reference_available():
            sim = checkers.Reference(rest, tets, float(density[0]), float(young[0]), float(poisson[0]),
                                     float(damping[0]))
            if not (young == young[0]).all():
                sim.set_young(young)
        else:
            sim = checkers.Oracle(32, rest, tets, young, poisson, damping, density)
        sim.set_params(gravity, 1, tol, max_iter)
        sim.set_fixed(fixed)
        sim.set_state(x0, v0)
        t0 = time.perf_counter()
        for _ in range(job["steps"]):
            sim.step(dt)
        el += time.perf_counter() - t0
        tet_steps += len(tets) * job["steps"]
        sim.close()
    return el, tet_steps


def cpu_parallel_upper_bound(w, name, steps, workers):
    """Every body of the batch as its own World on `workers` host processes
    (bodies spread by longest-processing-time on tet count): total
    tet-steps over the slowest worker's step time. An upper bound for a
    host engine that runs one reference World per body per core, not the
    reference itself (which has no threading)."""
    import multiprocessing as mp
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    loads = [0] * workers
    jobs = [{"bodies": [], "steps": steps} for _ in range(workers)]
    for b in sorted(range(w.n_bodies), key=lambda b: -(w.body_tets[b][1] - w.body_tets[b][0])):
        wb = w.body(b)
        k = loads.index(min(loads))
        loads[k] += wb.n_tets
        jobs[k]["bodies"].append((wb.rest, wb.tets, wb.young, wb.poisson, wb.damping, wb.density, wb.fixed, wb.x0,
                                  wb.v0, wb.dt, wb.tol, wb.max_iter, wb.gravity))
    ctx = mp.get_context("spawn")  # no fork of a process that holds the GPU
    with ctx.Pool(workers) as pool:
        res = pool.map(_ref_worker, jobs)
    slowest = max(el for el, _ in res)
    total = sum(ts for _, ts in res)
    return {
        "value": total / max(slowest, 1e-9),
        "unit": "tet·steps/s",
        "cores": workers,
        "kind": "reference" if checkers.reference_available() else "port",
        "label": "upper bound, not the reference: one World per body, bodies spread over host processes",
        "sample": f"all {w.n_bodies} bodies of {name}, {steps} substeps each from the initial state, FP32 forward, "
                  f"{workers} processes, slowest {slowest:.1f} s of step time; setup excluded",
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
    ap.add_argument("--cg-launch", default="kernels", choices=["kernels", "graphs", "steps"],
                    help="issue each CG launch from the host, replay captured 4-iteration HIP graphs, or replay "
                         "each implicit substep as one captured HIP graph (forward steps; unprofiled)")
    ap.add_argument("--no-oracle-error", action="store_true", help="skip the per-body FP64-oracle error")
    ap.add_argument("--oracle-budget", type=float, default=10.0, help="seconds of oracle work for the error sample")
    ap.add_argument("--no-batch-forward", action="store_true", help="skip the batch256 forward sub-record")
    ap.add_argument("--no-cpu-parallel", action="store_true", help="skip the process-parallel CPU upper bound")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the batched and single-body "
              f"paths shard across GPUs by body or by slab, which this build does not launch)", file=sys.stderr)
        sys.exit(2)

    name = args.workload
    prec_name = args.precision or DEFAULT_PRECISION.get(name, "fp64")
    steps = args.steps or DEFAULT_STEPS[name]
    diff = name in DIFFERENTIABLE
    progress(f"headline: {name} {prec_name}, {steps} steps")
    head, w = (run_fwd_adj if diff else run_plain)(name, prec_name, steps, args.warmup, args)
    internal = ("final_state", "gradients")
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
                                                                 "assembly_roofline", "kernels_ms", "ms_per_step",
                                                                 "steps", *internal)}},
        "roofline": head["roofline"],
        "assembly_roofline": head["assembly_roofline"],
        "kernels_ms": head["kernels_ms"],
    }
    if not args.no_oracle_error:
        progress("oracle error of the headline run")
        line["oracle_error"] = oracle_error(w, head, args.oracle_budget)
    if diff and prec_name == "fp32u" and not args.no_full_storage:
        # the same run with the full matrix stored (identical results): its
        # SpMV moves the whole matrix, the headline's reads the upper half
        progress("full-storage run")
        full, _ = run_fwd_adj(name, "fp32", steps, args.warmup, args)
        line["full_storage_step"] = {k: full[k] for k in ("value", "ms_per_step", "precision", "cg_launch_shape",
                                                          "roofline", "assembly_roofline", "kernels_ms")}
    if not args.no_plain and name != "twist64":
        progress("plain step: twist64")
        plain, _ = run_plain("twist64", DEFAULT_PRECISION["twist64"], DEFAULT_STEPS["twist64"], args.warmup, args)
        for k in internal:
            plain.pop(k, None)
        line["plain_step"] = plain
        line["vs_plain_step"] = head["value"] / plain["value"]
    if not args.no_batch_forward and name != "batch256":
        # the batched forward workload (256 bodies, FP64), launched kernel by
        # kernel and with every substep replayed as one captured HIP graph
        progress("batched forward: batch256, kernels")
        bf, wbf = run_plain("batch256", DEFAULT_PRECISION["batch256"], DEFAULT_STEPS["batch256"], args.warmup, args,
                            export=False)
        rec = {k: v for k, v in bf.items() if k not in internal}
        if not args.no_oracle_error:
            rec["oracle_error"] = oracle_error(wbf, bf, args.oracle_budget / 2)
        gargs = argparse.Namespace(**{**vars(args), "cg_launch": "steps"})
        progress("batched forward: batch256, substeps captured as HIP graphs")
        bg, _ = run_plain("batch256", DEFAULT_PRECISION["batch256"], DEFAULT_STEPS["batch256"], args.warmup, gargs,
                          export=False)
        rec["graphs"] = {k: bg[k] for k in ("value", "ms_per_step", "cg_launch", "cg_iterations_last")}
        # the same steps launched kernel by kernel without profiling, beside it
        kargs = argparse.Namespace(**{**vars(args), "cg_launch": "kernels"})
        progress("batched forward: batch256, kernels unprofiled")
        ku, _ = run_plain("batch256", DEFAULT_PRECISION["batch256"], DEFAULT_STEPS["batch256"], args.warmup, kargs,
                          export=False, profile=False)
        rec["graphs"]["kernels_unprofiled_ms_per_step"] = ku["ms_per_step"]
        line["batch_forward"] = rec
    if not args.no_cpu_baseline:
        progress("CPU baselines")
        info = cpu_info()
        if diff:
            line["cpu_baseline"] = cpu_fwd_adj(w, name, steps, args.cpu_budget)
            line["cpu_reference_forward"] = cpu_reference_forward(w, name, steps, args.cpu_budget / 2)
        else:
            line["cpu_baseline"] = cpu_reference_forward(w, name, steps, args.cpu_budget)
        line["cpu_baseline"].update({"cpu_model": info["model"], "nproc": info["nproc"]})
        if w.n_bodies > 1 and not args.no_cpu_parallel:
            workers = max(1, min(16, info["nproc"] or 1))
            progress(f"CPU process-parallel figure: {workers} processes")
            line["cpu_parallel_upper_bound"] = cpu_parallel_upper_bound(w, name, min(steps, 2), workers)
            line["cpu_parallel_upper_bound"].update({"cpu_model": info["model"], "nproc": info["nproc"]})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
