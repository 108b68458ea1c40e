"""Summaries of rocprofv3 output for profiles/ (text only, a few KB each).

    python -m physik_amd.profsum traffic --fetch DIR --write DIR --diag FILE \
        --key batch128/fp32u --out profiles/hbm_traffic_batch128_fp32u.md --cmd "..." --cmd "..."
    python -m physik_amd.profsum stats --dir DIR --out profiles/k_spmv_..._trace.md --cmd "..." [--title T]

`traffic` reads two counter passes over `python -m physik_amd.diag` (one
with FETCH_SIZE, one with WRITE_SIZE; each directory holds the pass's
*counter_collection.csv) and the JSON line that diag run printed. The
calibration copy (k_copy, exactly copy_bytes each way per dispatch) gives the
scale of each counter; every kernel's bytes per dispatch are scaled by it.
The SpMV's bytes are set against the algorithmic bytes of the run's own work
(physik_amd.roofline), the assembly kernels' against one assembly's. The
result is a markdown table and an entry of profiles/spmv_traffic.json, which
bench.py reads for its `traffic` fields.

`stats` turns a --kernel-trace --stats run's *kernel_stats.csv into a table
(calls, total, average, share per kernel).
"""

from __future__ import annotations

import argparse
import csv
import glob
import json
import os

from .roofline import assembly_bytes, spmv_bytes

# one assembly's kernels: the slot gather is k_assemble_pairs (by row, tet
# pairs) or, for a system with a row larger than a chunk, k_assemble
ASSEMBLY_KERNELS = {"element": ("k_element",), "assemble": ("k_assemble_pairs", "k_assemble"), "rhs": ("k_rhs",)}


def _csv_rows(directory, suffix):
    paths = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not paths:
        raise SystemExit(f"no *{suffix} under {directory}")
    rows = []
    for p in paths:
        with open(p, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def _counter_per_dispatch(directory, counter):
    """{kernel name: [KiB per dispatch, in dispatch order]} of one pass."""
    per = {}
    seen = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        key = (r["Kernel_Name"], int(r.get("Dispatch_Id", len(seen))))
        # a counter summed over its instances may come as several lines per dispatch
        seen[key] = seen.get(key, 0.0) + float(r["Counter_Value"])
    for (name, disp), val in sorted(seen.items(), key=lambda kv: kv[0][1]):
        per.setdefault(name, []).append(val)
    return per


def _short(name):
    return name.split("(")[0].strip()


def _find(per, stem):
    for name in per:
        if stem in _short(name).split("::")[-1].split("<")[0]:
            yield name


def traffic(args):
    with open(args.diag) as f:
        diag = json.loads([ln for ln in f if ln.strip().startswith("{")][-1])
    fetch = _counter_per_dispatch(args.fetch, "FETCH_SIZE")
    write = _counter_per_dispatch(args.write, "WRITE_SIZE")
    copy_bytes = float(diag["copy_bytes"])
    cf = [v for n in _find(fetch, "k_copy") for v in fetch[n]]
    cw = [v for n in _find(write, "k_copy") for v in write[n]]
    if not cf or not cw:
        raise SystemExit("the calibration copy k_copy is missing from a pass")
    f_scale = copy_bytes / (1024.0 * sum(cf) / len(cf))
    w_scale = copy_bytes / (1024.0 * sum(cw) / len(cw))
    prec = args.key.split("/")[1]
    s = 4 if prec.startswith("fp32") else 8
    upper = prec.endswith("u")
    n_spmv = int(diag["spmv_launches"])
    spmv_alg = spmv_bytes(diag["spmv_blocks"], diag["spmv_rows"], diag["spmv_body_launches"], s,
                          diag["spmv_upper_blocks"] if upper else None) / max(n_spmv, 1)
    stored = diag["nnzb_upper"] if upper else diag["nnzb_full"]
    asm_alg = assembly_bytes(diag["tets"], diag["nodes"], stored, diag["dynamic_rows"], s)
    table, asm, spmv = [], {}, None
    for name in fetch:
        if "k_copy" in name:
            continue
        fv = [v * 1024.0 * f_scale for v in fetch[name]]
        wv = [v * 1024.0 * w_scale for v in write.get(name, [0.0] * len(fv))]
        if "spmv" in _short(name):
            keep = list(range(min(n_spmv, len(fv))))  # the launches that did work come first
            note = f"first {len(keep)} (working launches)"
        else:
            top = max(fv) if fv else 0.0
            keep = [i for i, v in enumerate(fv) if v > 0.05 * top] or list(range(len(fv)))
            note = f"{len(keep)} of {len(fv)}"
        fm = sum(fv[i] for i in keep) / max(len(keep), 1)
        wm = sum(wv[i] for i in keep if i < len(wv)) / max(len(keep), 1)
        alg = None
        if "spmv" in _short(name):
            alg = spmv_alg
            spmv = (fm, wm)
        for k, stems in ASSEMBLY_KERNELS.items():
            if _short(name).split("::")[-1].split("<")[0] in stems:
                asm[k] = fm + wm
        table.append((_short(name), len(fv), note, fm, wm, alg))
    lines = [f"# HBM traffic per dispatch: {args.key}", "",
             "Two counter passes over the same program (`physik_amd.diag`: a "
             f"{copy_bytes / 2**20:.0f} MiB device copy, then {diag['steps']} implicit step(s) with CG launch shape "
             f"{','.join(str(v) for v in diag['shape'])}, staging {diag['staging']}):", "", "```"]
    lines += args.cmd + ["```", "", "## Calibration", "",
                         f"`k_copy` moves exactly {int(copy_bytes):,} B each way per dispatch ({len(cf)} dispatches). "
                         f"Counter readings per dispatch (KiB): FETCH_SIZE {min(cf):,.0f}-{max(cf):,.0f}, WRITE_SIZE "
                         f"{min(cw):,.0f}-{max(cw):,.0f}. Scale factors applied below: FETCH x {f_scale:.4f}, "
                         f"WRITE x {w_scale:.4f}.", "", "## Per kernel (calibrated, MB = 1e6 B)", "",
                         "| kernel | dispatches | averaged over | fetch MB | write MB | total MB | algorithmic MB |",
                         "|---|---|---|---|---|---|---|"]
    for name, nd, note, fm, wm, alg in table:
        lines.append(f"| `{name}` | {nd} | {note} | {fm / 1e6:.1f} | {wm / 1e6:.1f} | {(fm + wm) / 1e6:.1f} | "
                     f"{'' if alg is None else f'{alg / 1e6:.1f}'} |")
    lines.append("")
    if spmv:
        lines.append(f"SpMV work of the run: {diag['spmv_blocks']:,} blocks ({diag['spmv_upper_blocks']:,} stored "
                     f"upper blocks) and {diag['spmv_rows']:,} rows over {n_spmv} launches; the SpMV moves "
                     f"{(spmv[0] + spmv[1]) / spmv_alg:.2f}x its algorithmic bytes per launch.")
    if len(asm) == 3:
        tot = sum(asm.values())
        lines.append(f"Assembly (element + slot gather + right-hand side): {tot / 1e6:.1f} MB per substep against "
                     f"{asm_alg / 1e6:.1f} MB algorithmic ({tot / asm_alg:.2f}x).")
    lines.append("")
    with open(args.out, "w") as f:
        f.write("\n".join(lines))
    path = os.path.join(os.path.dirname(args.out), "spmv_traffic.json")
    try:
        with open(path) as f:
            db = json.load(f)
    except (OSError, ValueError):
        db = {}
    entry = {"source": args.out}
    if spmv:
        entry.update({"bytes": spmv[0] + spmv[1], "fetch": spmv[0], "write": spmv[1], "algorithmic": spmv_alg,
                      "shape": diag["shape"]})
    if len(asm) == 3:
        entry["assembly"] = {**asm, "algorithmic": asm_alg}
    db[args.key] = entry
    with open(path, "w") as f:
        json.dump(db, f, indent=1)
    print("\n".join(lines))


def stats(args):
    rows = _csv_rows(args.dir, "kernel_stats.csv")
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    # per kernel, the dispatches that did work: a CG launch issued after its
    # body stopped exits in a few microseconds
    long = {}
    for r in _csv_rows(args.dir, "kernel_trace.csv"):
        d = (float(r["End_Timestamp"]) - float(r["Start_Timestamp"])) / 1e3
        if d >= args.min_us:
            n, t = long.get(r["Kernel_Name"], (0, 0.0))
            long[r["Kernel_Name"]] = (n + 1, t + d)
    lines = [f"# {args.title}", "", "```"] + args.cmd + ["```", "",
                                                         f"| kernel | calls | total ms | average us | share | "
                                                         f"calls >= {args.min_us:g} us | their average us |",
                                                         "|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        n, t = long.get(r["Name"], (0, 0.0))
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {100 * float(r['TotalDurationNs']) / total:.1f} % | "
                     f"{n} | {t / max(n, 1):.2f} |")
    lines.append("")
    lines += args.note
    with open(args.out, "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--diag", required=True)
    t.add_argument("--key", required=True, help="workload/precision, e.g. batch128/fp32u")
    t.add_argument("--out", required=True)
    t.add_argument("--cmd", action="append", default=[])
    st = sub.add_parser("stats")
    st.add_argument("--dir", required=True)
    st.add_argument("--out", required=True)
    st.add_argument("--title", default="rocprofv3 kernel trace")
    st.add_argument("--cmd", action="append", default=[])
    st.add_argument("--note", action="append", default=[], help="a paragraph after the table")
    st.add_argument("--min-us", type=float, default=5.0)
    args = ap.parse_args()
    (traffic if args.mode == "traffic" else stats)(args)


if __name__ == "__main__":
    main()
