"""The GPU linear solve (include/physik_amd_linear.h, the reference's
LinearSolver plugin slot) against the oracle's FP64 restatement of
CurrentLinearSolver::Solve and the compiled reference's FP32 one.

Bars:
  pk_linear_solve (host, the reference's float layout) vs the FP64 oracle:
      same outcome (solution kept or cleared, converged), iterations +-1,
      x within 1e-6 relative (the float rounding of the FP64 result)
  the same vs the reference's FP32 solve: x within 1e-4 relative
  pk_bsr_pcg (device memory, FP64) on a 24^3-cell twisted Kuhn system
      (the GPU system's own assembled matrix, 15,000 block rows, ~230
      iterations at rtol 1e-6) vs the FP64 oracle and vs the system's own CG
      solution of the same substep: iterations +-1, the solution dv within
      1e-7 relative and the final residual norm within 5 % (three FP64 CGs
      that sum in different orders: the last iterations' rounding, amplified
      by the system's conditioning, measured 1.5e-8 and 2 %; the positions
      x + dt dv they lead to agree to 1e-9 as in the step tests)
"""

import ctypes as C
import os

import numpy as np
import pytest
import torch

from checkers import oracle_linear_solve, ref_linear_solve, reference_available
from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import Pattern

from test_oracle_linear import config1, small_cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def gpu_linear_solve(rs, col, vals, b, max_it, tol, jacobi=True):
    """pk_linear_solve with the reference's layout (column-major float Mat3)."""
    n = len(rs) - 1
    rs32 = np.ascontiguousarray(rs, np.int32)
    col32 = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
    cm = np.ascontiguousarray(np.transpose(np.asarray(vals, np.float64), (0, 2, 1)), np.float32)
    cm = cm if len(cm) else np.zeros((1, 3, 3), np.float32)
    bb = np.ascontiguousarray(b, np.float32) if n else np.zeros((1, 3), np.float32)
    x = np.full((max(n, 1), 3), np.float32(7.0))
    res = N.LinearResult()
    N.check(N.lib().pk_linear_solve(n, N.ptr(rs32), N.ptr(col32), N.ptr(cm), N.ptr(bb), N.ptr(x), int(max_it),
                                    float(tol), int(jacobi), C.byref(res)), "pk_linear_solve")
    return (x[:n] if res.solution_valid else None), res.iterations, res.residual_norm, bool(res.converged)


def full_from_upper(rs, col, vals):
    """Both triangles, rows sorted by column (row-major blocks)."""
    n = len(rs) - 1
    rows = np.repeat(np.arange(n), np.diff(rs))
    off = rows != col
    r = np.concatenate([rows, col[off]])
    c = np.concatenate([col, rows[off]])
    v = np.concatenate([vals, np.transpose(vals[off], (0, 2, 1))])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    start = np.zeros(n + 1, np.int64)
    np.add.at(start, r + 1, 1)
    return np.cumsum(start).astype(np.int32), c.astype(np.int32), np.ascontiguousarray(v)


def upper_from_full(start, fcol, fvals):
    n = len(start) - 1
    rows = np.repeat(np.arange(n), np.diff(start))
    keep = fcol >= rows
    rs = np.zeros(n + 1, np.int64)
    np.add.at(rs, rows[keep] + 1, 1)
    return np.cumsum(rs).astype(np.int32), fcol[keep], fvals[keep]


def bsr_pcg(start, fcol, fvals, b, max_it, tol, jacobi=True):
    dev = torch.device("cuda")
    t_rs = torch.from_numpy(np.ascontiguousarray(start, np.int32)).to(dev)
    t_col = torch.from_numpy(np.ascontiguousarray(fcol, np.int32)).to(dev)
    t_vals = torch.from_numpy(np.ascontiguousarray(fvals, np.float64)).to(dev)
    t_b = torch.from_numpy(np.ascontiguousarray(b, np.float64)).to(dev)
    t_x = torch.empty_like(t_b)
    torch.cuda.synchronize()
    res = N.LinearResult()
    N.check(N.lib().pk_bsr_pcg(len(start) - 1, t_rs.data_ptr(), t_col.data_ptr(), t_vals.data_ptr(), t_b.data_ptr(),
                               t_x.data_ptr(), int(max_it), float(tol), int(jacobi), None, C.byref(res)), "pk_bsr_pcg")
    return t_x.cpu().numpy(), res


def check_against_oracle(name, rs, col, vals, b, max_it, tol, jacobi):
    xg, ig, rg, cg = gpu_linear_solve(rs, col, vals, b, max_it, tol, jacobi)
    xo, io, ro, co = oracle_linear_solve(rs, col, vals, b, max_it, tol, jacobi)
    print(f"  {name}: gpu iterations {ig} converged {cg} residual {rg:.3e}; oracle {io} {co} {ro:.3e}", flush=True)
    assert (xg is None) == (xo is None), name
    assert cg == co and abs(ig - io) <= 1, name
    if xo is not None:
        assert rel(xg, xo) <= 1e-6, (name, rel(xg, xo))
        if np.isfinite(ro) and ro > 0:
            assert abs(rg - ro) <= 1e-3 * ro + 1e-12, name
    if reference_available() and xo is not None and co:
        xr, ir, _, cr = ref_linear_solve(rs, col, vals, b, max_it, tol, jacobi)
        assert cr and rel(xg, xr) <= 1e-4, (name, rel(xg, xr))


def test_plugin_solve_on_the_config1_system(golden_dir):
    print("case linear solve: config-1 system", flush=True)
    rs, col, vals, b = config1(golden_dir)
    check_against_oracle("config1", rs, col, vals, b, 200, 1e-6, True)


@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c[0])
def test_plugin_solve_on_small_cases(case):
    name, rs, col, vals, b, max_it, tol, jacobi = case
    print(f"case linear solve: {name}", flush=True)
    check_against_oracle(name, rs, col, vals, b, max_it, tol, jacobi)


def test_device_solve_matches_the_oracle_and_the_system_cg():
    w = W.twist64(n=24, h=1.0 / 24)
    print(f"case device solve: twist24, {w.n_tets} tets", flush=True)
    s = W.make_system(w, N.PK_FP64)
    s.assemble(w.dt)
    fvals, b, _ = s.matrix()
    s.step(w.dt)
    dv = s.solution()
    its_sys = int(s.results()[0][0])
    s.close()
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed)
    x, res = bsr_pcg(p.full_start, p.full_col, fvals, b, w.max_iter, w.tol)
    rs, col, vals = upper_from_full(p.full_start, p.full_col, fvals)
    xo, io, ro, co = oracle_linear_solve(rs, col, vals, b, w.max_iter, w.tol, True)
    print(f"  rows {p.nd}: gpu {res.iterations} iterations, residual {res.residual_norm:.4e}; oracle {io}, "
          f"{ro:.4e}; system CG {its_sys}; rel err vs oracle {rel(x, xo):.2e}, vs system {rel(x, dv):.2e}",
          flush=True)
    assert res.solution_valid and res.converged and co
    assert abs(res.iterations - io) <= 1 and abs(res.iterations - its_sys) <= 1
    assert rel(x, xo) <= 1e-7 and rel(x, dv) <= 1e-7
    assert abs(res.residual_norm - ro) <= 0.05 * ro


def test_device_solve_of_an_empty_system_and_bad_arguments():
    res = N.LinearResult()
    assert N.lib().pk_bsr_pcg(0, None, None, None, None, None, 10, 1e-6, 1, None, C.byref(res)) == N.PK_OK
    assert res.converged and res.solution_valid and res.iterations == 0
    assert N.lib().pk_bsr_pcg(-1, None, None, None, None, None, 10, 1e-6, 1, None, C.byref(res)) == N.PK_ERR_ARG
    assert N.lib().pk_bsr_pcg(4, None, None, None, None, None, 10, 1e-6, 1, None, C.byref(res)) == N.PK_ERR_ARG
    # a refused tolerance needs no device memory at all
    assert N.lib().pk_bsr_pcg(4, None, None, None, None, None, 10, 0.0, 1, None, C.byref(res)) == N.PK_OK
    assert not res.solution_valid and not res.converged
