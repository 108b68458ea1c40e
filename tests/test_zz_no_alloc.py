"""pk_bsr_pcg_noalloc (work arrays in the library's file-scope device array,
no allocation) against pk_bsr_pcg (the same kernels, work arrays allocated
per call), through the C ABI: bit-identical solutions and results on two
calls in a row, the second on a smaller system, then the first system again;
a refused solve; and a system above PK_LINEAR_NOALLOC_MAX_ROWS refused with
PK_ERR_ARG."""

import ctypes as C
import os

import numpy as np
import pytest
import torch

from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import Pattern

pytestmark = pytest.mark.gpu

MAX_ROWS = 262144  # PK_LINEAR_NOALLOC_MAX_ROWS


def system(n):
    """The assembled implicit system of a twisted n^3-cell Kuhn beam (full BSR)."""
    w = W.twist64(n=n, h=1.0 / n)
    s = W.make_system(w, N.PK_FP64)
    s.assemble(w.dt)
    vals, b, _ = s.matrix()
    s.close()
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed)
    return p.full_start, p.full_col, vals, b, w.max_iter, w.tol


def solve(entry, start, col, vals, b, max_it, tol):
    dev = torch.device("cuda")
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (start.astype(np.int32), col.astype(np.int32),
                                                                     vals.astype(np.float64), b.astype(np.float64))]
    x = torch.full_like(t[3], 3.0)
    torch.cuda.synchronize()
    res = N.LinearResult()
    rc = getattr(N.lib(), entry)(len(start) - 1, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(),
                                 x.data_ptr(), int(max_it), float(tol), 1, None, C.byref(res))
    return rc, x.cpu().numpy(), (res.iterations, res.converged, res.solution_valid, res.residual_norm)


def same(a, b):
    assert a[0] == b[0] == N.PK_OK
    assert a[2] == b[2], (a[2], b[2])
    if a[2][2]:
        assert np.array_equal(a[1], b[1])


def test_noalloc_equals_the_allocating_solve_on_two_calls_in_a_row():
    big, small = system(24), system(10)
    print(f"case no-alloc: {len(big[0]) - 1} rows, then {len(small[0]) - 1} rows", flush=True)
    r1 = solve("pk_bsr_pcg_noalloc", *big)
    same(r1, solve("pk_bsr_pcg", *big))
    assert r1[2][1] and r1[2][3] > 0
    r2 = solve("pk_bsr_pcg_noalloc", *small)
    same(r2, solve("pk_bsr_pcg", *small))
    same(solve("pk_bsr_pcg_noalloc", *big), r1)  # nothing of the smaller call carries over
    print(f"  iterations {r1[2][0]} and {r2[2][0]}, bit-identical", flush=True)


def test_noalloc_refusals():
    start, col, vals, b, max_it, tol = system(4)
    vals = vals.copy()
    diag = [k for r in range(len(start) - 1) for k in range(start[r], start[r + 1]) if col[k] == r]
    vals[diag[3]] = 0.0  # a singular diagonal block: the preconditioner refuses
    a = solve("pk_bsr_pcg_noalloc", start, col, vals, b, max_it, tol)
    same(a, solve("pk_bsr_pcg", start, col, vals, b, max_it, tol))
    assert a[2][2] == 0
    res = N.LinearResult()
    rc = N.lib().pk_bsr_pcg_noalloc(MAX_ROWS + 1, None, None, None, None, None, 10, 1e-6, 1, None, C.byref(res))
    assert rc == N.PK_ERR_ARG and b"exceed" in N.lib().pk_last_error()
