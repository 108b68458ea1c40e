"""The oracle's restatement of the reference's LinearSolver plugin slot
(CurrentLinearSolver::Solve with BuildInversePreconditioner,
LinearSolver.cpp:30-131) against the compiled reference's own
GetCurrentLinearSolver().Solve, bit for bit in FP32, and its FP64
instantiation against a dense solve.

Cases: the implicit system of the config-1 beam at step 1 (1,210 block rows,
from the reference fixture), and small systems that take every branch: a
missing diagonal block, a singular one, one with 1e-8 < |det| <= 1e-6 (the
Inverse floor gives an Identity block), Jacobi off, maxIterations <= 0
(3 n), an iteration cap that stops early, non-finite b, tolerance <= 0, and
an empty system.
"""

import os

import numpy as np
import pytest

from checkers import oracle_linear_solve, ref_linear_solve, reference_available

needs_ref = pytest.mark.skipif(not reference_available(), reason="reference build not present")


def dense(rs, col, vals):
    n = len(rs) - 1
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        for k in range(rs[r], rs[r + 1]):
            c = col[k]
            A[3 * r:3 * r + 3, 3 * c:3 * c + 3] = vals[k]
            if c != r:
                A[3 * c:3 * c + 3, 3 * r:3 * r + 3] = vals[k].T
    return A


def config1(golden_dir):
    s = np.load(os.path.join(golden_dir, "config1_system_step1.npz"))
    return s["row_start"], s["col"], s["vals"].astype(np.float64), s["rhs"].astype(np.float64)


def small_cases():
    """(name, row_start, col, vals, b, max_it, tol, jacobi) of small systems."""
    rng = np.random.default_rng(7)
    n = 4
    rs = np.array([0, 2, 4, 6, 7], np.int32)
    col = np.array([0, 1, 1, 2, 2, 3, 3], np.int32)

    def spd():
        vals = np.zeros((7, 3, 3))
        for k, (r, c) in enumerate([(0, 0), (0, 1), (1, 1), (1, 2), (2, 2), (2, 3), (3, 3)]):
            vals[k] = (np.eye(3) * 6.0 + 0.3 * rng.normal(size=(3, 3))) if r == c else 0.5 * rng.normal(size=(3, 3))
            if r == c:
                vals[k] = 0.5 * (vals[k] + vals[k].T)
        return vals

    b = rng.normal(size=(n, 3))
    base = spd()
    out = [("spd", rs, col, base, b, 64, 1e-6, True),
           ("jacobi_off", rs, col, base, b, 64, 1e-6, False),
           ("max_it_zero_means_3n", rs, col, base, b, 0, 1e-6, True),
           ("capped", rs, col, base, b, 2, 1e-9, True),
           ("tol_zero", rs, col, base, b, 64, 0.0, True),
           ("tol_nan", rs, col, base, b, 64, float("nan"), True)]
    bad = b.copy()
    bad[2, 1] = np.inf
    out.append(("nonfinite_b", rs, col, base, bad, 64, 1e-6, True))
    sing = base.copy()
    sing[4] = np.diag([1.0, 1.0, 0.0])
    out.append(("singular_diagonal", rs, col, sing, b, 64, 1e-6, True))
    floor = base.copy()
    floor[4] = np.diag([1e-2, 1e-2, 5e-3])  # det 5e-7: accepted, Inverse gives Identity
    floor[3] *= 1e-3                        # weak couplings keep the system SPD
    floor[5] *= 1e-3
    out.append(("inverse_floor", rs, col, floor, b, 64, 1e-6, True))
    nodiag_rs = np.array([0, 2, 3, 5, 6], np.int32)
    nodiag_col = np.array([0, 1, 2, 2, 3, 3], np.int32)  # row 1 has no (1, 1) block
    nd = np.stack([base[0], base[1], base[3], base[4], base[5], base[6]])
    out.append(("missing_diagonal", nodiag_rs, nodiag_col, nd, b, 64, 1e-6, True))
    out.append(("empty", np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros((0, 3, 3)), np.zeros((0, 3)), 64,
                1e-6, True))
    return out


@needs_ref
def test_linear_solve_fp32_bitwise_on_the_config1_system(golden_dir):
    rs, col, vals, b = config1(golden_dir)
    xo, io, ro, co = oracle_linear_solve(rs, col, vals, b, 200, 1e-6, True, precision=32)
    xr, ir, rr, cr = ref_linear_solve(rs, col, vals, b, 200, 1e-6, True)
    assert (io, co) == (ir, cr) and np.float32(ro) == np.float32(rr)
    assert np.array_equal(xo.astype(np.float32), xr)


@needs_ref
@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c[0])
def test_linear_solve_fp32_bitwise_on_small_cases(case):
    name, rs, col, vals, b, max_it, tol, jacobi = case
    xo, io, ro, co = oracle_linear_solve(rs, col, vals, b, max_it, tol, jacobi, precision=32)
    xr, ir, rr, cr = ref_linear_solve(rs, col, vals, b, max_it, tol, jacobi)
    assert (xo is None) == (xr is None), name
    assert (io, co) == (ir, cr), name
    assert np.float32(ro) == np.float32(rr) or (np.isinf(ro) and np.isinf(rr)), name
    if xo is not None:
        assert np.array_equal(xo.astype(np.float32), xr), name


def test_linear_solve_branches_fp64():
    cases = {c[0]: c for c in small_cases()}
    for name in ("missing_diagonal", "singular_diagonal", "nonfinite_b", "tol_zero", "tol_nan"):
        _, rs, col, vals, b, max_it, tol, jacobi = cases[name]
        x, it, res, conv = oracle_linear_solve(rs, col, vals, b, max_it, tol, jacobi)
        assert x is None and it == 0 and not conv, name
    x, it, res, conv = oracle_linear_solve(*cases["empty"][1:])
    assert conv and it == 0
    _, rs, col, vals, b, max_it, tol, jacobi = cases["capped"]
    x, it, res, conv = oracle_linear_solve(rs, col, vals, b, max_it, tol, jacobi)
    assert it == 2 and not conv
    for name in ("spd", "jacobi_off", "max_it_zero_means_3n", "inverse_floor"):
        _, rs, col, vals, b, max_it, tol, jacobi = cases[name]
        x, it, res, conv = oracle_linear_solve(rs, col, vals, b, max_it, 1e-12, jacobi)
        A = dense(rs, col, vals)
        ref = np.linalg.solve(A, b.reshape(-1)).reshape(-1, 3)
        assert conv and np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max(), name
        assert np.isclose(res, np.linalg.norm(b.reshape(-1) - A @ x.reshape(-1)), rtol=1e-3, atol=1e-14), name


def test_linear_solve_fp64_on_the_config1_system_matches_a_dense_solve(golden_dir):
    rs, col, vals, b = config1(golden_dir)
    x, it, res, conv = oracle_linear_solve(rs, col, vals, b, 1000, 1e-12, True)
    A = dense(rs, col, vals)
    ref = np.linalg.solve(A, b.reshape(-1)).reshape(-1, 3)
    assert conv and np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()
