"""The headline workload (batched + differentiable bodies, FP32 storage of the
upper triangle, forward + adjoint) against the FP64 oracle, body by body.

Reduced size: the batch128 generator (Kuhn split, h 1/32 m, per-tet
log E ~ N(log 1e5, 0.2), nu 0.3, x = 0 face fixed, per-body targets on the +x
face, dt 1/60 s, forward rtol 1e-6, adjoint rtol 1e-7) with 6 bodies of
n ~ U{3..6} cells per axis, 8 substeps checkpointed every 4, run as ONE
batched PK_FP32_UPPER system. Each body is compared with its own FP64 oracle
world (the oracle has no batching: one world per body is the reference's
semantics):
  final x and v ............................ relative L2 <= 1e-4
  dL/dlog E and dL/dv0 (oracle adjoint) .... relative L2 <= 1e-4
  forward CG iterations of the last substep  within +-2 (FP32 storage rounds
                                             A; the FP64 oracle does not)
Full size (128 bodies, 3,203,808 tets), one substep through properties that
hold at any size: the fixed faces stay at rest, every body's solution meets
the stopping rule on the TRUE residual b - A dv of the stored matrix, and the
recorded forward + adjoint substep gives finite gradients.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from checkers import Oracle
from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import Pattern

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def reduced():
    w = W.batch128(bodies=6, lo=3, hi=6, steps=8)
    w.checkpoint = 4
    return w


def oracle_body(wb):
    """FP64 oracle forward (x_S, v_S, last CG iterations) and adjoint of one body."""
    def world():
        o = Oracle(64, wb.rest, wb.tets, wb.young, wb.poisson, wb.damping, wb.density)
        o.set_params(wb.gravity, 1, wb.tol, wb.max_iter)
        o.set_fixed(wb.fixed)
        o.set_state(wb.x0, wb.v0)
        return o

    fwd = world()
    for _ in range(wb.steps):
        assert fwd.step(wb.dt) == 0
    x, v = fwd.state()
    its = fwd.cg()[0]
    fwd.close()
    _, xbar = W.loss_and_seed(wb, x)
    adj = world()
    gE, _, gv, _ = adj.adjoint(wb.dt, wb.steps, xbar, None, tol=wb.adj_tol, max_iter=wb.max_iter)
    adj.close()
    return x, v, its, gE, gv


def test_headline_fp32_upper_fwd_adj_matches_fp64_oracle_per_body():
    w = reduced()
    print(f"case headline reduced: {w.n_bodies} bodies, {w.n_tets} tets, fp32 upper, {w.steps} substeps",
          flush=True)
    s = W.make_system(w, N.PK_FP32_UPPER)
    s.record(w.dt, w.steps, w.checkpoint)
    x, v = s.state()
    it, _, conv, st = s.results()
    assert (st == N.GROUP_INTEGRATED).all() and conv.all()
    _, xbar = W.loss_and_seed(w, x)
    gE, _, gv, adj_its = s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter)
    assert (adj_its > 0).all() and (adj_its < w.max_iter).all()
    s.close()
    for b in range(w.n_bodies):
        wb = w.body(b)
        n0, n1 = w.body_nodes[b]
        t0, t1 = w.body_tets[b]
        xo, vo, ito, oE, ov = oracle_body(wb)
        ex, ev = rel(x[n0:n1], xo), rel(v[n0:n1], vo)
        eE, eV = rel(gE[t0:t1], oE), rel(gv[n0:n1], ov)
        print(f"  body {b}: {t1 - t0} tets, rel err x {ex:.2e} v {ev:.2e} dL/dlogE {eE:.2e} dL/dv0 {eV:.2e}, "
              f"CG {it[b]} vs {ito}", flush=True)
        assert ex <= 1e-4 and ev <= 1e-4, (b, ex, ev)
        assert eE <= 1e-4 and eV <= 1e-4, (b, eE, eV)
        assert abs(int(it[b]) - ito) <= 2, (b, it[b], ito)


def test_headline_full_size_substep_properties():
    w = W.batch128()
    print(f"case headline full size: {w.n_bodies} bodies, {w.n_tets} tets, fp32 upper, one substep", flush=True)
    s = W.make_system(w, N.PK_FP32_UPPER)
    s.assemble(w.dt)
    vals, b, minv = s.matrix()
    s.step(w.dt)
    dv = s.solution()
    it, _, conv, st = s.results()
    assert (st == N.GROUP_INTEGRATED).all() and conv.all()
    x, v = s.state()
    fixed = w.fixed == 1
    assert np.array_equal(x[fixed], w.x0[fixed]) and not v[fixed].any()
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed, node_group=w.node_group)
    A = sp.bsr_matrix((vals, p.full_col, p.full_start), shape=(3 * p.nd, 3 * p.nd))
    r = b - (A @ dv.reshape(-1)).reshape(-1, 3)
    Mr = np.einsum("nij,nj->ni", minv, r)
    Mb = np.einsum("nij,nj->ni", minv, b)
    body_of_row = w.node_group[np.flatnonzero(p.dyn >= 0)]
    rr = np.bincount(body_of_row, weights=np.sum(r * Mr, axis=1), minlength=w.n_bodies)
    bb = np.bincount(body_of_row, weights=np.sum(b * Mb, axis=1), minlength=w.n_bodies)
    ratio = np.sqrt(rr / bb)
    print(f"  true residual / initial (M^-1 norm): max {ratio.max():.3e} median {np.median(ratio):.3e}, "
          f"CG iterations {it.min()}..{it.max()}", flush=True)
    # the stopping rule r^T M^-1 r <= tol^2 r0^T M^-1 r0 holds for the CG's
    # recurrence residual; the true residual of the stored matrix may differ
    # from it by the FP32 rounding of the products, bounded here by tol again
    assert (ratio <= 2.0 * w.tol).all(), ratio.max()
    # one recorded forward + adjoint substep: finite gradients
    s.set_state(w.x0, w.v0)
    s.record(w.dt, 1, 1)
    xs, _ = s.state()
    _, xbar = W.loss_and_seed(w, xs)
    gE, gx, gv, adj_its = s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter)
    assert np.isfinite(gE).all() and np.isfinite(gx).all() and np.isfinite(gv).all()
    assert np.abs(gE).max() > 0 and (adj_its > 0).all()
    s.close()
