"""Reverse mode over contact forces: point connections and collision-sphere
contacts act in every recorded substep, and pk_system_backward
differentiates their forces. Checked against the FP64 oracle's adjoint,
which central finite differences of the oracle's forward map pin
(tests/test_oracle_adjoint.py, contact cases).

Bars:
  FP64 full and upper-triangle storage vs the oracle: dL/dlogE, dL/dx0 and
      dL/dv0 within 1e-8 relative L2, adjoint CG iterations +-1, the
      recorded end state within 1e-11
  FP32 upper-triangle storage vs the oracle: gradients within 1e-4
  a batch with a sphere touching every body: each body's gradients equal
      its single-body run bit for bit
"""

import numpy as np
import pytest

from checkers import Oracle
from physik_amd import native as N
from physik_amd import workloads as W

from test_gpu_adjoint import small

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def contacts(w):
    """Two point connections at the free end (the barycentre of its outermost
    tet's nodes to a fixed target, one loss-face node to another) and a sphere
    under the free end."""
    end = w.loss_nodes
    tet = np.asarray(w.tets[int(np.argmax(w.x0[np.asarray(w.tets)].mean(1)[:, 0]))])
    conns = (np.array([tet, [end[-1]] * 4]), np.array([[0.1, 0.2, 0.3, 0.4], [1.0, 0.0, 0.0, 0.0]]),
             np.array([w.x0[tet].mean(0) + [0.03, -0.02, 0.04], w.x0[end[-1]] + [0.0, 0.05, 0.0]]),
             np.array([300.0, 150.0]), np.array([0.7, 0.2]))
    lo, hi = w.x0.min(0), w.x0.max(0)
    sph = (np.array([[hi[0], 0.5 * (lo[1] + hi[1]), lo[2] - 0.02]]), np.array([0.15]), np.array([2000.0]),
           np.array([1.5]))
    return conns, sph


def gpu_run(w, precision, conns, sph, every=2):
    s = W.make_system(w, precision)
    if conns is not None:
        s.set_connections(*conns)
    if sph is not None:
        s.set_spheres(*sph)
    s.record(w.dt, w.steps, every)
    x, v = s.state()
    _, xbar = W.loss_and_seed(w, x)
    vbar = 0.2 * v
    gE, gx, gv, its = s.backward(xbar, vbar, tol=w.adj_tol, max_iter=w.max_iter)
    s.close()
    return x, xbar, vbar, gE, gx, gv, its


def oracle_run(w, conns, sph, xbar, vbar):
    o = Oracle(64, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
    o.set_params(w.gravity, 1, w.tol, w.max_iter)
    o.set_fixed(w.fixed)
    o.set_state(w.x0, w.v0)
    if sph is not None:
        o.set_spheres(*sph)
    if conns is not None:
        for i in range(len(conns[0])):
            o.add_point_connection(conns[0][i], conns[1][i], conns[2][i], conns[3][i], conns[4][i])
    gE, gx, gv, its = o.adjoint(w.dt, w.steps, xbar, vbar, tol=w.adj_tol, max_iter=w.max_iter)
    return o.state()[0], gE, gx, gv, its, o.last_contact_count()


@pytest.mark.parametrize("kind", ["connections", "sphere", "both"])
@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP64_UPPER])
def test_contact_adjoint_matches_oracle(precision, kind):
    w = small(tol=1e-11)
    conns, sph = contacts(w)
    conns = None if kind == "sphere" else conns
    sph = None if kind == "connections" else sph
    print(f"case contact adjoint {kind} precision {precision}", flush=True)
    x, xbar, vbar, gE, gx, gv, its = gpu_run(w, precision, conns, sph)
    xo, oE, ox, ov, oits, ncontact = oracle_run(w, conns, sph, xbar, vbar)
    print(f"  oracle contacts in the last substep {ncontact}; rel errors gE {rel(gE, oE):.2e} gx {rel(gx, ox):.2e} "
          f"gv {rel(gv, ov):.2e}", flush=True)
    if sph is not None:
        assert ncontact > 0
    assert rel(x, xo) <= 1e-11
    assert rel(gE, oE) <= 1e-8 and rel(gx, ox) <= 1e-8 and rel(gv, ov) <= 1e-8
    assert np.abs(its - oits).max() <= 1


def test_contact_terms_change_the_gpu_gradient():
    w = small(tol=1e-11)
    conns, sph = contacts(w)
    with_c = gpu_run(w, N.PK_FP64, conns, sph)
    without = gpu_run(w, N.PK_FP64, None, None)
    assert rel(with_c[4], without[4]) > 1e-3


def test_fp32_upper_contact_adjoint_close_to_oracle():
    w = small(tol=1e-9)
    conns, sph = contacts(w)
    print("case contact adjoint fp32 upper", flush=True)
    _, xbar, vbar, gE, gx, gv, _ = gpu_run(w, N.PK_FP32_UPPER, conns, sph)
    _, oE, ox, ov, _, _ = oracle_run(w, conns, sph, xbar, vbar)
    assert rel(gE, oE) <= 1e-4 and rel(gx, ox) <= 1e-4 and rel(gv, ov) <= 1e-4, (rel(gE, oE), rel(gx, ox),
                                                                                  rel(gv, ov))


def test_batched_contact_adjoint_equals_single_bodies():
    # the bodies of a batch all sit at the origin: one sphere touches each
    w = W.batch128(bodies=3, lo=2, hi=4, h=0.1, steps=4)
    w.young = w.young / 10.0
    sph = (np.array([[0.15, 0.15, 0.15]]), np.array([0.2]), np.array([3000.0]), np.array([2.0]))
    print("case batched contact adjoint", flush=True)
    _, _, _, gE, gx, gv, _ = gpu_run(w, N.PK_FP64, None, sph)
    for b in range(w.n_bodies):
        wb = w.body(b)
        _, _, _, e, gxb, gvb, _ = gpu_run(wb, N.PK_FP64, None, sph)
        n0, n1 = w.body_nodes[b]
        t0, t1 = w.body_tets[b]
        assert np.array_equal(e, gE[t0:t1]) and np.array_equal(gxb, gx[n0:n1]) and np.array_equal(gvb, gv[n0:n1])
