"""GPU reverse mode (pk_system_record / pk_system_backward) against the oracle.

Bars:
  FP64 system vs the FP64 oracle adjoint (same forward and adjoint tolerance)
      dL/dlogE, dL/dv0, dL/dx0: relative L2 <= 1e-8
  FP32-storage system vs the FP64 oracle adjoint: relative L2 <= 1e-4
  checkpoint interval: gradients bit-identical for every interval
  batched bodies: bit-identical to the same bodies reversed one by one
  the GPU adjoint vs central differences of the GPU forward itself (FP64,
      tight CG): relative error <= 1e-5
  diff32 at full size (196,608 tets, 50 substeps, checkpoint 10): 16 random
      parameters vs central differences (h = 1e-6) to 1e-4 relative above the
      measured difference noise, and random directional derivatives to 1e-6
  the same body at 16^3 cells vs the FP64 oracle adjoint: relative L2 <= 1e-8
The oracle adjoint is itself pinned by finite differences on the CPU
(tests/test_oracle_adjoint.py).
"""

import numpy as np
import pytest

from checkers import Oracle
from launch_shapes import shape_for
from physik_amd import meshgen as mg
from physik_amd import native as N
from physik_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def small(n=(4, 3, 3), steps=6, damping=0.3, seed=21, tol=1e-9):
    """A jittered, twisted Kuhn block with per-tet E and initial velocity."""
    nx, ny, nz = n
    h = 0.1
    rest, tets = mg.box_mesh(nx, ny, nz, h, "kuhn")
    rest = mg.jitter(rest, h, 0.1, mg.Pcg32(seed))
    rng = mg.Pcg32(seed, 1)
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(nx, ny, nz, 0)] = 1
    v0 = 0.3 * rng.normal(rest.size).reshape(-1, 3)
    v0[fixed == 1] = 0.0
    w = W._single("adjsmall", rest, tets, 1e4, 0.3, 1000.0, fixed, mg.twist_about_x(rest, nx, ny, nz, h, 0.5), v0,
                  1.0 / 60.0, steps, tol, 2000)
    w.young = np.exp(np.log(1e4) + 0.2 * rng.normal(len(tets)))
    w.damping = np.full(len(tets), damping)
    end = mg.face_nodes(nx, ny, nz, 1)
    w.loss_nodes = end.astype(np.int64)
    w.target = w.x0[end] + np.array([0.0, 0.02, -0.05])
    w.adj_tol = tol
    return w


def gpu_grad(w, precision=N.PK_FP64, every=None, vbar_scale=0.1):
    s = W.make_system(w, precision)
    s.record(w.dt, w.steps, every or w.checkpoint or w.steps)
    x, v = s.state()
    L, xbar = W.loss_and_seed(w, x)
    vbar = 2 * vbar_scale * v
    gE, gx, gv, its = s.backward(xbar, vbar, tol=w.adj_tol, max_iter=w.max_iter)
    x2, v2 = s.state()
    assert np.array_equal(x, x2) and np.array_equal(v, v2), "backward must leave the final state"
    s.close()
    return L + vbar_scale * float(np.sum(v * v)), gE, gx, gv, its, x


def oracle_grad(w, vbar_scale=0.1):
    o = Oracle(64, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
    o.set_params(w.gravity, 1, w.tol, w.max_iter)
    o.set_fixed(w.fixed)
    o.set_state(w.x0, w.v0)
    probe = Oracle(64, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
    probe.set_params(w.gravity, 1, w.tol, w.max_iter)
    probe.set_fixed(w.fixed)
    probe.set_state(w.x0, w.v0)
    for _ in range(w.steps):
        probe.step(w.dt)
    x, v = probe.state()
    _, xbar = W.loss_and_seed(w, x)
    gE, gx, gv, its = o.adjoint(w.dt, w.steps, xbar, 2 * vbar_scale * v, tol=w.adj_tol, max_iter=w.max_iter)
    return gE, gx, gv, its, x


def test_fp64_adjoint_matches_oracle():
    w = small()
    print("case adjsmall fp64", flush=True)
    _, gE, gx, gv, its, x = gpu_grad(w)
    oE, ox, ov, oits, xo = oracle_grad(w)
    assert rel(x, xo) <= 1e-11
    assert rel(gE, oE) <= 1e-8 and rel(gv, ov) <= 1e-8 and rel(gx, ox) <= 1e-8, (rel(gE, oE), rel(gv, ov), rel(gx, ox))
    assert np.abs(its - oits).max() <= 1


def test_fp32_storage_adjoint_close_to_oracle():
    w = small(tol=1e-7)
    print("case adjsmall fp32 storage", flush=True)
    _, gE, gx, gv, _, _ = gpu_grad(w, N.PK_FP32_STORAGE)
    oE, ox, ov, _, _ = oracle_grad(w)
    assert rel(gE, oE) <= 1e-4 and rel(gv, ov) <= 1e-4 and rel(gx, ox) <= 1e-4


def test_checkpoint_interval_does_not_change_gradients():
    w = small(steps=7)
    ref = gpu_grad(w, every=7)
    for every in (1, 2, 3):
        print(f"case checkpoint every {every}", flush=True)
        out = gpu_grad(w, every=every)
        for a, b in zip(out[1:5], ref[1:5]):
            assert np.array_equal(a, b), f"every={every}"


def test_batched_adjoint_equals_single_bodies():
    w = W.batch128(bodies=4, lo=2, hi=4, h=0.1, steps=4)
    w.young = w.young / 10.0  # softer: visible motion in four substeps
    print("case batched adjoint", flush=True)
    s = W.make_system(w)
    s.tune(16, 32, 64, 2)
    s.record(w.dt, w.steps, 2)
    x, _ = s.state()
    _, xbar = W.loss_and_seed(w, x)
    gE, gx, gv, its = s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter)
    for b in range(w.n_bodies):
        wb = w.body(b)
        sb = W.make_system(wb)
        sb.tune(16, 32, 64, 2)
        sb.record(wb.dt, wb.steps, 2)
        xb, _ = sb.state()
        _, xbarb = W.loss_and_seed(wb, xb)
        e, gxb, gvb, _ = sb.backward(xbarb, None, tol=wb.adj_tol, max_iter=wb.max_iter)
        n0, n1 = w.body_nodes[b]
        t0, t1 = w.body_tets[b]
        assert np.array_equal(e, gE[t0:t1]) and np.array_equal(gvb, gv[n0:n1]) and np.array_equal(gxb, gx[n0:n1])
        sb.close()


def test_gpu_adjoint_matches_gpu_finite_differences():
    w = small(steps=4, tol=1e-12)
    L0, gE, _, gv, _, _ = gpu_grad(w)
    h = 1e-6
    rng = np.random.default_rng(3)
    for e in rng.choice(w.n_tets, 3, replace=False):
        vals = []
        for sgn in (1, -1):
            w2 = small(steps=4, tol=1e-12)
            w2.young = w.young.copy()
            w2.young[e] *= np.exp(sgn * h)
            vals.append(gpu_grad(w2)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(gE[e] - fd) <= 1e-5 * abs(fd) + 1e-9 * np.abs(gE).max(), (e, gE[e], fd)
    i = int(np.flatnonzero(w.fixed == 0)[5])
    vals = []
    for sgn in (1, -1):
        w2 = small(steps=4, tol=1e-12)
        w2.young = w.young.copy()
        w2.v0 = w.v0.copy()
        w2.v0[i, 2] += sgn * h
        vals.append(gpu_grad(w2)[0])
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(gv[i, 2] - fd) <= 1e-5 * abs(fd) + 1e-9 * np.abs(gv).max()


def test_backward_without_record_is_rejected():
    s = W.make_system(small(steps=1))
    with pytest.raises(RuntimeError, match="no recorded run"):
        s.backward(None, None)
    s.record(1.0 / 60.0, 1)
    s.set_young(2e4)  # a material change discards the recording
    with pytest.raises(RuntimeError, match="no recorded run"):
        s.backward(None, None)


def test_diff32_gradients_match_finite_differences_on_16_parameters():
    # the differentiable single-body workload at full size: 32^3-cell Kuhn beam
    # (196,608 tets), 50 substeps, checkpoint every 10; forward and adjoint CG
    # tightened to 1e-10 so the differences measure the map, not the solves.
    # One log E or one v0 component moves the loss by ~1e-10 at h = 1e-6,
    # which is also the size of the solver noise in the loss: central
    # differences carry an absolute error of up to ~8e-10 here (measured, the
    # same at CG 1e-12), so each parameter is held to 1e-4 relative plus a
    # floor of 2e-6 * max|gradient| (3e-9 for log E, 1.4e-8 for v0).
    w = W.diff32()
    w.tol = w.adj_tol = 1e-10
    print("case diff32 record/backward", flush=True)
    s = W.make_system(w)
    s.record(w.dt, w.steps, w.checkpoint)
    x, _ = s.state()
    L0, xbar = W.loss_and_seed(w, x)
    gE, _, gv, its = s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter)
    assert (its > 0).all() and (its < w.max_iter).all()

    def loss(young=None, v0=None):
        s.set_young(w.young if young is None else young)
        s.set_state(w.x0, w.v0 if v0 is None else v0)
        s.step(w.dt, 1, w.steps)
        return W.loss_and_seed(w, s.state()[0])[0]

    assert loss() == L0, "a replayed forward must reproduce the recorded run"
    h = 1e-6
    rng = np.random.default_rng(32)
    floorE, floorV = 2e-6 * np.abs(gE).max(), 2e-6 * np.abs(gv).max()
    for e in rng.choice(w.n_tets, 8, replace=False):
        fd = (loss(young=w.young * np.exp(h * (np.arange(w.n_tets) == e)))
              - loss(young=w.young * np.exp(-h * (np.arange(w.n_tets) == e)))) / (2 * h)
        print(f"  log E[{e}] adjoint {gE[e]: .8e} fd {fd: .8e}", flush=True)
        assert abs(gE[e] - fd) <= 1e-4 * abs(fd) + floorE, (e, gE[e], fd)
    free = np.flatnonzero(w.fixed == 0)
    for i in rng.choice(free, 8, replace=False):
        k = int(rng.integers(3))
        vp, vm = w.v0.copy(), w.v0.copy()
        vp[i, k] += h
        vm[i, k] -= h
        fd = (loss(v0=vp) - loss(v0=vm)) / (2 * h)
        print(f"  v0[{i},{k}] adjoint {gv[i, k]: .8e} fd {fd: .8e}", flush=True)
        assert abs(gv[i, k] - fd) <= 1e-4 * abs(fd) + floorV, (i, k, gv[i, k], fd)
    # every parameter at once: directional derivatives along random directions,
    # where the signal is the whole gradient (measured agreement ~4e-8)
    dE = rng.normal(size=w.n_tets)
    dv = rng.normal(size=w.v0.shape) * (w.fixed == 0)[:, None]
    h = 1e-5
    fdE = (loss(young=w.young * np.exp(h * dE)) - loss(young=w.young * np.exp(-h * dE))) / (2 * h)
    fdV = (loss(v0=w.v0 + h * dv) - loss(v0=w.v0 - h * dv)) / (2 * h)
    print(f"  directional: log E {gE @ dE:.10e} vs {fdE:.10e}, v0 {np.sum(gv * dv):.10e} vs {fdV:.10e}", flush=True)
    assert abs(gE @ dE - fdE) <= 1e-6 * abs(fdE)
    assert abs(np.sum(gv * dv) - fdV) <= 1e-6 * abs(fdV)
    s.close()


def test_diff16_gradients_match_the_cpu_oracle_adjoint():
    # the same 0.5 m body at 16^3 cells (24,576 tets), GPU against the FP64 CPU adjoint
    w = W.diff32(n=16, h=1.0 / 32.0, steps=20)
    w.checkpoint = 5
    print("case diff16 vs oracle", flush=True)
    _, gE, gx, gv, its, x = gpu_grad(w, vbar_scale=0.0)
    oE, ox, ov, oits, xo = oracle_grad(w, vbar_scale=0.0)
    assert rel(x, xo) <= 1e-11
    assert rel(gE, oE) <= 1e-8 and rel(gv, ov) <= 1e-8 and rel(gx, ox) <= 1e-8, (rel(gE, oE), rel(gv, ov), rel(gx, ox))
    assert np.abs(its - oits).max() <= 1


@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP32_STORAGE, N.PK_FP32_UPPER, N.PK_FP64_UPPER])
def test_segment_matrix_store_gives_identical_gradients(precision):
    # the reverse steps solve with the matrices the replay kept, or assemble
    # them again (store off): the same matrices, so bit-identical gradients;
    # 7 substeps checkpointed every 3 leave a shorter last segment
    w = W.batch128(bodies=3, lo=3, hi=5, h=0.1, steps=7)
    w.young = w.young / 10.0
    print(f"case matrix store, precision {precision}", flush=True)
    out = []
    for cap, graphs in ((0, False), (1 << 30, False), (1 << 30, True), (1 << 30, False)):
        s = W.make_system(w, precision)
        s.tune(*shape_for(precision, 8, 32, 64, 2))
        s.set_matrix_store(cap)
        s.set_graphs(graphs)
        s.record(w.dt, w.steps, 3)
        x, _ = s.state()
        _, xbar = W.loss_and_seed(w, x)
        out.append(s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter))
        out.append(s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter))  # the kept buffers, reused
        s.close()
    for other in out[1:]:
        for u, v in zip(out[0], other):
            assert np.array_equal(u, v)

