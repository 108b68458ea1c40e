"""CG replayed from captured HIP graphs (pk_system_set_graphs) against the same
solves launched kernel by kernel.

The captured chunks hold the same kernels in the same order as the launches
they replace, so every bar here is bit-identity: positions, velocities,
per-body iteration counts, residuals and gradients, step after step. Every
step replays the captured chunks many times, with the host's own copies and
synchronizations in between, and each step is checked, not only the last.
"""

import numpy as np
import pytest

from physik_amd import meshgen as mg
from physik_amd import native as N
from physik_amd import workloads as W

pytestmark = pytest.mark.gpu


def twisted(n=(6, 4, 4), seed=11):
    nx, ny, nz = n
    h = 0.1
    rest, tets = mg.box_mesh(nx, ny, nz, h, "kuhn")
    rest = mg.jitter(rest, h, 0.1, mg.Pcg32(seed))
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(nx, ny, nz, 0)] = 1
    x0 = mg.twist_about_x(rest, nx, ny, nz, h, 0.8)
    return W._single("graphs", rest, tets, 1e5, 0.35, 1000.0, fixed, x0, np.zeros_like(rest), 1.0 / 60.0, 8, 1e-7,
                     400)


def pair(w, precision=N.PK_FP64, shape=None):
    a = W.make_system(w, precision)
    b = W.make_system(w, precision)
    if shape:
        a.tune(*shape)
        b.tune(*shape)
    b.set_graphs(True)
    return a, b


def same_step(a, b, w, steps):
    for n in range(steps):
        print(f"  step {n}", flush=True)
        a.step(w.dt)
        b.step(w.dt)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"step {n}"
        ra, rb = a.results(), b.results()
        for u, v in zip(ra, rb):
            assert np.array_equal(u, v), f"step {n} results"


@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP32_STORAGE])
def test_single_body_graph_steps_equal_launched_steps(precision):
    w = twisted()
    print(f"case single body, precision {precision}", flush=True)
    a, b = pair(w, precision)
    same_step(a, b, w, w.steps)
    # a maximum that is not a multiple of the chunk: replays run past it
    a.set_solver(1e-12, 7)
    b.set_solver(1e-12, 7)
    same_step(a, b, w, 3)
    assert (a.results()[0] == 7).all()


def test_batched_graph_steps_equal_launched_steps():
    w = W.batch256(bodies=12, lo=3, hi=9, steps=6)
    print("case batch of 12 bodies", flush=True)
    a, b = pair(w, shape=(8, 32, 64, 2))
    it = None
    for n in range(w.steps):
        same_step(a, b, w, 1)
        it = a.results()[0]
    assert len(set(it.tolist())) > 1, "bodies should stop at different iterations"


@pytest.mark.parametrize("staging", [0, 1])
def test_graphs_recaptured_after_tune_and_profiling(staging):
    w = W.batch256(bodies=6, lo=3, hi=7, steps=4)
    print(f"case recapture, staging {staging}", flush=True)
    a, b = pair(w, shape=(16, 64, 128, 1))
    a.set_staging(staging)
    b.set_staging(staging)
    same_step(a, b, w, 1)
    a.tune(4, 32, 64, 4)
    b.tune(4, 32, 64, 4)
    same_step(a, b, w, 1)
    a.profile(True)
    b.profile(True)
    same_step(a, b, w, 2)
    # the stamped spans and the work they cover agree launch for launch
    for k in (N.K_SPMV, N.K_UPDATE, N.K_DIRECTION):
        assert a.kernel_time(k)[0] == b.kernel_time(k)[0] > 0
    assert a.spmv_work() == b.spmv_work()
    b.set_graphs(False)
    same_step(a, b, w, 1)


def test_graph_adjoint_equals_launched_adjoint():
    w = W.batch128(bodies=3, lo=3, hi=5, h=0.1, steps=5)
    w.young = w.young / 10.0
    print("case adjoint", flush=True)
    out = []
    for graphs in (False, True):
        s = W.make_system(w)
        s.tune(16, 32, 64, 2)
        s.set_graphs(graphs)
        s.record(w.dt, w.steps, 2)
        x, _ = s.state()
        _, xbar = W.loss_and_seed(w, x)
        out.append(s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter))
        s.close()
    for u, v in zip(*out):
        assert np.array_equal(u, v)


# ---- whole substeps captured (pk_system_set_graphs mode 2) ----

def substep_pair(w, precision=N.PK_FP64, shape=None):
    a = W.make_system(w, precision)
    b = W.make_system(w, precision)
    if shape:
        a.tune(*shape)
        b.tune(*shape)
    b.set_graphs(2)
    return a, b


@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP32_UPPER])
def test_substep_graphs_equal_launched_steps(precision):
    # the first step finds no previous iteration count: its graph holds two
    # chunks and the host finishes the solve; later graphs hold the solve
    w = twisted()
    print(f"case substep graphs, single body, precision {precision}", flush=True)
    a, b = substep_pair(w, precision)
    same_step(a, b, w, w.steps)
    # a maximum that is not a multiple of the chunk: the graph runs past it
    a.set_solver(1e-12, 7)
    b.set_solver(1e-12, 7)
    same_step(a, b, w, 3)
    assert (a.results()[0] == 7).all()
    # settings the graph bakes in: captured again
    a.set_solver(1e-7, 400)
    b.set_solver(1e-7, 400)
    a.set_gravity((0.0, -3.0, -9.81))
    b.set_gravity((0.0, -3.0, -9.81))
    same_step(a, b, w, 2)


def test_substep_graphs_with_a_batch_and_contacts():
    w = W.batch256(bodies=12, lo=3, hi=9, steps=6)
    sph = (np.array([[0.2, 0.2, 0.2]]), np.array([0.25]), np.array([3000.0]), np.array([5.0]))
    print("case substep graphs, batch of 12 bodies with a sphere", flush=True)
    a, b = substep_pair(w, N.PK_FP64_UPPER)
    a.set_spheres(*sph)
    b.set_spheres(*sph)
    for n in range(w.steps):
        same_step(a, b, w, 1)
    it = a.results()[0]
    assert len(set(it.tolist())) > 1, "bodies should stop at different iterations"
    # a shorter time step: fewer iterations, the graph captured again
    w.dt = w.dt / 4
    same_step(a, b, w, 2)


def test_substep_graphs_leave_record_backward_and_profiling_unchanged():
    w = W.batch128(bodies=3, lo=3, hi=5, h=0.1, steps=5)
    w.young = w.young / 10.0
    print("case substep graphs: record, backward, profiled step", flush=True)
    out = []
    for mode in (0, 2):
        s = W.make_system(w)
        s.set_graphs(mode)
        s.step(w.dt)          # a graph-captured step in mode 2
        s.profile(True)
        s.step(w.dt)          # profiled: launched kernels
        s.profile(False)
        s.record(w.dt, w.steps, 2)
        x, _ = s.state()
        _, xbar = W.loss_and_seed(w, x)
        out.append((x,) + tuple(s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter)))
        assert s.lib.pk_system_set_graphs(s.h, 3) == N.PK_ERR_ARG
        s.close()
    for u, v in zip(*out):
        assert np.array_equal(u, v)
