"""Upper-only FP32 storage (PK_FP32_UPPER) against full FP32 storage.

Each lower block is read as the transpose of its upper mirror, with the same
lanes, products and reduction order, so every bar here is bit-identity with
PK_FP32_STORAGE: matrix, SpMV, states, iteration counts, residuals and
gradients. Full FP32 storage is itself held to the FP64 oracle
(tests/test_gpu_step.py, tests/test_gpu_adjoint.py).
"""

import numpy as np
import pytest

from physik_amd import meshgen as mg
from physik_amd import native as N
from physik_amd import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [(4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (8, 4), (16, 1), (16, 2), (16, 4), (32, 1), (32, 2)]


def twisted(n=(8, 5, 5), seed=12):
    nx, ny, nz = n
    h = 0.1
    rest, tets = mg.box_mesh(nx, ny, nz, h, "kuhn")
    rest = mg.jitter(rest, h, 0.1, mg.Pcg32(seed))
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(nx, ny, nz, 0)] = 1
    x0 = mg.twist_about_x(rest, nx, ny, nz, h, 0.8)
    return W._single("upper", rest, tets, 1e5, 0.4, 1000.0, fixed, x0, np.zeros_like(rest), 1.0 / 60.0, 6, 1e-7, 500)


def pair(w, shape=None):
    a = W.make_system(w, N.PK_FP32_STORAGE)
    b = W.make_system(w, N.PK_FP32_UPPER)
    if shape:
        a.tune(*shape)
        b.tune(*shape)
    return a, b


def same_steps(a, b, w, steps):
    for n in range(steps):
        print(f"  step {n}", flush=True)
        a.step(w.dt)
        b.step(w.dt)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"step {n}"
        for u, v in zip(a.results(), b.results()):
            assert np.array_equal(u, v), f"step {n} results"


def test_upper_matrix_and_spmv_equal_full_storage():
    w = twisted()
    print("case matrix and spmv", flush=True)
    a, b = pair(w)
    assert b.nnzb_upper < b.nnzb_full
    a.assemble(w.dt)
    b.assemble(w.dt)
    for u, v in zip(a.matrix(), b.matrix()):
        assert np.array_equal(u, v)
    x = np.random.default_rng(5).normal(size=(a.nd, 3))
    assert np.array_equal(a.spmv(x), b.spmv(x))


def test_upper_single_body_steps_equal_full_storage():
    w = twisted()
    print("case single body", flush=True)
    a, b = pair(w)
    same_steps(a, b, w, w.steps)


@pytest.mark.parametrize("lanes,unroll", SHAPES)
def test_every_upper_launch_shape_equals_full_storage(lanes, unroll):
    w = twisted(n=(6, 4, 4), seed=13)
    print(f"case shape {lanes},{unroll}", flush=True)
    a, b = pair(w, (lanes, 32, 64, unroll))
    same_steps(a, b, w, 2)


def test_upper_batch_steps_equal_full_storage_and_count_stored_blocks():
    w = W.batch128(bodies=10, lo=3, hi=8, h=0.1, steps=4)
    print("case batch", flush=True)
    a, b = pair(w, (8, 32, 64, 2))
    a.profile(True)
    b.profile(True)
    same_steps(a, b, w, w.steps)
    wa, wb = a.spmv_work(), b.spmv_work()
    assert wa == wb
    blocks, upper, rows, bl = wb
    assert 0 < upper < blocks and rows > 0 and bl > 0
    # stored blocks over the lower ones: nd diagonal blocks plus one of each mirrored pair
    assert (upper - rows) * 2 + rows == blocks


def test_upper_graphs_and_adjoint_equal_full_storage():
    w = W.batch128(bodies=3, lo=3, hi=5, h=0.1, steps=5)
    w.young = w.young / 10.0
    print("case graphs and adjoint", flush=True)
    out = []
    for precision in (N.PK_FP32_STORAGE, N.PK_FP32_UPPER):
        s = W.make_system(w, precision)
        s.tune(16, 32, 64, 2)
        s.set_graphs(True)
        s.record(w.dt, w.steps, 2)
        x, _ = s.state()
        _, xbar = W.loss_and_seed(w, x)
        out.append(s.backward(xbar, None, tol=w.adj_tol, max_iter=w.max_iter))
        s.close()
    for u, v in zip(*out):
        assert np.array_equal(u, v)
