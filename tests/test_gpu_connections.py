"""Point connections (PHYSIK_AddPointConnection) on the GPU path against the
FP64 oracle, which is itself bit-exact against the compiled reference on the
committed fixture (tests/test_oracle_golden.py::test_point_connections_bitwise).

Bars: FP64 1e-9 relative on x and v per frame with CG iterations within +-1;
upper storage bit-identical to full; FP32 storage 1e-4.
"""

import numpy as np
import pytest

from checkers import Oracle
from launch_shapes import PAIR_SHAPE
from physik_amd import meshgen as mg
from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.world import World, material

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def beam(n=(4, 3, 3), substeps=2):
    rest, tets = mg.box_mesh(*n, 0.1, "five")
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(*n, 0)] = 1
    w = W._single("conn", rest, tets, 1e5, 0.3, 1000.0, fixed, rest.copy(), np.zeros_like(rest), 1.0 / 60.0, 8,
                  1e-9, 400)
    w.damping = np.full(len(tets), 0.2)
    w.substeps = substeps
    return w


def frames(w, far=True):
    """Per-frame connections: a point inside a tet pulled toward a moving
    target, two nodes of another tet with a zero weight, and (far=True) a
    spring over nodes no tet joins, one of them fixed, with a repeated node."""
    t0, t1 = w.tets[10], w.tets[30]
    out = []
    for f in range(w.steps):
        s = 0.01 * f
        c = [([*t0], [0.1, 0.2, 0.3, 0.4], w.rest[t0].mean(0) + [0.0, s, 0.05], 500.0, 2.0),
             ([t1[0], t1[1], 3, 0], [0.5, 0.5, 0.0, 0.0], w.rest[t1[0]] + [0.02, 0.0, -s], 800.0, 0.0)]
        if far:
            c.append(([5, len(w.rest) - 1, 7, 7], [0.25, 0.25, 0.3, 0.2], [0.2 + s, 0.15, 0.1], 300.0, 1.0))
        out.append(c)
    return out


def far_pairs(w):
    n = [5, len(w.rest) - 1, 7]
    return [(a, b) for i, a in enumerate(n) for b in n[i + 1:]]


def run(w, precision, conn, implicit=True, extra_pairs=None, shape=None):
    s = W.make_system(w, precision) if extra_pairs is None else make_with_pairs(w, precision, extra_pairs)
    if shape:
        s.tune(*shape)
    s.set_mode(implicit)
    o = Oracle(64, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
    o.set_params(w.gravity, w.substeps, w.tol, w.max_iter, implicit=implicit)
    o.set_fixed(w.fixed)
    o.set_state(w.x0, w.v0)
    out = []
    for f, cs in enumerate(conn):
        arrays = [np.array([c[i] for c in cs]) for i in range(5)]
        s.set_connections(*arrays)
        for c in cs:
            o.add_point_connection(np.array(c[0]), np.array(c[1], float), np.array(c[2], float), c[3], c[4])
        s.step(w.dt, w.substeps, 1)
        assert o.step(w.dt) == 0
        out.append((s.state(), o.state(), s.results()[0][0], o.cg()[0]))
    s.close()
    return out


def make_with_pairs(w, precision, pairs):
    from physik_amd.system import System
    s = System(w.rest, w.tets, w.young, w.poisson, w.damping, w.density, precision=precision, node_fixed=w.fixed,
               extra_pairs=pairs)
    s.set_state(w.x0, w.v0)
    s.set_gravity(w.gravity)
    s.set_solver(w.tol, w.max_iter)
    return s


@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP64_UPPER])
def test_connections_inside_the_pattern_match_the_oracle(precision):
    w = beam()
    print(f"case connections in pattern, precision {precision}", flush=True)
    for f, ((xg, vg), (xo, vo), ig, io) in enumerate(run(w, precision, frames(w, far=False))):
        assert rel(xg, xo) <= 1e-9 and rel(vg, vo) <= 1e-9, (f, rel(xg, xo), rel(vg, vo))
        assert abs(int(ig) - io) <= 1


def test_connection_outside_the_pattern_is_rejected_then_accepted_with_extra_pairs():
    w = beam()
    print("case connection outside the pattern", flush=True)
    s = W.make_system(w)
    c = frames(w)[0]
    with pytest.raises(RuntimeError, match="not coupled in the BSR pattern"):
        s.set_connections(*[np.array([x[i] for x in c]) for i in range(5)])
    for f, ((xg, vg), (xo, vo), ig, io) in enumerate(run(w, N.PK_FP64, frames(w), extra_pairs=far_pairs(w))):
        assert rel(xg, xo) <= 1e-9 and rel(vg, vo) <= 1e-9, (f, rel(xg, xo), rel(vg, vo))
        assert abs(int(ig) - io) <= 1


def test_upper_and_fp32_storage_with_connections():
    w = beam()
    print("case connections, upper and fp32 storage", flush=True)
    full = run(w, N.PK_FP64, frames(w), extra_pairs=far_pairs(w), shape=PAIR_SHAPE)
    upper = run(w, N.PK_FP64_UPPER, frames(w), extra_pairs=far_pairs(w), shape=PAIR_SHAPE)
    for a, b in zip(full, upper):
        assert np.array_equal(a[0][0], b[0][0]) and np.array_equal(a[0][1], b[0][1]) and a[2] == b[2]
    for f, ((xg, vg), (xo, vo), _, _) in enumerate(run(w, N.PK_FP32_UPPER, frames(w), extra_pairs=far_pairs(w))):
        assert rel(xg, xo) <= 1e-4 and rel(vg, vo) <= 1e-4, (f, rel(xg, xo), rel(vg, vo))


def test_explicit_step_with_connections_matches_the_oracle():
    w = beam()
    w.dt = 1.0 / 2000.0
    print("case explicit connections", flush=True)
    for f, ((xg, vg), (xo, vo), _, _) in enumerate(run(w, N.PK_FP64, frames(w), implicit=False,
                                                       extra_pairs=far_pairs(w))):
        assert rel(xg, xo) <= 1e-12 and rel(vg, vo) <= 1e-10, (f, rel(xg, xo), rel(vg, vo))


def test_clearing_connections_restores_the_plain_step():
    w = beam()
    print("case clear connections", flush=True)
    a = W.make_system(w)
    b = W.make_system(w)
    c = frames(w, far=False)[0]
    a.set_connections(*[np.array([x[i] for x in c]) for i in range(5)])
    a.set_connections()
    a.step(w.dt, 2, 1)
    b.step(w.dt, 2, 1)
    assert np.array_equal(a.state()[0], b.state()[0]) and np.array_equal(a.state()[1], b.state()[1])


def test_physik_point_connection_pulls_the_point_toward_the_target():
    # PointConnectionTests.cpp:680-740: one tet, a spring from its centroid to
    # a target above it; after a step the point is nearer and higher
    w = World()
    w.SetSolverMode(1)
    unit = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    m = w.GenerateTetMesh(unit, [[0, 1, 2, 3]])
    c = w.CreateTetMeshPhysicsComponent(m, material(24.0, 25.0, 0.3, 0.0))
    w.DestroyGeneratedTetMesh(m)
    nodes = [w.GetTetMeshGlobalNodeIndex(c, i) for i in range(4)]
    target = np.array([0.25, 0.25, 1.25], np.float32)
    before = np.mean([w.GetNodePosition(n) for n in nodes], axis=0)
    w.AddPointConnection(nodes, [0.25] * 4, target, 100.0, 0.0)
    assert w.GetPointConnectionCount() == 1
    w.Step(0.1)
    assert w.GetPointConnectionCount() == 0
    after = np.mean([w.GetNodePosition(n) for n in nodes], axis=0)
    assert np.sum((after - target) ** 2) < np.sum((before - target) ** 2)
    assert after[2] > before[2]


def test_physik_connections_over_far_nodes_match_the_oracle():
    # through the PHYSIK_* boundary: connections added before every frame,
    # including a spring over nodes no tet joins (the world rebuilds its
    # pattern with that pair), in FP64 against the oracle fed the same
    # float32 values the PHYSIK_* calls carry
    wl = beam()
    print("case PHYSIK connections", flush=True)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    w = World()
    w.SetSolverMode(1)
    w.SetSubstepCount(wl.substeps)
    w.SetGravity(*wl.gravity)
    w.SetConjugateGradientTolerance(1e-6)
    w.SetConjugateGradientMaxIterations(400)
    m = w.GenerateTetMeshF64(wl.rest, wl.tets)
    comp = w.CreateTetMeshPhysicsComponent(m, material(1000.0, 1e5, 0.3, 0.2))
    w.DestroyGeneratedTetMesh(m)
    base = w.GetTetMeshGlobalNodeIndex(comp, 0)
    for i in np.flatnonzero(wl.fixed):
        w.SetNodeFixed(base + int(i), 1)
    o = Oracle(64, wl.rest, wl.tets, 1e5, float(f32(0.3)), float(f32(0.2)), 1000.0)
    o.set_params(tuple(f32(wl.gravity)), wl.substeps, float(f32(1e-6)), 400)
    o.set_fixed(wl.fixed)
    n = len(wl.rest)
    for f, cs in enumerate(frames(wl)):
        for nodes, wt, tgt, k, c in cs:
            w.AddPointConnection([base + int(q) for q in nodes], wt, tgt, k, c)
            o.add_point_connection(np.array(nodes), f32(wt), f32(tgt), float(f32(k)), float(f32(c)))
        assert w.GetPointConnectionCount() == 3
        w.Step(float(f32(wl.dt)))
        assert o.step(float(f32(wl.dt))) == 0
        x, v = w.node_state(base + n)
        xo, vo = o.state()
        assert rel(x[base:], xo) <= 1e-9 and rel(v[base:], vo) <= 1e-9, (f, rel(x[base:], xo), rel(v[base:], vo))
