"""Collision spheres (CollisionSphereComponent) on the GPU path against the
FP64 oracle, which is itself bit-exact against the compiled reference on the
committed fixture (tests/test_oracle_golden.py::test_sphere_contacts_bitwise).

Bars: FP64 1e-9 relative on x and v per frame with CG iterations within +-1;
upper storage bit-identical to full; FP32 storage 1e-4; a batch bit-identical
to its bodies run alone.
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


def beam(substeps=2, steps=10):
    n = (4, 3, 3)
    rest, tets = mg.box_mesh(*n, 0.1, "five")
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(*n, 0)] = 1
    w = W._single("spheres", rest, tets, 1e5, 0.3, 1000.0, fixed, rest.copy(), np.zeros_like(rest), 1.0 / 60.0,
                  steps, 1e-9, 400)
    w.damping = np.full(len(tets), 0.2)
    w.substeps = substeps
    return w


def scene(f):
    """Per frame: a sphere under the beam's free end, into which the beam
    sags, and one moved in from the side (centres, radii, stiffness, damping)."""
    return (np.array([[0.32, 0.15, -0.12], [0.22, 0.42 - 0.015 * f, 0.15]]), np.array([0.2, 0.13]),
            np.array([1500.0, 2500.0]), np.array([4.0, 0.0]))


def connections(w):
    t0 = w.tets[10]
    return (np.array([list(t0)]), np.array([[0.1, 0.2, 0.3, 0.4]]), np.array([w.rest[t0].mean(0) + [0.0, 0.02, 0.05]]),
            np.array([500.0]), np.array([2.0]))


def run(w, precision, implicit=True, with_connections=False, shape=None):
    s = W.make_system(w, precision)
    if shape:
        s.tune(*shape)
    s.set_mode(implicit)
    o = Oracle(64, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
    o.set_params(w.gravity, w.substeps, w.tol, w.max_iter, implicit=implicit)
    o.set_fixed(w.fixed)
    o.set_state(w.x0, w.v0)
    out = []
    for f in range(w.steps):
        c, r, k, d = scene(f)
        s.set_spheres(c, r, k, d)
        o.set_spheres(c, r, k, d)
        if with_connections:
            cn = connections(w)
            s.set_connections(*cn)
            o.add_point_connection(cn[0][0], cn[1][0], cn[2][0], cn[3][0], cn[4][0])
        s.step(w.dt, w.substeps, 1)
        assert o.step(w.dt) == 0
        assert o.last_contact_count() > 0
        out.append((s.state(), o.state(), s.results()[0][0], o.cg()[0]))
    s.close()
    return out


@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP64_UPPER])
def test_spheres_match_the_oracle(precision):
    w = beam()
    print(f"case spheres, precision {precision}", flush=True)
    for f, ((xg, vg), (xo, vo), ig, io) in enumerate(run(w, precision)):
        assert rel(xg, xo) <= 1e-9 and rel(vg, vo) <= 1e-9, (f, rel(xg, xo), rel(vg, vo))
        assert abs(int(ig) - io) <= 1


def test_explicit_step_with_spheres_matches_the_oracle():
    w = beam()
    w.dt = 1.0 / 2000.0
    print("case explicit spheres", flush=True)
    for f, ((xg, vg), (xo, vo), _, _) in enumerate(run(w, N.PK_FP64, implicit=False)):
        assert rel(xg, xo) <= 1e-12 and rel(vg, vo) <= 1e-10, (f, rel(xg, xo), rel(vg, vo))


def test_spheres_with_connections_match_the_oracle():
    w = beam()
    print("case spheres and connections", flush=True)
    for f, ((xg, vg), (xo, vo), ig, io) in enumerate(run(w, N.PK_FP64, with_connections=True)):
        assert rel(xg, xo) <= 1e-9 and rel(vg, vo) <= 1e-9, (f, rel(xg, xo), rel(vg, vo))
        assert abs(int(ig) - io) <= 1


def test_upper_and_fp32_storage_with_spheres():
    w = beam()
    print("case spheres, upper and fp32 storage", flush=True)
    full = run(w, N.PK_FP64, shape=PAIR_SHAPE)
    upper = run(w, N.PK_FP64_UPPER, shape=PAIR_SHAPE)
    for a, b in zip(full, upper):
        assert np.array_equal(a[0][0], b[0][0]) and np.array_equal(a[0][1], b[0][1]) and a[2] == b[2]
    for f, ((xg, vg), (xo, vo), _, _) in enumerate(run(w, N.PK_FP32_UPPER)):
        assert rel(xg, xo) <= 1e-4 and rel(vg, vo) <= 1e-4, (f, rel(xg, xo), rel(vg, vo))


def test_removing_spheres_restores_the_plain_step():
    w = beam()
    print("case remove spheres", flush=True)
    a = W.make_system(w)
    b = W.make_system(w)
    a.set_spheres(*scene(0))
    a.set_spheres()
    a.step(w.dt, 2, 1)
    b.step(w.dt, 2, 1)
    assert np.array_equal(a.state()[0], b.state()[0]) and np.array_equal(a.state()[1], b.state()[1])
    with pytest.raises(RuntimeError, match="finite"):
        a.set_spheres([[0.0, 0.0, 0.0]], [-1.0])


def test_batched_bodies_with_spheres_equal_single_body_runs():
    # the bodies of a batch all sit at the origin, so one sphere touches each
    w = W.batch256(bodies=5, lo=2, hi=5, steps=4)
    print("case batch with spheres", flush=True)
    sph = (np.array([[0.2, 0.2, 0.2]]), np.array([0.25]), np.array([3000.0]), np.array([5.0]))
    sb = W.make_system(w)
    sb.set_spheres(*sph)
    sb.step(w.dt, 1, w.steps)
    xb, vb = sb.state()
    itb = sb.results()[0]
    for bi in range(w.n_bodies):
        wb = w.body(bi)
        s1 = W.make_system(wb)
        s1.set_spheres(*sph)
        s1.step(wb.dt, 1, wb.steps)
        x1, v1 = s1.state()
        n0, n1 = w.body_nodes[bi]
        assert np.array_equal(xb[n0:n1], x1) and np.array_equal(vb[n0:n1], v1), f"body {bi}"
        assert s1.results()[0][0] == itb[bi]
        o = Oracle(64, wb.rest, wb.tets, wb.young, wb.poisson, wb.damping, wb.density)
        o.set_params(wb.gravity, 1, wb.tol, wb.max_iter)
        o.set_fixed(wb.fixed)
        o.set_state(wb.x0, wb.v0)
        o.set_spheres(*sph)
        for _ in range(wb.steps):
            assert o.step(wb.dt) == 0
        xo, vo = o.state()
        assert rel(x1, xo) <= 1e-9 and rel(v1, vo) <= 1e-9, (bi, rel(x1, xo), rel(v1, vo))


def single_tet(w, young, damping, density):
    unit = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    m = w.GenerateTetMesh(unit, [[0, 1, 2, 3]])
    c = w.CreateTetMeshPhysicsComponent(m, material(density, young, 0.3, damping))
    w.DestroyGeneratedTetMesh(m)
    return [w.GetTetMeshGlobalNodeIndex(c, i) for i in range(4)]


def centroid(w, nodes):
    return np.mean([w.GetNodePosition(n) for n in nodes], axis=0)


def test_physik_sphere_contact_moves_the_tet():
    # PointConnectionTests.cpp:792-815 and :859-877 (default explicit mode,
    # no gravity): the sphere pushes the tet's centroid up; with zero
    # stiffness and damping nothing moves; no connection outlives the step
    print("case PHYSIK sphere contact", flush=True)
    for settings, moves in (((1000.0, 10.0), True), ((0.0, 0.0), False)):
        w = World()
        nodes = single_tet(w, 25.0, 0.25, 24.0)
        s = w.CreateCollisionSphereComponent(0.25, 0.25, 0.25, 0.75)
        w.SetCollisionSphereConnectionSettings(s, *settings)
        assert w.GetPointConnectionCount() == 0
        before = centroid(w, nodes)
        w.Step(0.1)
        after = centroid(w, nodes)
        assert w.GetPointConnectionCount() == 0
        if moves:
            assert after[2] > before[2]
        else:
            assert np.sum((after - before) ** 2) < 1e-6


def test_physik_spheres_with_kinematic_targets_match_the_oracle():
    # through the PHYSIK_* boundary in FP64: two spheres, one moved by a
    # kinematic target before every frame, a third destroyed after frame 3,
    # against the oracle fed the same float32 values
    wl = beam()
    print("case PHYSIK spheres", flush=True)
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
    c, r, k, d = scene(0)
    handles = []
    for i in range(2):
        h = w.CreateCollisionSphereComponent(*[float(v) for v in c[i]], float(r[i]))
        w.SetCollisionSphereConnectionSettings(h, float(k[i]), float(d[i]))
        handles.append(h)
    extra = w.CreateCollisionSphereComponent(0.05, 0.15, 0.3, 0.08)
    o = Oracle(64, wl.rest, wl.tets, 1e5, float(f32(0.3)), float(f32(0.2)), 1000.0)
    o.set_params(tuple(f32(wl.gravity)), wl.substeps, float(f32(1e-6)), 400)
    o.set_fixed(wl.fixed)
    n = len(wl.rest)
    for f in range(wl.steps):
        c, r, k, d = scene(f)
        w.SetCollisionComponentKinematicTarget(handles[1], *[float(v) for v in c[1]])
        if f == 4:
            w.DestroyComponent(extra)
        cs, rs, ks, ds = f32(c), f32(r), f32(k), f32(d)
        if f < 4:
            cs = np.vstack([cs, f32([[0.05, 0.15, 0.3]])])
            rs, ks, ds = np.append(rs, f32(0.08)), np.append(ks, 1000.0), np.append(ds, 10.0)
        o.set_spheres(cs, rs, ks, ds)
        w.Step(float(f32(wl.dt)))
        assert o.step(float(f32(wl.dt))) == 0
        assert o.last_contact_count() > 0
        x, v = w.node_state(base + n)
        xo, vo = o.state()
        assert rel(x[base:], xo) <= 1e-9 and rel(v[base:], vo) <= 1e-9, (f, rel(x[base:], xo), rel(v[base:], vo))
        assert w.GetPointConnectionCount() == 0
