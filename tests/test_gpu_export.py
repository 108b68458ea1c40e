"""Frame export: the float readback after a step.

The reference keeps its node state in float on the host and its per-frame
readers take it from there (World::PostUpdate, World.cpp:539-549;
PHYSIK_GetNodePosition/Velocity, PhysiKAPI.cpp:1510-1545). Here the state is
FP64 on the device, and a frame export converts it to float triples on the
device and copies them into pinned host memory. Every value must be exactly
static_cast<float> of the FP64 state (numpy's float64 -> float32 cast rounds
to nearest even, as the host cast and the device conversion do), for every
node range, and the export must never be served once the state has changed.
"""

import numpy as np
import pytest

from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import System
from physik_amd.world import World, material

pytestmark = pytest.mark.gpu


def test_export_is_the_float_cast_of_the_state():
    w = W.beam10(n=4)
    s = W.make_system(w)
    for step in range(3):
        s.step(w.dt)
        s.export_begin()
        x, v = s.state()
        fx, fv = s.export_read()
        print(f"step {step}: nodes {s.n}", flush=True)
        assert fx.dtype == np.float32 and fx.shape == (s.n, 3)
        assert np.array_equal(fx, x.astype(np.float32))
        assert np.array_equal(fv, v.astype(np.float32))
        # node ranges, including one that ends at the last node
        for first, count in ((0, 1), (7, 13), (s.n - 5, 5), (s.n, 0)):
            px, pv = s.export_read(first, count)
            assert np.array_equal(px, fx[first:first + count])
            assert np.array_equal(pv, fv[first:first + count])
    s.close()


def test_export_kinds_and_stale_state_are_refused():
    w = W.beam10(n=3)
    s = W.make_system(w)
    s.step(w.dt)
    assert s.export_status() == N.PK_ERR_STATE  # nothing exported yet
    s.export_begin(positions=True, velocities=False)
    x, v = s.state()
    px, _ = s.export_read(velocities=False)
    assert np.array_equal(px, x.astype(np.float32))
    assert s.export_status(velocities=True, positions=False) == N.PK_ERR_STATE
    s.export_begin(positions=False, velocities=True)
    _, pv = s.export_read(positions=False)
    assert np.array_equal(pv, v.astype(np.float32))
    assert s.export_status(positions=True, velocities=False) == N.PK_ERR_STATE
    # any change of the state makes the export stale
    s.export_begin()
    assert s.export_status() == N.PK_OK
    s.step(w.dt)
    assert s.export_status() == N.PK_ERR_STATE
    s.export_begin()
    s.set_state(w.x0, w.v0)
    assert s.export_status() == N.PK_ERR_STATE
    # bad arguments
    assert s.export_status(first=-1, count=1) == N.PK_ERR_ARG
    assert s.export_status(first=s.n - 1, count=2) == N.PK_ERR_ARG
    assert s.lib.pk_system_export_begin(s.h, 0) == N.PK_ERR_ARG
    assert s.lib.pk_system_export_begin(s.h, 4) == N.PK_ERR_ARG
    s.close()


def test_export_of_many_nodes_covers_every_value():
    # 900,001 nodes (an odd 3N, and more values than one sweep of the capped
    # grid converts), one tet: the export needs no step
    n = 900_001
    rng = np.random.default_rng(5)
    rest = np.zeros((n, 3))
    rest[1:4] = np.eye(3)
    rest[4:] = rng.normal(size=(n - 4, 3))
    tets = np.array([[0, 1, 2, 3]], np.int32)
    s = System(rest, tets, 1e5, 0.3, 0.0, 1000.0)
    x = rng.normal(scale=10.0, size=(n, 3))
    v = rng.normal(scale=1e-3, size=(n, 3))
    x[5] = [np.inf, -np.inf, np.nan]
    v[6] = [1e300, -1e-300, 3.4028235677973366e38]  # overflow, underflow, the float max rounding boundary
    s.set_state(x, v)
    s.export_begin()
    fx, fv = s.export_read()
    with np.errstate(over="ignore"):
        ex, ev = x.astype(np.float32), v.astype(np.float32)
    assert np.array_equal(fx, ex, equal_nan=True)
    assert np.array_equal(fv, ev, equal_nan=True)
    s.close()


def _physik_beam(n=3, mode=1):
    from physik_amd import meshgen as mg
    rest, tets = mg.box_mesh(n, n, n, 0.1, "five")
    w = World()
    mesh = w.GenerateTetMesh(rest.astype(np.float32), tets)
    comp = w.CreateTetMeshPhysicsComponent(mesh, material(1000.0, 1e5, 0.3, 0.0))
    w.DestroyGeneratedTetMesh(mesh)
    for i in mg.face_nodes(n, n, n, 0):
        w.SetNodeFixed(int(i), 1)
    w.SetSolverMode(mode)
    w.SetConjugateGradientTolerance(1e-6)
    w.SetConjugateGradientMaxIterations(200)
    return w, comp, len(rest)


def _per_node(w, n):
    px = np.array([w.GetNodePosition(i) for i in range(n)], dtype=np.float32)
    pv = np.array([w.GetNodeVelocity(i) for i in range(n)], dtype=np.float32)
    return px, pv


@pytest.mark.parametrize("mask", [3, 1, 2, 0])
def test_physik_getters_read_the_frame_export(mask):
    w, comp, n = _physik_beam()
    assert w.SetFrameExport(mask) == 1
    for step in range(3):
        w.Step(1.0 / 60.0)
        print(f"mask {mask} step {step}", flush=True)
        px, pv = _per_node(w, n)
        bx, bv = w.GetNodePositions(0, n), w.GetNodeVelocities(0, n)
        x, v = w.node_state(n)
        assert np.array_equal(px, x.astype(np.float32))
        assert np.array_equal(pv, v.astype(np.float32))
        assert np.array_equal(bx, px) and np.array_equal(bv, pv)
    # after a host edit the getters read the host state
    w.Step(1.0 / 60.0)
    w.SetNodePosition(5, 1.0, 2.0, 3.0)
    assert np.array_equal(w.GetNodePosition(5), np.float32([1.0, 2.0, 3.0]))
    assert np.array_equal(w.GetNodePositions(5, 1)[0], np.float32([1.0, 2.0, 3.0]))
    assert np.array_equal(w.GetNodeVelocities(5, 1)[0], np.float32([0.0, 0.0, 0.0]))
    w.Step(1.0 / 60.0)
    px, pv = _per_node(w, n)
    x, v = w.node_state(n)
    assert np.array_equal(px, x.astype(np.float32)) and np.array_equal(pv, v.astype(np.float32))
    w.close()


def test_physik_bulk_getters_clip_and_refuse_bad_ranges():
    w, comp, n = _physik_beam()
    w.Step(1.0 / 60.0)
    assert len(w.GetNodePositions(n - 3, 10)) == 3
    assert len(w.GetNodePositions(n, 1)) == 0
    assert len(w.GetNodePositions(-1, 2)) == 0
    assert len(w.GetNodeVelocities(0, 0)) == 0
    assert w.SetFrameExport(4) == 0 and w.SetFrameExport(-1) == 0
    px, _ = _per_node(w, n)
    assert np.array_equal(w.GetNodePositions(n - 3, 10), px[n - 3:])
    w.close()


def test_physik_frame_export_after_explicit_steps():
    # the explicit integrator (the reference's default mode) exports the same way
    w, comp, n = _physik_beam(mode=0)
    w.SetGravity(0.0, 0.0, -9.81)
    w.SetSubstepCount(8)
    for step in range(3):
        w.Step(1.0 / 600.0)
        px, pv = _per_node(w, n)
        bx, bv = w.GetNodePositions(0, n), w.GetNodeVelocities(0, n)
        x, v = w.node_state(n)
        print(f"explicit step {step}: max |v| {np.abs(v).max():.3e}", flush=True)
        assert np.abs(v).max() > 0.0  # gravity moved the free nodes
        assert np.array_equal(px, x.astype(np.float32)) and np.array_equal(pv, v.astype(np.float32))
        assert np.array_equal(bx, px) and np.array_equal(bv, pv)
    w.close()
