"""The drop-in PHYSIK_* path on the GPU, driven like the reference's own tests.

Each small case of tests/golden/small_cases_fp32.npz (stepped by the compiled
reference) is rebuilt through PHYSIK_GenerateTetMesh /
PHYSIK_CreateTetMeshPhysicsComponent / PHYSIK_SetNodeFixed / ... and stepped
with PHYSIK_Step. Positions and velocities must match the FP32 reference to
1e-4 and the FP64 oracle (same float inputs) to 1e-9; CG iteration counts
must match the FP64 oracle within +-1.
"""

import os

import numpy as np
import pytest

from checkers import Oracle
from physik_amd.world import World, material

pytestmark = pytest.mark.gpu

CASES = ["single_tet_gravity", "single_tet_stretched_damped", "two_tets_substeps", "two_tets_maxit1",
         "two_tets_one_inactive", "two_tets_deactivate_free_node", "kuhn2_twist"]


def rel(a, b, floor=1e-30):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), floor)


def load_case(golden_dir, name):
    z = np.load(os.path.join(golden_dir, "small_cases_fp32.npz"), allow_pickle=False)
    return {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}


@pytest.mark.parametrize("name", CASES)
def test_small_case_through_physik_api(golden_dir, name):
    c = load_case(golden_dir, name)
    print(f"case {name}", flush=True)
    w = World()
    mesh = w.GenerateTetMesh(c["rest"], c["tets"])
    comp = w.CreateTetMeshPhysicsComponent(mesh, material(float(c["density"]), float(c["young"]),
                                                          float(c["poisson"]), float(c["damping"])))
    w.DestroyGeneratedTetMesh(mesh)
    assert w.IsComponentHandleValid(comp)
    n = len(c["rest"])
    w.SetGravity(*[float(v) for v in c["gravity"]])
    w.SetSubstepCount(int(c["substeps"]))
    w.SetSolverMode(1)
    w.SetConjugateGradientTolerance(float(c["tol"]))
    w.SetConjugateGradientMaxIterations(int(c["max_iter"]))
    for i in range(n):
        if c["fixed"][i]:
            w.SetNodeFixed(i, 1)
    for i in range(n):
        w.SetNodePosition(i, *[float(v) for v in c["x_init"][i]])
        w.SetNodeVelocity(i, *[float(v) for v in c["v_init"][i]])
    for t, a in enumerate(c["active"]):
        if not a:
            w.SetTetActive(comp, t, 0)

    o = Oracle(64, c["rest"].astype(np.float64), c["tets"], float(c["young"]), float(c["poisson"]),
               float(c["damping"]), float(c["density"]))
    o.set_params(tuple(float(v) for v in c["gravity"]), int(c["substeps"]), float(c["tol"]), int(c["max_iter"]))
    o.set_fixed(c["fixed"])
    o.set_state(c["x_init"].astype(np.float64), c["v_init"].astype(np.float64))
    o.set_tet_active(c["active"])

    dt = float(c["dt"])
    for s in range(len(c["x"])):
        w.Step(dt)
        assert o.step(dt) == 0
        x, v = w.node_state(n)
        xo, vo = o.state()
        assert rel(x, xo) <= 1e-9 and rel(v, vo, 1e-12) <= 1e-9, f"step {s}: vs oracle {rel(x, xo):.2e}"
        assert rel(x, c["x"][s]) <= 1e-4 and np.abs(v - c["v"][s]).max() <= 1e-4 * max(1.0, np.abs(c["v"][s]).max())
        it = w.GetLastConjugateGradientIterations()
        assert abs(it - o.cg()[0]) <= 1
        assert bool(w.DidLastConjugateGradientSolveConverge()) == o.cg()[2]
        assert abs(w.GetLastConjugateGradientResidualNorm() - o.cg()[1]) <= 1e-6 * max(1.0, o.cg()[1])
        # the float getters read the same state
        assert np.allclose(w.GetNodePosition(n - 1), x[n - 1], rtol=1e-6, atol=1e-7)


def test_explicit_is_the_default_mode():
    # World.h:133: SolverMode::Explicit until PHYSIK_SetSolverMode(world, 1)
    w = World()
    m = w.GenerateTetMesh(np.eye(4, 3), [[0, 1, 2, 3]])
    w.CreateTetMeshPhysicsComponent(m, material(1.0, 0.0, 0.3, 0.0))
    w.SetGravity(0.0, 0.0, -10.0)
    w.Step(0.1)
    v = w.GetNodeVelocity(0)
    assert np.allclose(v, [0, 0, -1.0]) and w.GetLastConjugateGradientIterations() == 0


def test_missing_mass_raises_like_the_reference():
    w = World()
    w.AddNode(0.0, 0.0, 0.0)
    w.SetSolverMode(1)
    with pytest.raises(RuntimeError, match="no physical mass"):
        w.Step(0.1)


def test_state_changes_between_steps_reach_the_gpu():
    w = World()
    m = w.GenerateTetMesh(np.eye(4, 3), [[0, 1, 2, 3]])
    c = w.CreateTetMeshPhysicsComponent(m, material(24.0, 25.0, 0.3, 0.0))
    w.SetSolverMode(1)
    w.Step(0.05)
    w.SetNodeFixed(0, 1)
    w.SetNodePosition(3, 0.0, 0.0, 1.2)
    w.Step(0.05)
    assert np.array_equal(w.GetNodePosition(0), np.eye(4, 3)[0])
    p = w.GetNodePosition(3)
    assert p[2] < 1.2  # pulled back towards rest
    w.SetTetMeshMaterial(c, material(24.0, 0.0, 0.3, 0.0))
    before = w.GetNodePosition(3)
    w.SetNodeVelocity(3, 0.0, 0.0, 0.0)
    w.Step(0.05)
    assert np.allclose(w.GetNodePosition(3), before)  # no stiffness, no gravity: at rest
