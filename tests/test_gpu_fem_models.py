"""The linear FEM model and the NeoHookean setting on the GPU path
(pk_system_set_fem_models, PHYSIK_AMD_SetTetMeshFemModel) against the FP64
oracle and the reference's FP32 fixtures (tests/golden/linear_fp32.npz).

Bars: element blocks and forces vs the FP64 oracle 1e-12 of the case's
scale, vs the FP32 fixture 2e-5; trajectories vs the FP64 oracle 1e-9 per
step (CG iterations +-1), vs the FP32 fixture 1e-4; a batch of a linear and a
corotational body equals the two bodies run alone, bit for bit; the reverse
mode refuses a linear system."""

import os

import numpy as np
import pytest

from checkers import Oracle
from physik_amd import meshgen as mg
from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import System
from physik_amd.world import World, material

pytestmark = pytest.mark.gpu

G = (0.0, 0.0, -9.81)
PAIRS = [(a, b) for a in range(4) for b in range(a, 4)]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def load(golden_dir):
    z = np.load(os.path.join(golden_dir, "linear_fp32.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def test_linear_elements_on_the_gpu(golden_dir):
    g = load(golden_dir)
    n = len(g["elements__young"])
    print(f"case linear elements: {n} tets", flush=True)
    rest = g["elements__rest"].reshape(-1, 3).astype(np.float64)
    tets = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    s = System(rest, tets, g["elements__young"].astype(np.float64), g["elements__poisson"].astype(np.float64),
               g["elements__damping"].astype(np.float64), 1.0, precision=N.PK_FP64)
    s.set_fem_models(0)
    s.set_state(g["elements__x"].reshape(-1, 3).astype(np.float64), g["elements__v"].reshape(-1, 3).astype(np.float64))
    s.assemble(0.01)
    blocks, forces, _ = s.elements()
    R, kind = s.rotations()
    s.close()
    assert (kind == 3).all() and np.array_equal(R, np.broadcast_to(np.eye(3), R.shape))
    for i in range(n):
        o = Oracle(64, g["elements__rest"][i].astype(np.float64), np.array([[0, 1, 2, 3]], np.int32),
                   float(g["elements__young"][i]), float(g["elements__poisson"][i]), float(g["elements__damping"][i]),
                   1.0)
        o.set_fem_model(0)
        o.set_state(g["elements__x"][i].astype(np.float64), g["elements__v"][i].astype(np.float64))
        o.assemble(0.01)
        _, _, fo, Ko = o.elements()
        o.close()
        Kp = np.stack([Ko[0][a, b] for a, b in PAIRS])
        Kr = np.stack([g["elements__K"][i][a, b] for a, b in PAIRS]).astype(np.float64)
        scale = np.abs(Kp).max()
        fscale = max(np.abs(fo[0]).max(), scale * np.abs(g["elements__x"][i] - g["elements__rest"][i]).max())
        assert np.abs(blocks[i] - Kp).max() <= 1e-12 * scale and np.abs(forces[i] - fo[0]).max() <= 1e-12 * fscale
        assert np.abs(blocks[i] - Kr).max() <= 2e-5 * scale
        assert np.abs(forces[i] - g["elements__f"][i]).max() <= 2e-5 * fscale


@pytest.mark.parametrize("name", ["implicit", "explicit"])
def test_linear_trajectories_on_the_gpu(golden_dir, name):
    g = load(golden_dir)
    print(f"case linear trajectory {name}", flush=True)
    rest = g[f"{name}__rest"].astype(np.float64)
    tets = g[f"{name}__tets"]
    s = System(rest, tets, 2e4, 0.35, 0.05, 500.0, precision=N.PK_FP64, node_fixed=g[f"{name}__fixed"])
    s.set_fem_models(0)
    s.set_gravity(G)
    s.set_solver(1e-6, 256)
    s.set_mode(name == "implicit")
    s.set_state(g[f"{name}__x_init"].astype(np.float64), g[f"{name}__v_init"].astype(np.float64))
    o = Oracle(64, rest, tets, 2e4, 0.35, 0.05, 500.0)
    o.set_fem_model(0)
    o.set_params(G, 2, 1e-6, 256, implicit=name == "implicit")
    o.set_fixed(g[f"{name}__fixed"])
    o.set_state(g[f"{name}__x_init"].astype(np.float64), g[f"{name}__v_init"].astype(np.float64))
    dt = float(g[f"{name}__dt"])
    for k in range(len(g[f"{name}__x"])):
        s.step(dt, 2)
        assert o.step(dt) == 0
        x, v = s.state()
        xo, vo = o.state()
        assert rel(x, xo) <= 1e-9 and rel(v, vo) <= 1e-9, (k, rel(x, xo), rel(v, vo))
        assert rel(x, g[f"{name}__x"][k]) <= 1e-4 and rel(v, g[f"{name}__v"][k]) <= 1e-4, k
        if name == "implicit":
            assert abs(int(s.results()[0][0]) - o.cg()[0]) <= 1
    s.close()
    o.close()


def test_mixed_models_batch_equals_bodies_alone_and_reverse_mode_refuses():
    w = W.batch128(bodies=2, lo=3, hi=4, h=0.1, steps=3)
    print("case mixed fem models batch", flush=True)
    models = np.ones(w.n_tets, np.uint8)
    t0, t1 = w.body_tets[0]
    models[t0:t1] = 0
    s = W.make_system(w)
    s.set_fem_models(models)
    s.step(w.dt, 1, w.steps)
    x, v = s.state()
    with pytest.raises(RuntimeError, match="corotational"):
        s.record(w.dt, 1)
    s.close()
    for b in range(2):
        wb = w.body(b)
        sb = W.make_system(wb)
        sb.set_fem_models(0 if b == 0 else 1)
        sb.step(wb.dt, 1, wb.steps)
        xb, vb = sb.state()
        n0, n1 = w.body_nodes[b]
        assert np.array_equal(xb, x[n0:n1]) and np.array_equal(vb, v[n0:n1])
        o = Oracle(64, wb.rest, wb.tets, wb.young, wb.poisson, wb.damping, wb.density)
        o.set_fem_model(0 if b == 0 else 1)
        o.set_params(wb.gravity, 1, wb.tol, wb.max_iter)
        o.set_fixed(wb.fixed)
        for _ in range(wb.steps):
            assert o.step(wb.dt) == 0
        assert rel(xb, o.state()[0]) <= 1e-9
        sb.close()


@pytest.mark.parametrize("model", [0, 2])
def test_fem_model_through_the_physik_api(model):
    rest, tets = mg.box_mesh(2, 2, 3, 0.25, "five")
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(2, 2, 3, 4)] = 1
    print(f"case PHYSIK fem model {model}", flush=True)
    w = World()
    mesh = w.GenerateTetMesh(rest, tets)
    comp = w.CreateTetMeshPhysicsComponent(mesh, material(800.0, 5e3, 0.3, 0.1))
    w.DestroyGeneratedTetMesh(mesh)
    assert w.AMD_SetTetMeshFemModel(comp, model) == 1 and w.AMD_SetTetMeshFemModel(comp, 3) == 0
    w.SetGravity(*G)
    w.SetSolverMode(1)
    w.SetConjugateGradientTolerance(1e-6)
    for i in np.flatnonzero(fixed):
        w.SetNodeFixed(int(i), 1)
    o = Oracle(64, rest.astype(np.float32).astype(np.float64), tets, 5e3, 0.3, 0.1, 800.0)
    o.set_fem_model(model)
    o.set_params(G, 1, float(np.float32(1e-6)), 128)
    o.set_fixed(fixed)
    for k in range(5):
        w.Step(1.0 / 60.0)
        assert o.step(float(np.float32(1.0 / 60.0))) == 0
        x = np.array([w.GetNodePosition(i) for i in range(len(rest))], np.float64)
        xo = o.state()[0].astype(np.float32).astype(np.float64)
        assert rel(x, xo) <= 1e-6, (k, rel(x, xo))
    w.close()
