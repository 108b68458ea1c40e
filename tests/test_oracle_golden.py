"""The oracle's FP32 instantiation against fixtures from the compiled reference.

Every comparison here is bit-exact: the restatement follows the reference's
operation order and both are built without FMA contraction, so any deviation
(a dropped term, a swapped index, a different summation order) shows up as a
nonzero difference.
"""

import os

import numpy as np
import pytest

from checkers import Oracle


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name), allow_pickle=False)


def make_config1(g, precision=32):
    o = Oracle(precision, g["rest"].astype(np.float64), g["tets"], 1e5, 0.3, 0.0, 1000.0)
    o.set_params((0.0, 0.0, -9.81), 1, 1e-6, 200)
    o.set_fixed(g["fixed"])
    return o


def test_config1_trajectory_bitwise(golden_dir):
    g = load(golden_dir, "config1_fp32.npz")
    o = make_config1(g)
    dt = float(g["dt"])
    for s in range(1, 101):
        assert o.step(dt) == 0
        it, res, conv = o.cg()
        assert it == g["cg_iterations"][s - 1]
        assert np.float32(res) == g["cg_residual"][s - 1]
        assert conv == bool(g["cg_converged"][s - 1])
        if f"x{s}" in g:
            x, v = o.state()
            assert np.array_equal(x.astype(np.float32), g[f"x{s}"]), f"positions differ at step {s}"
            assert np.array_equal(v.astype(np.float32), g[f"v{s}"]), f"velocities differ at step {s}"


def test_config1_system_bitwise(golden_dir):
    g = load(golden_dir, "config1_fp32.npz")
    s = load(golden_dir, "config1_system_step1.npz")
    o = make_config1(g)
    sysm = o.assemble(float(g["dt"]))
    assert sysm["rc"] == 0
    assert np.array_equal(sysm["dyn"], s["dyn"])
    assert np.array_equal(sysm["row_start"], s["row_start"])
    assert np.array_equal(sysm["col"], s["col"])
    assert np.array_equal(o.scatter(), s["scatter"])
    assert np.array_equal(sysm["vals"].astype(np.float32), s["vals"])
    assert np.array_equal(sysm["rhs"].astype(np.float32), s["rhs"])
    assert np.array_equal(sysm["mass"].astype(np.float32), s["mass"])


def test_elements_bitwise(golden_dir):
    e = load(golden_dir, "elements_fp32.npz")
    tets = np.array([[0, 1, 2, 3]], np.int32)
    for i in range(len(e["young"])):
        o = Oracle(32, e["rest"][i].astype(np.float64), tets, float(e["young"][i]), float(e["poisson"][i]),
                   float(e["damping"][i]), 1.0)
        o.set_state(e["x"][i].astype(np.float64), e["v"][i].astype(np.float64))
        o.assemble(0.01)
        valid, R, f, K = o.elements()
        vol, dminv, _ = o.rest()
        assert np.float32(vol[0]) == e["vol"][i]
        assert np.array_equal(dminv[0].astype(np.float32), e["dminv"][i])
        assert valid[0] == bool(e["valid"][i]), f"case {i}"
        if valid[0]:
            assert np.array_equal(f[0].astype(np.float32), e["f"][i]), f"forces differ in case {i}"
            assert np.array_equal(K[0].astype(np.float32), e["K"][i]), f"blocks differ in case {i}"


@pytest.mark.parametrize("name", ["single_tet_gravity", "single_tet_stretched_damped", "two_tets_substeps",
                                  "two_tets_maxit1", "two_tets_one_inactive",
                                  "two_tets_deactivate_free_node", "kuhn2_twist"])
def test_small_cases_bitwise(golden_dir, name):
    z = load(golden_dir, "small_cases_fp32.npz")
    c = {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}
    o = Oracle(32, c["rest"].astype(np.float64), c["tets"], float(c["young"]), float(c["poisson"]),
               float(c["damping"]), float(c["density"]))
    o.set_params(tuple(float(a) for a in c["gravity"]), int(c["substeps"]), float(c["tol"]), int(c["max_iter"]))
    o.set_fixed(c["fixed"])
    o.set_state(c["x_init"].astype(np.float64), c["v_init"].astype(np.float64))
    o.set_tet_active(c["active"])
    for s in range(len(c["x"])):
        assert o.step(float(c["dt"])) == 0
        x, v = o.state()
        it, res, conv = o.cg()
        assert it == c["cg_iterations"][s]
        assert np.float32(res) == c["cg_residual"][s]
        assert conv == bool(c["cg_converged"][s])
        assert np.array_equal(x.astype(np.float32), c["x"][s]), f"{name}: positions differ at step {s}"
        assert np.array_equal(v.astype(np.float32), c["v"][s]), f"{name}: velocities differ at step {s}"


@pytest.mark.parametrize("mode", ["implicit", "explicit"])
def test_sphere_contacts_bitwise(golden_dir, mode):
    # two collision spheres, one moved every frame (CollisionSphereComponent.cpp:112-193,
    # CollisionComponent.cpp:5-15): per-substep centroid contacts turned into
    # point connections; without the spheres the trajectory differs
    z = load(golden_dir, "spheres_fp32.npz")
    runs = []
    for with_spheres in (True, False):
        o = Oracle(32, z["rest"].astype(np.float64), z["tets"], 1e5, 0.3, 0.2, 1000.0)
        o.set_params((0.0, 0.0, -9.81), 2, 1e-6, 200, implicit=mode == "implicit")
        o.set_fixed(z["fixed"])
        xs, contacts = [], []
        for f in range(len(z[f"{mode}_x"])):
            if with_spheres:
                o.set_spheres(z["centers"][f].astype(np.float64), z["radii"][f].astype(np.float64),
                              z["stiffness"][f].astype(np.float64), z["damping"][f].astype(np.float64))
            assert o.step(float(z[f"{mode}_dt"])) == 0
            assert o.point_connection_count() == 0
            contacts.append(o.last_contact_count())
            x, v = o.state()
            xs.append(x.astype(np.float32))
            if with_spheres:
                if mode == "implicit":
                    assert o.cg()[0] == z["implicit_cg_iterations"][f]
                assert np.array_equal(x.astype(np.float32), z[f"{mode}_x"][f]), f"positions differ at frame {f}"
                assert np.array_equal(v.astype(np.float32), z[f"{mode}_v"][f]), f"velocities differ at frame {f}"
        runs.append((xs, contacts))
    assert min(runs[0][1]) >= 10 and max(runs[1][1]) == 0
    assert not np.array_equal(runs[0][0][-1], runs[1][0][-1])


@pytest.mark.parametrize("mode", ["implicit", "explicit"])
def test_point_connections_bitwise(golden_dir, mode):
    # PHYSIK_AddPointConnection before every frame (PointConnection.cpp:54-130,
    # World.cpp:32-73): forces and k w_a w_b I blocks, including a spring over
    # nodes outside the FEM pattern, a fixed node and a repeated node
    z = load(golden_dir, "connections_fp32.npz")
    o = Oracle(32, z["rest"].astype(np.float64), z["tets"], 1e5, 0.3, 0.2, 1000.0)
    o.set_params((0.0, 0.0, -9.81), 2, 1e-6, 200, implicit=mode == "implicit")
    o.set_fixed(z["fixed"])
    for f in range(len(z[f"{mode}_x"])):
        for c in range(z["nodes"].shape[1]):
            o.add_point_connection(z["nodes"][f, c], z["weights"][f, c].astype(np.float64),
                                   z["targets"][f, c].astype(np.float64), float(z["stiffness"][f, c]),
                                   float(z["damping"][f, c]))
        assert o.point_connection_count() == z[f"{mode}_counts"][f]
        assert o.step(float(z[f"{mode}_dt"])) == 0
        assert o.point_connection_count() == 0
        x, v = o.state()
        if mode == "implicit":
            assert o.cg()[0] == z["implicit_cg_iterations"][f]
        assert np.array_equal(x.astype(np.float32), z[f"{mode}_x"][f]), f"positions differ at frame {f}"
        assert np.array_equal(v.astype(np.float32), z[f"{mode}_v"][f]), f"velocities differ at frame {f}"
