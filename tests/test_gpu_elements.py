"""Element edge cases through the GPU element kernel (k_element's arithmetic,
read back with the inspection records of pk_system_get_elements and
pk_system_get_rotations).

Cases: the 18 single tets of tests/golden/elements_fp32.npz, written by the
compiled reference (12 random stretched and rotated tets, a 1.4 rad twist, an
inverted tet, a nearly flat tet, a rigid rotation, nu = 0.99 (clamped to
0.49) and E < 0 (clamped to 0)), plus a tet whose rest volume is <= 1e-8
(InitializeTetRestData zeroes it, FEMModel.cpp:504-548) and one deformed to
|det F| <= 1e-8 (the Identity fallback of ExtractRotationPolar,
FEMModel.cpp:115-163). All 20 are disjoint tets of ONE system, so the
per-tet Lame parameters, clamps and rest data take the product path
(pk_topology.cpp) too.

Bars:
  GPU FP64 vs the FP64 oracle: R, blocks R K_ab R^T, forces  <= 1e-12 of the
      case's scale (largest block entry; forces: largest block entry times
      the largest node displacement), polar branch equal
  GPU vs the reference's FP32 fixture: blocks and forces     <= 2e-5 of the
      same scale (the FP32 rounding of the reference)
  every branch is taken: Newton limit, column 2 negated, Identity, and a
      tet that contributes nothing (zero volume)
"""

import os

import numpy as np
import pytest

from checkers import Oracle, oracle_polar_vjp
from physik_amd import native as N
from physik_amd.system import System

pytestmark = pytest.mark.gpu

PAIRS = [(a, b) for a in range(4) for b in range(a, 4)]
UNIT = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)


def cases(golden_dir):
    e = np.load(os.path.join(golden_dir, "elements_fp32.npz"))
    out = []
    for i in range(len(e["young"])):
        out.append(dict(rest=e["rest"][i].astype(np.float64), x=e["x"][i].astype(np.float64),
                        v=e["v"][i].astype(np.float64), young=float(e["young"][i]), poisson=float(e["poisson"][i]),
                        damping=float(e["damping"][i]), ref=dict(valid=bool(e["valid"][i]), f=e["f"][i],
                                                                 K=e["K"][i])))
    flat_rest = UNIT.copy()
    flat_rest[3] = [0.3, 0.3, 3e-8]  # V = 5e-9 <= 1e-8: zeroed rest data
    out.append(dict(rest=flat_rest, x=flat_rest + 0.01, v=np.ones((4, 3)), young=1e4, poisson=0.3, damping=0.2,
                    ref=None, fixed=True))
    squashed = UNIT.copy()
    squashed[3] = [0.1, 0.2, 1e-9]  # det F = 1e-9 <= 1e-8: R = Identity
    out.append(dict(rest=UNIT.copy(), x=squashed, v=0.5 * np.ones((4, 3)), young=2e4, poisson=0.35, damping=0.1,
                    ref=None))
    return out


def oracle_element(c):
    tets = np.array([[0, 1, 2, 3]], np.int32)
    o = Oracle(64, c["rest"], tets, c["young"], c["poisson"], c["damping"], 1.0)
    o.set_state(c["x"], c["v"])
    o.assemble(0.01)
    valid, R, f, K = o.elements()
    vol, dminv, _ = o.rest()
    o.close()
    F = (c["x"][1:] - c["x"][0]).T @ dminv[0].T
    kind = oracle_polar_vjp(F, np.zeros((3, 3)))[2] if valid[0] and vol[0] > 0 else -1
    return bool(valid[0] and vol[0] > 0), R[0], f[0], K[0], kind


def test_element_edge_cases_match_oracle_and_reference(golden_dir):
    cs = cases(golden_dir)
    n = len(cs)
    print(f"case element edge cases: {n} disjoint tets, FP64", flush=True)
    rest = np.concatenate([c["rest"] for c in cs])
    x = np.concatenate([c["x"] for c in cs])
    v = np.concatenate([c["v"] for c in cs])
    tets = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    fixed = np.zeros(4 * n, np.uint8)
    for i, c in enumerate(cs):
        if c.get("fixed"):
            fixed[4 * i:4 * i + 4] = 1  # its nodes get no mass: keep them out of the system
    s = System(rest, tets, [c["young"] for c in cs], [c["poisson"] for c in cs], [c["damping"] for c in cs], 1.0,
               precision=N.PK_FP64, node_fixed=fixed)
    s.set_state(x, v)
    s.assemble(0.01)
    blocks, forces, _ = s.elements()
    R, kind = s.rotations()
    s.close()
    seen = set()
    for i, c in enumerate(cs):
        valid, Ro, fo, Ko, ko = oracle_element(c)
        seen.add(int(kind[i]))
        assert kind[i] == ko, (i, kind[i], ko)
        if not valid:
            assert not blocks[i].any() and not forces[i].any() and not R[i].any()
            continue
        Kp = np.stack([Ko[a, b] for a, b in PAIRS])
        scale = max(np.abs(Kp).max(), 1e-300)
        fscale = max(np.abs(fo).max(), scale * np.abs(c["x"] - c["rest"]).max(), 1e-300)
        assert np.abs(R[i] - Ro).max() <= 1e-12, (i, np.abs(R[i] - Ro).max())
        eK = np.abs(blocks[i] - Kp).max() / scale
        ef = np.abs(forces[i] - fo).max() / fscale
        assert eK <= 1e-12 and ef <= 1e-12, (i, eK, ef)
        if c["ref"] is not None:
            assert c["ref"]["valid"]
            Kr = np.stack([c["ref"]["K"][a, b] for a, b in PAIRS]).astype(np.float64)
            rK = np.abs(blocks[i] - Kr).max() / scale
            rf = np.abs(forces[i] - c["ref"]["f"]).max() / fscale
            print(f"  case {i}: kind {kind[i]}, vs oracle K {eK:.1e} f {ef:.1e}, vs reference K {rK:.1e} f {rf:.1e}",
                  flush=True)
            assert rK <= 2e-5 and rf <= 2e-5, (i, rK, rf)
    assert {0, 1, 2, -1} <= seen, seen
    # the clamps: nu = 0.99 acts as 0.49 (finite, nonzero blocks); E < 0 as 0
    assert np.isfinite(blocks[16]).all() and np.abs(blocks[16]).max() > 0
    assert not blocks[17].any() and not forces[17].any()
