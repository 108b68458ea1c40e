"""Host-side dynamic numbering, BSR patterns and element->slot map (no GPU).

Bit-exact against the reference's SolverData matrix and FindBlockIndex
(fixture from the compiled reference) and against the oracle on assorted
meshes; closed-form counts for the benchmark meshes.
"""

import os

import numpy as np
import pytest

from checkers import Oracle
from physik_amd import meshgen as mg
from physik_amd import workloads as W
from physik_amd.system import Pattern


def test_config1_pattern_matches_reference_fixture(golden_dir):
    g = np.load(os.path.join(golden_dir, "config1_fp32.npz"))
    s = np.load(os.path.join(golden_dir, "config1_system_step1.npz"))
    p = Pattern(g["rest"], g["tets"], 1e5, 0.3, 0.0, 1000.0, node_fixed=g["fixed"])
    assert np.array_equal(p.dyn, s["dyn"])
    assert np.array_equal(p.upper_start, s["row_start"])
    assert np.array_equal(p.upper_col, s["col"])
    assert np.array_equal(p.scatter, s["scatter"])


def full_from_upper(p):
    rows = [[] for _ in range(p.nd)]
    for r in range(p.nd):
        for k in range(p.upper_start[r], p.upper_start[r + 1]):
            c = p.upper_col[k]
            rows[r].append(c)
            if c != r:
                rows[c].append(r)
    return [sorted(x) for x in rows]


def check_against_oracle(rest, tets, fixed=None, tet_active=None):
    o = Oracle(64, rest, tets, 1e4, 0.3, 0.0, 10.0)
    if fixed is not None:
        o.set_fixed(fixed)
    if tet_active is not None:
        o.set_tet_active(tet_active)
    s = o.assemble(0.01)
    flags = {}
    if fixed is not None:
        flags["node_fixed"] = fixed
    if tet_active is not None:
        flags["tet_active"] = tet_active
        _, v = o.state()
        # the oracle deactivates unreferenced nodes like PHYSIK_SetTetActive does
        act = np.zeros(len(rest), np.uint8)
        for t, a in zip(tets, tet_active):
            if a:
                act[t] = 1
        flags["node_active"] = np.where(np.isin(np.arange(len(rest)), np.asarray(tets)[np.asarray(tet_active) == 0]),
                                        act, 1).astype(np.uint8)
    p = Pattern(rest, tets, **flags)
    assert np.array_equal(p.dyn, s["dyn"])
    assert np.array_equal(p.upper_start, s["row_start"])
    assert np.array_equal(p.upper_col, s["col"])
    assert np.array_equal(p.scatter, o.scatter())
    full = full_from_upper(p)
    assert np.array_equal(p.full_start, np.concatenate([[0], np.cumsum([len(x) for x in full])]))
    assert np.array_equal(p.full_col, np.concatenate(full) if full else [])
    return p


@pytest.mark.parametrize("split", ["five", "kuhn"])
def test_grid_patterns_match_oracle(split):
    rest, tets = mg.box_mesh(3, 2, 4, 0.1, split)
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(3, 2, 4, 0)] = 1
    check_against_oracle(rest, tets, fixed)


def test_shuffled_tets_and_nodes_match_oracle():
    rest, tets = mg.box_mesh(3, 3, 2, 0.2, "kuhn")
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(rest))
    inv = np.argsort(perm)
    rest2 = rest[perm]
    tets2 = inv[tets][rng.permutation(len(tets))]
    tets2 = np.array([rng.permutation(t) for t in tets2], np.int32)
    fixed = (rng.random(len(rest)) < 0.15).astype(np.uint8)
    check_against_oracle(rest2, tets2, fixed)


def test_inactive_tets_keep_the_pattern_superset():
    rest, tets = mg.box_mesh(2, 2, 2, 0.1, "five")
    active = np.ones(len(tets), np.uint8)
    active[[0, 7, 13]] = 0
    check_against_oracle(rest, tets, tet_active=active)


def test_single_and_two_tet_counts():
    # reference tests: PointConnectionTests.cpp:2072-2113
    p = Pattern(np.eye(4, 3), [[0, 1, 2, 3]])
    assert p.nnzb_upper == 10 and list(np.diff(p.upper_start)) == [4, 3, 2, 1] and p.nnzb_full == 16
    two = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1]], float)
    p = Pattern(two, [[0, 1, 2, 3], [1, 2, 3, 4]])
    assert p.nnzb_upper == 14


def kuhn_counts(n):
    """Closed forms for an n^3 Kuhn grid with the x = 0 face fixed."""
    nodes = (n + 1) ** 3
    nd = nodes - (n + 1) ** 2
    # Kuhn edges: axis edges 3 n (n+1)^2, face diagonals 3 n^2 (n+1), one body diagonal per cell
    edges = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    # edges with both ends dynamic: drop those touching the fixed face (x = 0)
    fixed_face = (n + 1) ** 2
    e_in_face = 2 * n * (n + 1) + n * n  # axis edges in y, z plus one diagonal per face cell
    e_leaving = fixed_face + n * (n + 1) * 2  # x edges plus the xy and xz diagonals leaving the face
    e_leaving += n * n  # body diagonals leaving the face
    dyn_edges = edges - e_in_face - e_leaving
    return nd, nd + dyn_edges, nd + 2 * dyn_edges


@pytest.mark.parametrize("n", [2, 4, 8])
def test_kuhn_closed_forms(n):
    w = W.twist64(n=n, h=1.0 / n)
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed)
    assert (p.nd, p.nnzb_upper, p.nnzb_full) == kuhn_counts(n)


def test_twist64_counts():
    # the benchmark mesh: 274,625 nodes, 270,400 dynamic rows
    nd, nu, nf = kuhn_counts(64)
    assert (nd, nu, nf) == (270400, 2113407, 3956414)
    w = W.twist64()
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed)
    assert (p.nd, p.nnzb_upper, p.nnzb_full) == (nd, nu, nf)
    # every upper slot is hit by at least one live tet and the scatter map is canonical
    sc = p.scatter.reshape(-1)
    hit = np.zeros(nu, bool)
    hit[sc[sc >= 0]] = True
    assert hit.all()


def test_batched_pattern_is_block_diagonal():
    w = W.batch256(bodies=4, lo=2, hi=4)
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed, node_group=w.node_group)
    row_begin = 0
    for b in range(w.n_bodies):
        pb = Pattern(*[getattr(w.body(b), k) for k in ("rest", "tets")], node_fixed=w.body(b).fixed)
        n0, n1 = w.body_nodes[b]
        assert np.array_equal(p.dyn[n0:n1][pb.dyn >= 0], pb.dyn[pb.dyn >= 0] + row_begin)
        rs = p.upper_start[row_begin:row_begin + pb.nd + 1]
        assert np.array_equal(rs - rs[0], pb.upper_start)
        assert np.array_equal(p.upper_col[rs[0]:rs[-1]] - row_begin, pb.upper_col)
        row_begin += pb.nd
    assert row_begin == p.nd


def test_invalid_meshes_are_rejected():
    with pytest.raises(RuntimeError, match="out of range"):
        Pattern(np.eye(4, 3), [[0, 1, 2, 4]])
    two = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1]], float)
    with pytest.raises(RuntimeError, match="spans two groups"):
        Pattern(two, [[0, 1, 2, 3], [1, 2, 3, 4]], node_group=[0, 0, 0, 0, 1])
