"""The assembly gather on meshes whose rows are far from the usual ~24 tets
per node, against the FP64 oracle.

The gather runs by (row, tet) pairs in chunks of consecutive rows (at most
256 pairs, 384 off-diagonal contributions, 48 rows and 224 upper slots per
chunk); a system with a row larger than a chunk gathers by rows instead.
A fan -- a ring of n nodes around a centre node, with an apex above and one
below -- puts 2n tets and up to 6n off-diagonal contributions on the centre
node's row: n = 40 and 60 fill one pair chunk with it, n = 150 takes the row
gather. Either way the matrix matches the oracle's entrywise to 1e-12 of the
largest entry and an FP64 step to 1e-9, in full and upper-only storage.
"""

import numpy as np
import pytest

from checkers import Oracle
from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import Pattern

pytestmark = pytest.mark.gpu


def fan(n, radius=0.1, height=0.08):
    """Centre node 0, apexes 1 (above) and 2 (below), ring nodes 3 .. n + 2;
    tets (centre, ring i, ring i+1, apex) oriented to positive volume. The
    first quarter of the ring is fixed; the rest hangs in gravity."""
    ang = 2.0 * np.pi * np.arange(n) / n
    ring = np.stack([radius * np.cos(ang), radius * np.sin(ang), np.zeros(n)], axis=1)
    rest = np.vstack([[0.0, 0.0, 0.0], [0.0, 0.0, height], [0.0, 0.0, -height], ring])
    tets = []
    for i in range(n):
        a, b = 3 + i, 3 + (i + 1) % n
        for apex in (1, 2):
            tet = [0, a, b, apex]
            p = rest[tet]
            if np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])) < 0:
                tet = [0, b, a, apex]
            tets.append(tet)
    tets = np.array(tets, np.int32)
    fixed = np.zeros(len(rest), np.uint8)
    fixed[3:3 + max(1, n // 4)] = 1
    x0 = rest.copy()
    x0[0] += [0.0, 0.0, 0.01]  # a dent at the centre
    w = W._single(f"fan{n}", rest, tets, 1e5, 0.3, 1000.0, fixed, x0, np.zeros_like(rest), 1.0 / 60.0, 3, 1e-8,
                  500)
    w.gravity = (0.0, 0.0, -9.81)
    return w


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("n", [40, 60, 150])
@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP64_UPPER])
def test_fan_assembly_and_steps_match_the_oracle(n, precision):
    w = fan(n)
    print(f"case fan {n}, precision {precision}", flush=True)
    s = W.make_system(w, precision)
    o = Oracle(64, w.rest, w.tets, w.young, w.poisson, w.damping, w.density)
    o.set_params(w.gravity, 1, w.tol, w.max_iter)
    o.set_fixed(w.fixed)
    o.set_state(w.x0, w.v0)
    # the matrix of the first substep, entrywise (full BSR: every upper
    # block and its mirror), and the block-Jacobi inverse
    s.assemble(w.dt)
    vals, rhs, minv = s.matrix()
    ref = o.assemble(w.dt)
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed)
    scale = np.abs(ref["vals"]).max()
    for r in range(p.nd):
        for k in range(p.full_start[r], p.full_start[r + 1]):
            c = p.full_col[k]
            lo, hi = min(r, c), max(r, c)
            j = ref["row_start"][lo] + np.searchsorted(ref["col"][ref["row_start"][lo]:ref["row_start"][lo + 1]], hi)
            expect = ref["vals"][j] if r <= c else ref["vals"][j].T
            assert np.abs(vals[k] - expect).max() <= 1e-12 * scale, f"row {r} column {c}"
    assert np.abs(rhs - ref["rhs"]).max() <= 1e-11 * np.abs(ref["rhs"]).max()
    diag = np.array([ref["vals"][ref["row_start"][r]] for r in range(p.nd)])
    assert np.abs(minv - np.linalg.inv(diag)).max() <= 1e-12 * np.abs(minv).max()
    s.set_state(w.x0, w.v0)
    o.set_state(w.x0, w.v0)
    for k in range(w.steps):
        s.step(w.dt)
        assert o.step(w.dt) == 0
        xg, vg = s.state()
        xo, vo = o.state()
        it, res, conv, st = s.results()
        io, ro, co = o.cg()
        assert st[0] == N.GROUP_INTEGRATED
        assert rel(xg, xo) <= 1e-9 and rel(vg, vo) <= 1e-9, f"step {k}: {rel(xg, xo):.2e} {rel(vg, vo):.2e}"
        assert abs(int(it[0]) - io) <= 1
    s.close()


PAIRS = [(a, b) for a in range(4) for b in range(a, 4)]  # the 10 stored local pairs


def tet_order_sums(p, tets, blocks):
    """Each upper slot's sum of its tets' element blocks (the element kernel's
    inspection records), added in tet order in float64 from zero: the
    block of the local pair (min, max), transposed where the row's corner
    is the larger one (SolverData.cpp:407-450 adds them in this order)."""
    index = {}
    for r in range(p.nd):
        for k in range(p.full_start[r], p.full_start[r + 1]):
            index[(r, int(p.full_col[k]))] = k
    stored = {ab: i for i, ab in enumerate(PAIRS)}
    sums = {}
    for t, tet in enumerate(tets):
        rows = p.dyn[tet]
        for a in range(4):
            for b in range(4):
                ra, rb = int(rows[a]), int(rows[b])
                if ra < 0 or rb < 0 or ra > rb or (a != b and ra == rb):
                    continue
                blk = blocks[t, stored[(min(a, b), max(a, b))]]
                if a > b:
                    blk = blk.T
                k = index[(ra, rb)]
                sums[k] = sums[k] + blk if k in sums else 0.0 + blk
    return index, sums


@pytest.mark.parametrize("mesh", ["fan40", "fan150", "beam10", "twist8"])
@pytest.mark.parametrize("precision", [N.PK_FP64, N.PK_FP64_UPPER])
def test_matrix_is_the_tet_order_sum_of_the_element_blocks(mesh, precision):
    """Bitwise: A's entries are dt^2 times the tet-order sums of the blocks
    the inspection kernel reports (all kernels form a block with the same
    arithmetic; the pair gather and, for fan150, the row gather add them in
    tet order). The three entries of a diagonal block that carry the mass
    are left out (the lumped masses are not read back); mirrors are the
    transposes, and diagonal blocks their own."""
    # fans (pair gather; fan150 takes the row gather), a 5-tet box at rest,
    # a jittered Kuhn box twisted by pi/2
    w = fan(int(mesh[3:])) if mesh.startswith("fan") else W.beam10() if mesh == "beam10" else W.twist64(n=8)
    print(f"case {mesh}, precision {precision}", flush=True)
    s = W.make_system(w, precision)
    s.assemble(w.dt)
    vals, rhs, minv = s.matrix()
    blocks, forces, kv = s.elements()
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed)
    index, sums = tet_order_sums(p, w.tets, blocks)
    dt2 = w.dt * w.dt
    off = np.ones((3, 3), bool)
    off[[0, 1, 2], [0, 1, 2]] = False
    for (r, c), k in index.items():
        if r > c:
            continue
        expect = sums[k] * dt2
        if r == c:
            assert np.array_equal(vals[k][off], expect[off]), f"diagonal block of row {r}"
            # symmetric bit for bit: the inverse is kept as its upper triangle
            assert np.array_equal(vals[k], vals[k].T), f"diagonal block of row {r} not symmetric"
        else:
            assert np.array_equal(vals[k], expect), f"block ({r}, {c})"
            assert np.array_equal(vals[index[(c, r)]], expect.T), f"mirror of ({r}, {c})"
    s.close()
