"""SpMV direction segments staged in LDS (full and upper-triangle storage).

When every SpMV tile's column window fits the capacity, tiles copy their
windows of d into LDS; otherwise they copy up to four segments of d, and a
tile needing more than the capacity gathers from global memory. The products
and sums are the same, so a system with staging on must step bit for bit like
the same system with it off, at every launch shape, both in a batch whose
windows all fit and in one whose segments put tiles on each side of the
capacity.
"""

import numpy as np
import pytest

from physik_amd import meshgen as mg
from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.system import Pattern

from launch_shapes import shape_for

pytestmark = pytest.mark.gpu

def mixed_batch(seed=21, bar=True):
    """Two small boxes, whose tiles' windows fit, and (bar) a 1,400 x 3 x 3
    cell bar, whose tiles reference columns in clusters 1,401 (y) and 5,604
    (z) rows apart, so that even their four segments need about 7,000 rows
    (beyond any capacity: at most a third of a CU's 160 KB, 4,437 FP32 rows;
    the corner lines' tiles need fewer)."""
    rng = mg.Pcg32(seed)
    rests, tetss, fixeds, v0s, groups, body_nodes, body_tets = [], [], [], [], [], [], []
    n_off = t_off = 0
    for b, (nx, ny, nz) in enumerate([(4, 3, 3), (8, 6, 5)] + ([(1400, 3, 3)] if bar else [])):
        h = 0.05
        rest, tets = mg.box_mesh(nx, ny, nz, h, "kuhn")
        rest = mg.jitter(rest, h, 0.1, rng)
        fixed = np.zeros(len(rest), np.uint8)
        fixed[mg.face_nodes(nx, ny, nz, 0)] = 1
        v0 = 0.2 * rng.normal(rest.size).reshape(-1, 3)
        v0[fixed == 1] = 0.0
        rests.append(rest)
        tetss.append(tets + n_off)
        fixeds.append(fixed)
        v0s.append(v0)
        groups.append(np.full(len(rest), b, np.int32))
        body_nodes.append((n_off, n_off + len(rest)))
        body_tets.append((t_off, t_off + len(tets)))
        n_off += len(rest)
        t_off += len(tets)
    rest = np.concatenate(rests)
    tets = np.concatenate(tetss).astype(np.int32)
    t = len(tets)
    return W.Workload("mixed", rest, tets, np.full(t, 2e5), np.full(t, 0.35), np.zeros(t), np.full(t, 1000.0),
                      np.concatenate(fixeds), rest.copy(), np.concatenate(v0s), np.concatenate(groups), 1.0 / 60.0,
                      4, 1e-7, 600, W.G, body_nodes, body_tets)


def staged_rows(cols):
    """Rows the library stages for a tile referencing `cols`: the distinct
    columns cut at their three widest gaps (ties: the earlier gap) into up to
    four runs, each staged from its first to its last column."""
    cols = np.unique(cols)
    gaps = [(cols[i + 1] - cols[i] - 1, i) for i in range(len(cols) - 1) if cols[i + 1] - cols[i] > 1]
    gaps.sort(key=lambda g: (-g[0], g[1]))
    ends = sorted(g[1] for g in gaps[:3]) + [len(cols) - 1]
    total, first = 0, 0
    for e in ends:
        total += cols[e] - cols[first] + 1
        first = e + 1
    return total


def windows(w, rows_per_tile):
    """Window widths and segment totals of every SpMV tile, as the library
    forms them."""
    p = Pattern(w.rest, w.tets, node_fixed=w.fixed, node_group=w.node_group)
    dyn_group = w.node_group[np.flatnonzero(p.dyn >= 0)]
    wins, segs = [], []
    for g in range(w.n_bodies):
        rows = np.flatnonzero(dyn_group == g)
        for r0 in range(rows[0], rows[-1] + 1, rows_per_tile):
            r1 = min(r0 + rows_per_tile, rows[-1] + 1)
            cols = np.concatenate([p.full_col[p.full_start[r0]:p.full_start[r1]], np.arange(r0, r1)])
            wins.append(cols.max() - cols.min() + 1)
            segs.append(staged_rows(cols))
    return np.array(wins), np.array(segs)


@pytest.mark.parametrize("bar", [True, False])
@pytest.mark.parametrize("precision", [N.PK_FP32_UPPER, N.PK_FP64_UPPER, N.PK_FP32_STORAGE, N.PK_FP64])
@pytest.mark.parametrize("shape", [(8, 128, 512, 2), (4, 64, 256, 1), (16, 32, 256, 4), (32, 256, 1024, 2)])
def test_staged_tiles_step_like_global_gathers(precision, shape, bar):
    w = mixed_batch(bar=bar)
    shape = shape_for(precision, *shape)
    a = W.make_system(w, precision)
    b = W.make_system(w, precision)
    a.tune(*shape)
    b.tune(*shape)
    a.set_staging(1)
    b.set_staging(0)
    got_mode, rows, segmented = a.get_staging()
    cap = a.stage_capacity()
    assert cap > 0
    wins, segs = windows(w, shape[1])
    print(f"precision {precision} shape {shape} bar {bar}: staged rows {rows}, segmented {segmented}, "
          f"windows {wins.min()}..{wins.max()}, segments {segs.min()}..{segs.max()}", flush=True)
    assert got_mode == 1
    if wins.max() <= cap:
        assert not segmented and rows == wins.max()
    else:
        assert segmented and rows == min(segs.max(), cap)
    if bar:
        assert segmented and (segs <= rows).any() and (segs > rows).any(), "tiles on both sides of the capacity"
    else:
        assert not segmented
    assert b.get_staging() == (0, 0, False)
    for n in range(w.steps):
        a.step(w.dt)
        b.step(w.dt)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"step {n}"
        for u, v in zip(a.results(), b.results()):
            assert np.array_equal(u, v), f"step {n} results"
    assert a.results()[0].min() > 5  # the solves did iterate
    a.close()
    b.close()


def test_staging_is_switchable():
    w = mixed_batch()
    for precision in (N.PK_FP32_STORAGE, N.PK_FP32_UPPER):
        s = W.make_system(w, precision)
        mode, rows, segmented = s.get_staging()
        assert mode == 1 and segmented and rows == s.stage_capacity()  # the bar's tiles need more than it
        s.set_staging(0)
        assert s.get_staging() == (0, 0, False) and s.stage_capacity() == 0
        s.set_staging(1)
        assert s.get_staging() == (1, rows, True)
        assert s.lib.pk_system_set_spmv_staging(s.h, 2) == N.PK_ERR_ARG
        s.close()
    s = W.make_system(mixed_batch(bar=False), N.PK_FP32_UPPER)
    assert s.get_staging()[1:] == (s.get_staging()[1], False)
    s.close()
