"""Launch-shape autotuner for the CG kernels of a ``System``.

The fused SpMV is the dominant kernel of the step. Its launch shape -- lanes
per BSR row, rows per workgroup, and rows each lane group loads before
reducing -- decides how many bytes each wave keeps in flight and how well the
row lengths fill the lanes. The best shape depends on the row-length
distribution, the storage precision and the number of rows, so it is measured
rather than guessed:

  stage 1  lanes per row, with (128 rows, unroll 2) (upper-only storage:
           one lane per row, no search);
  stage 2  rows per workgroup x unroll, with the best lane count (of shapes
           within 1 % of the fastest, the one with the fewest rows);
  stage 3  rows per vector-kernel workgroup: the grid of the direction and
           update kernels, whose every workgroup first re-sums the other
           kernel's partials; this stage minimises SpMV + direction + update.

Each candidate runs one implicit step from the same state with kernel timing
on (the SpMV's time is its in-kernel wall-clock span); the state is restored
afterwards, so tuning leaves the simulation where it was. Results are cached
per matrix signature (rows, blocks, bodies, precision) in a JSON file so a
second system of the same shape skips the search.
"""

from __future__ import annotations

import json
import os

from . import native as N

LANES = (4, 8, 16, 32)
ROWS = (64, 128, 256, 512)
UNROLL = (1, 2, 4)
VEC_ROWS = (256, 512, 1024, 2048)


def _valid(lanes, unroll):
    return unroll in (1, 2) or (unroll == 4 and lanes <= 16)


def signature(system):
    return f"nd{system.nd}-nnzf{system.nnzb_full}-g{system.groups}-p{system.precision}"


def _measure(system, shape, dt, x, v):
    system.tune(*shape)
    system.set_state(x, v)
    system.profile(True)
    system.step(dt)
    n_s, ms_s = system.kernel_time(N.K_SPMV)
    n_u, ms_u = system.kernel_time(N.K_UPDATE)
    n_d, ms_d = system.kernel_time(N.K_DIRECTION)
    system.profile(False)
    spmv = ms_s / n_s if n_s else float("inf")
    vec = (ms_u + ms_d) / n_u if n_u else float("inf")  # update + direction per iteration
    return spmv, vec


def autotune(system, dt, cache=None, verbose=False):
    """Pick and apply the fastest CG launch shape for ``system``.

    ``dt`` is the step used for the trial steps (the workload's own step, so
    the matrix is the one the solves will see). ``cache`` is a JSON path; an
    entry for this system's signature is applied without measuring. Returns
    the chosen (lanes_per_row, spmv_rows, vector_rows, unroll)."""
    key = signature(system)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
        if key in table:
            shape = tuple(table[key])
            system.tune(*shape)
            return shape
    if system.nd == 0:
        return system.get_tune()
    x, v = system.state()
    _, _, vec0, _ = system.get_tune()
    log = []

    def trial(shape):
        spmv, upd = _measure(system, shape, dt, x, v)
        log.append((shape, spmv, upd))
        if verbose:
            print(f"autotune {shape}: spmv {spmv * 1e3:.1f} us, direction + update {upd * 1e3:.1f} us", flush=True)
        return spmv, upd

    # upper-only storage runs one lane per row over 64-row slices: no lane stage
    upper = system.precision in (N.PK_FP32_UPPER, N.PK_FP64_UPPER)
    best_l = 1 if upper else min(LANES, key=lambda lanes: trial((lanes, 128, vec0, 2))[0])
    # upper-only storage does not use unroll
    cands = [(best_l, r, vec0, u) for r in ROWS for u in ((1,) if upper else UNROLL) if _valid(best_l, u)]
    times = {s: trial(s)[0] for s in cands}
    fastest = min(times.values())
    # near-ties (within 1 %) go to the smaller tile: each SpMV tile re-reads the
    # band of the direction vector around its rows, so smaller tiles that run
    # as fast move no more bytes (batch128 FP32 upper: 256-row tiles read
    # 1.03x the algorithmic bytes, 512-row tiles 1.29x, in the same time)
    best = min((s for s, t in times.items() if t <= 1.01 * fastest), key=lambda s: (s[1], times[s]))
    best = min(((best[0], best[1], vr, best[3]) for vr in VEC_ROWS), key=lambda s: sum(trial(s)))
    system.tune(*best)
    system.set_state(x, v)
    if cache:
        table[key] = list(best)
        d = os.path.dirname(os.path.abspath(cache))
        os.makedirs(d, exist_ok=True)
        with open(cache, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return best
