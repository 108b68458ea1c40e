"""Seeded synthetic tet meshes for the implicit-Euler step.

Node numbering is x-fastest: node (i, j, k) of an (nx+1) x (ny+1) x (nz+1) grid
has index i + (nx+1) * (j + (ny+1) * k). Cells are numbered the same way over
nx x ny x nz and emit their tets consecutively, so tet order follows cell order.

Two hexahedron splits are provided:

* ``five``: the 5-tet split. Cell corners are numbered b = bx + 2 by + 4 bz.
  Cells with even (i + j + k) use the even corners {0, 3, 5, 6} for the central
  tet and cut one tet off each odd corner {1, 2, 4, 7}; odd cells swap the two
  sets, so neighbouring cells share face diagonals (a conforming mesh).
* ``kuhn``: the 6-tet Kuhn (Freudenthal) split along the cell diagonal 0 -> 7,
  one tet per axis permutation. It is translation invariant and conforming.

Every tet is emitted with positive orientation (det [x1-x0, x2-x0, x3-x0] > 0).

The random streams are PCG32 (XSH-RR, 64-bit LCG multiplier
6364136223846793005) seeded through SplitMix64: ``Pcg32(seed, stream)`` draws
initstate and initseq from SplitMix64 started at ``seed * 2 + stream`` and then
follows the reference ``pcg32_srandom_r`` seeding. Uniform reals are
u32 / 2**32, normals use Box-Muller on consecutive pairs.
"""

from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1
PCG_MULT = 6364136223846793005


def _splitmix64(x: int):
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


class Pcg32:
    """PCG32 XSH-RR generator with vectorised block draws (LCG jump-ahead)."""

    def __init__(self, seed: int, stream: int = 0):
        sm = _splitmix64((seed * 2 + stream) & MASK64)
        initstate = next(sm)
        initseq = next(sm)
        self.inc = ((initseq << 1) | 1) & MASK64
        self.state = 0
        self._step()
        self.state = (self.state + initstate) & MASK64
        self._step()

    def _step(self):
        self.state = (self.state * PCG_MULT + self.inc) & MASK64

    def u32(self, n: int) -> np.ndarray:
        """Next n outputs as uint32, in stream order."""
        if n <= 0:
            return np.zeros(0, dtype=np.uint32)
        states = np.empty(n, dtype=np.uint64)
        states[0] = self.state
        m = 1
        a, c = PCG_MULT, self.inc  # state_{k+m} = a_m * state_k + c_m
        with np.errstate(over="ignore"):
            while m < n:
                take = min(m, n - m)
                states[m:m + take] = states[:take] * np.uint64(a) + np.uint64(c)
                c = (a * c + c) & MASK64
                a = (a * a) & MASK64
                m += take
        last = int(states[-1])
        self.state = (last * PCG_MULT + self.inc) & MASK64
        old = states
        xorshifted = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = (old >> np.uint64(59)).astype(np.uint64)
        x = xorshifted
        out = ((x >> rot) | (x << ((np.uint64(32) - rot) & np.uint64(31)))) & np.uint64(0xFFFFFFFF)
        return out.astype(np.uint32)

    def uniform(self, n: int) -> np.ndarray:
        return self.u32(n).astype(np.float64) * (1.0 / 4294967296.0)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u = self.uniform(2 * m).reshape(m, 2)
        r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
        z = np.stack([r * np.cos(2.0 * np.pi * u[:, 1]), r * np.sin(2.0 * np.pi * u[:, 1])], axis=1)
        return z.reshape(-1)[:n]

    def randint(self, lo: int, hi: int, n: int) -> np.ndarray:
        """Uniform integers in [lo, hi] (inclusive)."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)


def grid_nodes(nx: int, ny: int, nz: int, h: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    k = np.arange(nz + 1, dtype=np.float64)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    pts = np.stack([I.ravel() * h, J.ravel() * h, K.ravel() * h], axis=1)
    return pts + np.asarray(origin, dtype=np.float64)


def grid_index(nx: int, ny: int, nz: int) -> np.ndarray:
    """(nz+1, ny+1, nx+1) array of node ids."""
    return np.arange((nx + 1) * (ny + 1) * (nz + 1), dtype=np.int64).reshape(nz + 1, ny + 1, nx + 1)


def _cell_corners(nx: int, ny: int, nz: int) -> np.ndarray:
    """(ncells, 8) node ids of cell corners b = bx + 2 by + 4 bz, cells x-fastest."""
    g = grid_index(nx, ny, nz)
    corners = []
    for b in range(8):
        bx, by, bz = b & 1, (b >> 1) & 1, (b >> 2) & 1
        corners.append(g[bz:bz + nz, by:by + ny, bx:bx + nx].ravel())
    return np.stack(corners, axis=1)


_FIVE_EVEN = np.array([[1, 0, 3, 5], [2, 0, 6, 3], [4, 0, 5, 6], [7, 3, 6, 5], [0, 3, 5, 6]])
_FIVE_ODD = np.array([[0, 1, 4, 2], [3, 1, 2, 7], [5, 1, 7, 4], [6, 2, 4, 7], [1, 2, 7, 4]])
_KUHN = np.array([[0, 1, 3, 7], [0, 1, 5, 7], [0, 2, 3, 7], [0, 2, 6, 7], [0, 4, 5, 7], [0, 4, 6, 7]])


def _orient(nodes: np.ndarray, tets: np.ndarray) -> np.ndarray:
    p = nodes[tets]
    d = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    out = tets.copy()
    neg = d < 0
    out[neg, 2], out[neg, 3] = tets[neg, 3], tets[neg, 2]
    return out


def tets_five(nx: int, ny: int, nz: int) -> np.ndarray:
    c = _cell_corners(nx, ny, nz)
    parity = ((np.arange(nx)[None, None, :] + np.arange(ny)[None, :, None] + np.arange(nz)[:, None, None]) % 2).ravel()
    even = c[:, _FIVE_EVEN]  # (ncells, 5, 4)
    odd = c[:, _FIVE_ODD]
    t = np.where(parity[:, None, None] == 0, even, odd)
    return t.reshape(-1, 4)


def tets_kuhn(nx: int, ny: int, nz: int) -> np.ndarray:
    c = _cell_corners(nx, ny, nz)
    return c[:, _KUHN].reshape(-1, 4)


def box_mesh(nx: int, ny: int, nz: int, h: float, split: str = "five", origin=(0.0, 0.0, 0.0)):
    """Rest node positions (N, 3) float64 and positively oriented tets (T, 4) int32."""
    nodes = grid_nodes(nx, ny, nz, h, origin)
    if split == "five":
        tets = tets_five(nx, ny, nz)
    elif split == "kuhn":
        tets = tets_kuhn(nx, ny, nz)
    else:
        raise ValueError(f"unknown split {split!r}")
    return nodes, _orient(nodes, tets).astype(np.int32)


def face_nodes(nx: int, ny: int, nz: int, face: int) -> np.ndarray:
    """Node ids on one box face: 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z."""
    g = grid_index(nx, ny, nz)
    sel = [g[:, :, 0], g[:, :, -1], g[:, 0, :], g[:, -1, :], g[0, :, :], g[-1, :, :]][face]
    return np.sort(sel.ravel())


def jitter(nodes: np.ndarray, h: float, frac: float, rng: Pcg32) -> np.ndarray:
    """Adds U(-frac*h, +frac*h) per coordinate, drawn node by node, x, y, z."""
    u = rng.uniform(nodes.size).reshape(nodes.shape)
    return nodes + (2.0 * u - 1.0) * (frac * h)


def twist_about_x(rest: np.ndarray, nx: int, ny: int, nz: int, h: float, angle: float) -> np.ndarray:
    """Rotates every node about the x axis through the box centre line by
    angle * x / (nx h), with x the node's own rest coordinate, so the map is a
    smooth twist (volume preserving up to discretisation) whose far face turns
    by ``angle``. Nodes of the x = 0 grid face keep their rest position."""
    g = grid_index(nx, ny, nz)
    theta = angle * rest[:, 0] / (nx * h)
    theta[g[:, :, 0].ravel()] = 0.0
    cy = 0.5 * ny * h
    cz = 0.5 * nz * h
    y = rest[:, 1] - cy
    z = rest[:, 2] - cz
    c, s = np.cos(theta), np.sin(theta)
    out = rest.copy()
    # displacement form: nodes with theta = 0 keep their exact coordinates
    out[:, 1] = rest[:, 1] + ((c - 1.0) * y - s * z)
    out[:, 2] = rest[:, 2] + (s * y + (c - 1.0) * z)
    return out
