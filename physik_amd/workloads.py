"""Seeded synthetic workloads for the implicit-Euler step.

Each function returns a ``Workload``: one concatenated node/tet set (bodies
occupy consecutive node ranges, ``node_group`` names the body of each node),
initial state, per-tet material, and the step settings. Parameters the task
statement leaves open are fixed here and recorded in the docstrings.

Config names follow the five benchmark settings of this project:

1. ``beam10``     10^3 cells of 0.1 m, 5-tet split (5,000 tets, 1,331 nodes),
                  E 1e5, nu 0.3, rho 1000, g = -9.81 z, x = 0 face fixed,
                  dt 1/60, 100 steps, rtol 1e-6, max 200 iterations.
2. ``twist64``    64^3 cells, Kuhn 6-tet split (1,572,864 tets, 274,625 nodes),
                  rest nodes jittered U(+-10 % of the edge) with seed 1,
                  E 1e6, nu 0.45, x = 0 face fixed, initial 90-degree twist,
                  50 steps dt 1/120, rtol 1e-6.
3. ``batch256``   256 bodies of n_x x n_y x n_z cells, n ~ U{4..32} per axis
                  (seed 2), 5-tet split, E log-uniform [1e4, 1e6], nu
                  U[0.3, 0.45], rho 1000, one random face fixed, v0 ~ N(0, 0.5)
                  per node and axis, 100 steps dt 1/60, rtol 1e-6, one CG per
                  body.

Choices not given by the statement:
  * cell edge: 0.1 m for beam10 and batch256, 1/32 m for twist64 and
    batch128 (keeps the Kuhn det(Dm) = h^3 = 3.1e-5 well above the 1e-6
    inverse floor the reference applies, even after jitter), 1/64 m for
    diff32 (det(Dm) = 3.8e-6 with no jitter; the smaller body stays clear of
    inverted tets, see diff32);
  * gravity -9.81 z and rho 1000 everywhere, damping 0 everywhere;
  * CG iteration cap 1024 (the reference's clamp) wherever none is stated;
  * the twist turns grid plane i about the x axis through the box centre by
    90 deg * i / n (linear in x: the fixed face stays at rest, the far face
    turns 90 degrees, no tet inverts);
  * the random face of batch256 is floor(6 u) in the order -x, +x, -y, +y,
    -z, +z; per-body draws come from Pcg32(2, stream 0) in the order n_x,
    n_y, n_z, E, nu, face; velocities from Pcg32(2, stream 1) body by body,
    node by node, x y z; fixed nodes start at rest (fixing zeroes velocity).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import meshgen as mg

G = (0.0, 0.0, -9.81)


@dataclass
class Workload:
    name: str
    rest: np.ndarray
    tets: np.ndarray
    young: np.ndarray
    poisson: np.ndarray
    damping: np.ndarray
    density: np.ndarray
    fixed: np.ndarray
    x0: np.ndarray
    v0: np.ndarray
    node_group: np.ndarray | None
    dt: float
    steps: int
    tol: float
    max_iter: int
    gravity: tuple = G
    body_nodes: list = field(default_factory=list)  # (begin, end) node range per body
    body_tets: list = field(default_factory=list)   # (begin, end) tet range per body
    # differentiable workloads: loss = sum over loss_nodes of |x_S - target|^2
    loss_nodes: np.ndarray | None = None
    target: np.ndarray | None = None
    checkpoint: int = 0                 # substeps between checkpoints of the recorded run
    adj_tol: float = 1e-6               # adjoint CG tolerance

    @property
    def n_nodes(self):
        return self.rest.shape[0]

    @property
    def n_tets(self):
        return self.tets.shape[0]

    @property
    def n_bodies(self):
        return max(1, len(self.body_nodes))

    def body(self, b):
        """The workload restricted to body b (node ids rebased to 0)."""
        n0, n1 = self.body_nodes[b]
        t0, t1 = self.body_tets[b]
        w = Workload(f"{self.name}[{b}]", self.rest[n0:n1], self.tets[t0:t1] - n0, self.young[t0:t1],
                     self.poisson[t0:t1], self.damping[t0:t1], self.density[t0:t1], self.fixed[n0:n1],
                     self.x0[n0:n1], self.v0[n0:n1], None, self.dt, self.steps, self.tol, self.max_iter,
                     self.gravity, [(0, n1 - n0)], [(0, t1 - t0)])
        if self.loss_nodes is not None:
            sel = (self.loss_nodes >= n0) & (self.loss_nodes < n1)
            w.loss_nodes = self.loss_nodes[sel] - n0
            w.target = self.target[sel]
            w.checkpoint, w.adj_tol = self.checkpoint, self.adj_tol
        return w


def _single(name, rest, tets, young, poisson, rho, fixed, x0, v0, dt, steps, tol, max_iter):
    t = tets.shape[0]
    return Workload(name, rest, tets.astype(np.int32), np.full(t, young, np.float64), np.full(t, poisson),
                    np.zeros(t), np.full(t, rho), fixed.astype(np.uint8), x0, v0, None, dt, steps, tol, max_iter,
                    G, [(0, rest.shape[0])], [(0, t)])


def beam10(n=10, h=0.1):
    rest, tets = mg.box_mesh(n, n, n, h, "five")
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(n, n, n, 0)] = 1
    return _single(f"beam{n}", rest, tets, 1e5, 0.3, 1000.0, fixed, rest.copy(), np.zeros_like(rest), 1.0 / 60.0,
                   100, 1e-6, 200)


def twist64(n=64, h=1.0 / 32.0, steps=50):
    rest, tets = mg.box_mesh(n, n, n, h, "kuhn")
    rest = mg.jitter(rest, h, 0.10, mg.Pcg32(1))
    x0 = mg.twist_about_x(rest, n, n, n, h, np.pi / 2)
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(n, n, n, 0)] = 1
    return _single(f"twist{n}", rest, tets, 1e6, 0.45, 1000.0, fixed, x0, np.zeros_like(rest), 1.0 / 120.0, steps,
                   1e-6, 1024)


def batch256(bodies=256, seed=2, lo=4, hi=32, h=0.1, steps=100):
    rng = mg.Pcg32(seed, 0)
    rngv = mg.Pcg32(seed, 1)
    rests, tetss, fixeds, v0s, youngs, poissons, groups = [], [], [], [], [], [], []
    body_nodes, body_tets = [], []
    n_off = t_off = 0
    for b in range(bodies):
        u = rng.uniform(6)
        nx, ny, nz = (lo + np.floor(u[:3] * (hi - lo + 1))).astype(int)
        young = float(np.exp(np.log(1e4) + u[3] * (np.log(1e6) - np.log(1e4))))
        poisson = 0.3 + 0.15 * u[4]
        face = int(np.floor(6 * u[5]))
        rest, tets = mg.box_mesh(nx, ny, nz, h, "five")
        fixed = np.zeros(len(rest), np.uint8)
        fixed[mg.face_nodes(nx, ny, nz, face)] = 1
        v0 = 0.5 * rngv.normal(rest.size).reshape(-1, 3)
        v0[fixed == 1] = 0.0
        rests.append(rest)
        tetss.append(tets + n_off)
        fixeds.append(fixed)
        v0s.append(v0)
        youngs.append(np.full(len(tets), young))
        poissons.append(np.full(len(tets), poisson))
        groups.append(np.full(len(rest), b, np.int32))
        body_nodes.append((n_off, n_off + len(rest)))
        body_tets.append((t_off, t_off + len(tets)))
        n_off += len(rest)
        t_off += len(tets)
    rest = np.concatenate(rests)
    tets = np.concatenate(tetss).astype(np.int32)
    t = len(tets)
    return Workload(f"batch{bodies}", rest, tets, np.concatenate(youngs), np.concatenate(poissons), np.zeros(t),
                    np.full(t, 1000.0), np.concatenate(fixeds), rest.copy(), np.concatenate(v0s),
                    np.concatenate(groups), 1.0 / 60.0, steps, 1e-6, 1024, G, body_nodes, body_tets)


def diff32(n=32, h=1.0 / 64.0, steps=50, seed=3):
    """Differentiable single body: n^3-cell Kuhn beam (6 n^3 tets), per-tet
    log E ~ N(log 1e5, 0.2) from Pcg32(seed, 0), nu 0.3, rho 1000, x = 0 face
    fixed, starting at rest; loss on the +x face against its rest positions
    moved 0.05 m down; 50 substeps of 1/60 s, checkpoint every 10, forward
    and adjoint CG rtol 1e-6.

    Cell edge 1/64 m (a 0.5 m cube at n = 32): its free end sags a few cm,
    the scale of the 0.05 m target, and no tet inverts over the 50 substeps
    (min det F = 0.47), so the run is a smooth function of log E and v0; a
    1 m cube at E = 1e5 collapses through det F = 0, where the polar flips a
    column and finite differences stop meaning anything.  Kuhn det Dm =
    h^3 = 3.8e-6 stays clear of the 1e-6 inverse floor (no jitter here)."""
    rest, tets = mg.box_mesh(n, n, n, h, "kuhn")
    t = len(tets)
    young = np.exp(np.log(1e5) + 0.2 * mg.Pcg32(seed, 0).normal(t))
    fixed = np.zeros(len(rest), np.uint8)
    fixed[mg.face_nodes(n, n, n, 0)] = 1
    end = mg.face_nodes(n, n, n, 1)
    w = _single(f"diff{n}", rest, tets, 1e5, 0.3, 1000.0, fixed, rest.copy(), np.zeros_like(rest), 1.0 / 60.0,
                steps, 1e-6, 1024)
    w.young = young
    w.loss_nodes = end.astype(np.int64)
    w.target = rest[end] - np.array([0.0, 0.0, 0.05])
    w.checkpoint = 10
    w.adj_tol = 1e-6
    return w


def batch128(bodies=128, seed=4, lo=8, hi=24, h=1.0 / 32.0, steps=40):
    """Batched differentiable bodies: n_x, n_y, n_z ~ U{lo..hi} per body
    (Pcg32(seed, 0)), Kuhn split, per-tet log E ~ N(log 1e5, 0.2)
    (Pcg32(seed, 1), body by body), nu 0.3, rho 1000, x = 0 face fixed, at
    rest; per-body loss on the +x face against its rest positions moved by
    a per-body offset U[-0.05, 0.05]^3 m (Pcg32(seed, 2)); 40 substeps of
    1/60 s, checkpoint every 8, forward CG rtol 1e-6, adjoint rtol 1e-7."""
    rng, rngE, rngT = mg.Pcg32(seed, 0), mg.Pcg32(seed, 1), mg.Pcg32(seed, 2)
    rests, tetss, fixeds, youngs, groups, ends, targets = [], [], [], [], [], [], []
    body_nodes, body_tets = [], []
    n_off = t_off = 0
    for b in range(bodies):
        nx, ny, nz = (int(v) for v in rng.randint(lo, hi, 3))
        rest, tets = mg.box_mesh(nx, ny, nz, h, "kuhn")
        fixed = np.zeros(len(rest), np.uint8)
        fixed[mg.face_nodes(nx, ny, nz, 0)] = 1
        end = mg.face_nodes(nx, ny, nz, 1)
        off = 0.05 * (2.0 * rngT.uniform(3) - 1.0)
        rests.append(rest)
        tetss.append(tets + n_off)
        fixeds.append(fixed)
        youngs.append(np.exp(np.log(1e5) + 0.2 * rngE.normal(len(tets))))
        groups.append(np.full(len(rest), b, np.int32))
        ends.append(end + n_off)
        targets.append(rest[end] + off)
        body_nodes.append((n_off, n_off + len(rest)))
        body_tets.append((t_off, t_off + len(tets)))
        n_off += len(rest)
        t_off += len(tets)
    rest = np.concatenate(rests)
    tets = np.concatenate(tetss).astype(np.int32)
    t = len(tets)
    w = Workload(f"batch{bodies}", rest, tets, np.concatenate(youngs), np.full(t, 0.3), np.zeros(t),
                 np.full(t, 1000.0), np.concatenate(fixeds), rest.copy(), np.zeros_like(rest),
                 np.concatenate(groups), 1.0 / 60.0, steps, 1e-6, 1024, G, body_nodes, body_tets)
    w.loss_nodes = np.concatenate(ends).astype(np.int64)
    w.target = np.concatenate(targets)
    w.checkpoint = 8
    w.adj_tol = 1e-7
    return w


def loss_and_seed(w: Workload, x: np.ndarray):
    """Loss of a final state and its gradient dL/dx_S (node_count x 3)."""
    d = x[w.loss_nodes] - w.target
    xbar = np.zeros_like(x)
    xbar[w.loss_nodes] = 2.0 * d
    return float(np.sum(d * d)), xbar


def make_system(w: Workload, precision=0, batched=True):
    """A System for the workload with its initial state and solver settings."""
    from .system import System
    s = System(w.rest, w.tets, w.young, w.poisson, w.damping, w.density, precision=precision, node_fixed=w.fixed,
               node_group=w.node_group if batched else None)
    s.set_state(w.x0, w.v0)
    s.set_gravity(w.gravity)
    s.set_solver(w.tol, w.max_iter)
    return s
