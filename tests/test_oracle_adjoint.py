"""The oracle's reverse-mode step against central finite differences.

The reference has no adjoint, so the FP64 CPU adjoint (oracle/physik_oracle.cpp,
``adjoint_run``) is pinned here by differentiating the oracle's own forward
map numerically: gradients of a loss of the final state with respect to
per-tet log E, initial velocities and initial positions must agree with
central differences (h = 1e-6) of full forward runs. Forward and adjoint CG
run to tol 1e-12 so the solve error stays far below the differencing error.
The polar rotation's vector-Jacobian product is checked on its own, including
the negated-column branch of inverted tets and the Identity fallback.
"""

import numpy as np
import pytest

from checkers import Oracle, oracle_polar, oracle_polar_vjp
from physik_amd import meshgen as mg

H = 1e-6


def polar_dot(F, G):
    return float(np.sum(G * oracle_polar(F)[0]))


@pytest.mark.parametrize("case", ["stretch", "shear", "inverted", "rotated"])
def test_polar_vjp_matches_finite_differences(case):
    rng = np.random.default_rng({"stretch": 1, "shear": 2, "inverted": 3, "rotated": 4}[case])
    F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    if case == "inverted":
        F[:, 2] *= -1.0
    if case == "rotated":
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        F = q @ np.diag([1.2, 0.9, 1.1])
    G = rng.standard_normal((3, 3))
    R, Fb, kind = oracle_polar_vjp(F, G)
    assert kind == (1 if np.linalg.det(F) < 0 else 0)
    fd = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += H
            Fm[i, j] -= H
            fd[i, j] = (polar_dot(Fp, G) - polar_dot(Fm, G)) / (2 * H)
    assert np.abs(Fb - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max()), (Fb, fd)


def test_polar_vjp_of_a_rotation_is_the_skew_projection():
    # for F = Q (a rotation), S = I and dR = Q skew(Q^T dF): dL/dF = Q skew(Q^T G)
    q, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    G = np.random.default_rng(6).standard_normal((3, 3))
    _, Fb, kind = oracle_polar_vjp(q, G)
    A = q.T @ G
    assert kind == 0 and np.allclose(Fb, q @ (0.5 * (A - A.T)), atol=1e-10)


def test_polar_vjp_identity_fallback_is_zero():
    F = np.diag([1.0, 1.0, 1e-9])  # |det F| <= 1e-8: R = I, independent of F
    R, Fb, kind = oracle_polar_vjp(F, np.ones((3, 3)))
    assert kind == 2 and np.array_equal(R, np.eye(3)) and not Fb.any()


class Problem:
    """A jittered 2x3x2 Kuhn block, x = 0 face fixed, per-tet E, damping,
    gravity, an initial twist and velocity; loss on the far face."""

    def __init__(self, steps=4, damping=0.4):
        nx, ny, nz, h = 2, 3, 2, 0.1
        rest, tets = mg.box_mesh(nx, ny, nz, h, "kuhn")
        self.rest = mg.jitter(rest, h, 0.1, mg.Pcg32(7))
        self.tets = tets
        self.n, self.t = len(rest), len(tets)
        rng = np.random.default_rng(8)
        self.logE = np.log(1e4) + 0.2 * rng.standard_normal(self.t)
        self.fixed = np.zeros(self.n, np.uint8)
        self.fixed[mg.face_nodes(nx, ny, nz, 0)] = 1
        self.end = mg.face_nodes(nx, ny, nz, 1)
        self.x0 = mg.twist_about_x(self.rest, nx, ny, nz, h, 0.4)
        self.v0 = 0.3 * rng.standard_normal((self.n, 3))
        self.v0[self.fixed == 1] = 0.0
        self.target = self.x0[self.end] + np.array([0.0, 0.02, -0.05])
        self.damping, self.steps, self.dt = damping, steps, 1.0 / 60.0

    def world(self, logE=None, x0=None, v0=None):
        o = Oracle(64, self.rest, self.tets, np.exp(self.logE if logE is None else logE), 0.3, self.damping, 1000.0)
        o.set_params((0.0, 0.0, -9.81), 1, 1e-12, 1000)
        o.set_fixed(self.fixed)
        o.set_state(self.x0 if x0 is None else x0, self.v0 if v0 is None else v0)
        return o

    def loss_terms(self, x, v):
        d = x[self.end] - self.target
        return float(np.sum(d * d) + 0.1 * np.sum(v * v))

    def loss(self, **kw):
        o = self.world(**kw)
        for _ in range(self.steps):
            assert o.step(self.dt) == 0
        return self.loss_terms(*o.state())

    def adjoint(self):
        o = self.world()
        # record from the initial state; seeds are dL/dx_S, dL/dv_S
        probe = self.world()
        for _ in range(self.steps):
            probe.step(self.dt)
        xs, vs = probe.state()
        xbar = np.zeros((self.n, 3))
        xbar[self.end] = 2 * (xs[self.end] - self.target)
        vbar = 0.2 * vs
        gE, gx, gv, its = o.adjoint(self.dt, self.steps, xbar, vbar, tol=1e-12, max_iter=1000)
        assert np.array_equal(o.state()[0], xs)  # left at the final state
        return gE, gx, gv, its


@pytest.fixture(scope="module")
def problem():
    p = Problem()
    return p, p.adjoint()


def central(f, h=H):
    return (f(h) - f(-h)) / (2 * h)


def test_log_young_gradient_matches_finite_differences(problem):
    p, (gE, _, _, its) = problem
    assert (its > 0).all()
    for e in np.random.default_rng(9).choice(p.t, 8, replace=False):
        def f(d, e=e):
            le = p.logE.copy()
            le[e] += d
            return p.loss(logE=le)
        fd = central(f)
        assert abs(gE[e] - fd) <= 1e-8 * np.abs(gE).max() + 1e-6 * abs(fd), (e, gE[e], fd)


def test_initial_velocity_gradient_matches_finite_differences(problem):
    p, (_, _, gv, _) = problem
    free = np.flatnonzero(p.fixed == 0)
    rng = np.random.default_rng(10)
    for i in rng.choice(free, 6, replace=False):
        for k in range(3):
            def f(d, i=i, k=k):
                v = p.v0.copy()
                v[i, k] += d
                return p.loss(v0=v)
            fd = central(f)
            assert abs(gv[i, k] - fd) <= 1e-8 * np.abs(gv).max() + 1e-6 * abs(fd), (i, k, gv[i, k], fd)


def test_fixed_node_velocity_gradient_matches_finite_differences(problem):
    # a fixed node keeps its velocity; it still enters b through -dt^2 K v
    p, (_, _, gv, _) = problem
    i = int(np.flatnonzero(p.fixed)[1])
    def f(d):
        v = p.v0.copy()
        v[i, 2] += d
        return p.loss(v0=v)
    fd = central(f)
    assert abs(gv[i, 2] - fd) <= 1e-8 * np.abs(gv).max() + 1e-6 * abs(fd)


def test_initial_position_gradient_matches_finite_differences(problem):
    p, (_, gx, _, _) = problem
    rng = np.random.default_rng(11)
    for i in rng.choice(np.flatnonzero(p.fixed == 0), 4, replace=False):
        for k in range(3):
            def f(d, i=i, k=k):
                x = p.x0.copy()
                x[i, k] += d
                return p.loss(x0=x)
            fd = central(f)
            assert abs(gx[i, k] - fd) <= 1e-8 * np.abs(gx).max() + 1e-6 * abs(fd), (i, k, gx[i, k], fd)


def test_zero_seed_gives_zero_gradient():
    p = Problem(steps=2)
    o = p.world()
    gE, gx, gv, _ = o.adjoint(p.dt, p.steps, np.zeros((p.n, 3)), np.zeros((p.n, 3)))
    assert not gE.any() and not gx.any() and not gv.any()


def test_one_step_velocity_seed_is_the_solve_transpose():
    # one substep, no stiffness change with v: vbar seeds u = vbar; with E -> 0
    # and no damping A = M, so dL/dv0 = vbar on dynamic nodes exactly
    p = Problem(steps=1, damping=0.0)
    o = Oracle(64, p.rest, p.tets, 0.0, 0.3, 0.0, 1000.0)
    o.set_params((0.0, 0.0, -9.81), 1, 1e-12, 1000)
    o.set_fixed(p.fixed)
    o.set_state(p.x0, p.v0)
    vbar = np.random.default_rng(12).standard_normal((p.n, 3))
    gE, gx, gv, _ = o.adjoint(p.dt, 1, np.zeros((p.n, 3)), vbar)
    free = p.fixed == 0
    assert np.allclose(gv[free], vbar[free], rtol=1e-12, atol=1e-14)
    assert not gx.any() and not gE.any()


class ContactProblem(Problem):
    """Problem with contact forces in every substep: two point connections
    (the far face's barycentre to a fixed target, and one node to a second
    target) and/or a collision sphere under the bottom far corner."""

    def __init__(self, connections=True, sphere=True, steps=4):
        super().__init__(steps=steps)
        far = self.end
        self.conns = []
        if connections:
            self.conns = [(far[:4], np.full(4, 0.25), self.x0[far[:4]].mean(0) + np.array([0.03, -0.02, 0.04]), 80.0,
                           0.7),
                          (np.array([far[-1]] * 4), np.array([1.0, 0.0, 0.0, 0.0]),
                           self.x0[far[-1]] + np.array([0.0, 0.05, 0.0]), 40.0, 0.2)]
        self.sphere = None
        if sphere:
            lo, hi = self.x0.min(0), self.x0.max(0)
            self.sphere = (np.array([hi[0], 0.5 * (lo[1] + hi[1]), lo[2] - 0.02]), 0.13, 300.0, 1.5)

    def world(self, **kw):
        o = super().world(**kw)
        if self.sphere is not None:
            c, r, k, d = self.sphere
            o.set_spheres(c[None], [r], [k], [d])
        return o

    def add_conns(self, o):
        for n, wts, t, k, c in self.conns:
            o.add_point_connection(np.asarray(n), wts, t, k, c)

    def loss(self, **kw):
        o = self.world(**kw)
        for _ in range(self.steps):
            self.add_conns(o)  # a step clears its frame's connections
            assert o.step(self.dt) == 0
        return self.loss_terms(*o.state())

    def adjoint(self):
        probe = self.world()
        for _ in range(self.steps):
            self.add_conns(probe)
            probe.step(self.dt)
            if self.sphere is not None:
                assert probe.last_contact_count() > 0
        xs, vs = probe.state()
        o = self.world()
        self.add_conns(o)  # the recorded substeps keep the connections
        xbar = np.zeros((self.n, 3))
        xbar[self.end] = 2 * (xs[self.end] - self.target)
        vbar = 0.2 * vs
        gE, gx, gv, its = o.adjoint(self.dt, self.steps, xbar, vbar, tol=1e-12, max_iter=1000)
        assert np.array_equal(o.state()[0], xs)
        return gE, gx, gv, its


@pytest.fixture(scope="module", params=["connections", "sphere", "both"])
def contact_problem(request):
    p = ContactProblem(connections=request.param != "sphere", sphere=request.param != "connections")
    return p, p.adjoint()


def test_contact_gradients_match_finite_differences(contact_problem):
    """dL/dlogE, dL/dv0 and dL/dx0 with point connections and sphere
    contacts in the forward map (the sphere's target moves with x)."""
    p, (gE, gx, gv, its) = contact_problem
    assert (its > 0).all()
    rng = np.random.default_rng(13)
    free = np.flatnonzero(p.fixed == 0)
    picks = [("E", e, 0) for e in rng.choice(p.t, 4, replace=False)]
    # the nodes the contacts pull on, and a few others
    near = list(p.end[:2]) + list(rng.choice(free, 2, replace=False))
    picks += [(kind, i, k) for kind in ("v", "x") for i in near for k in range(3)]
    for kind, i, k in picks:
        def f(d, kind=kind, i=i, k=k):
            if kind == "E":
                le = p.logE.copy()
                le[i] += d
                return p.loss(logE=le)
            a = (p.v0 if kind == "v" else p.x0).copy()
            a[i, k] += d
            return p.loss(**{("v0" if kind == "v" else "x0"): a})
        fd = central(f)
        g = {"E": gE, "v": gv, "x": gx}[kind]
        got = g[i] if kind == "E" else g[i, k]
        assert abs(got - fd) <= 1e-8 * np.abs(g).max() + 1e-6 * abs(fd), (kind, i, k, got, fd)


def test_contact_terms_change_the_gradient():
    """The contact terms are not negligible here: dropping the connections
    and the sphere changes dL/dx0 far beyond the tolerance above."""
    with_c = ContactProblem().adjoint()
    without = Problem().adjoint()
    assert np.abs(with_c[1] - without[1]).max() > 1e-3 * np.abs(without[1]).max()
