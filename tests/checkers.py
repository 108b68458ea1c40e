"""ctypes bindings for the checker libraries (test infrastructure).

* ``Oracle``   -> oracle/liboracle.so, the CPU restatement (float or double).
* ``Reference``-> oracle/_ref/libphysik_ref.so, the reference's own sources
  compiled in place plus oracle/ref_driver.cpp. It exists only where the
  reference tree was available at build time; tests that need it skip
  otherwise.

Both take float64 numpy arrays at the boundary (the reference binding converts
to float32, exactly as its FP32 ABI does).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libphysik_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _dp, _dp, _dp, _dp]
        lib.orc_destroy.argtypes = [C.c_void_p]
        lib.orc_set_params.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                       C.c_int, C.c_int]
        lib.orc_set_fixed.argtypes = [C.c_void_p, _bp]
        lib.orc_set_tet_active.argtypes = [C.c_void_p, _bp]
        lib.orc_set_node_active.argtypes = [C.c_void_p, _bp]
        lib.orc_set_young.argtypes = [C.c_void_p, _dp]
        lib.orc_set_fem_model.argtypes = [C.c_void_p, C.c_int]
        lib.orc_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_get_state.argtypes = [C.c_void_p, _dp, _dp]
        lib.orc_step.argtypes = [C.c_void_p, C.c_double]
        lib.orc_step.restype = C.c_int
        lib.orc_get_cg.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]
        lib.orc_assemble.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.orc_assemble.restype = C.c_int
        lib.orc_get_system.argtypes = [C.c_void_p, _ip, _ip, _ip, _dp, _dp, _dp]
        lib.orc_get_scatter.argtypes = [C.c_void_p, _ip]
        lib.orc_get_elements.argtypes = [C.c_void_p, _bp, _dp, _dp, _dp]
        lib.orc_get_rest.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.orc_polar.argtypes = [C.c_int, _dp, _dp]
        lib.orc_polar.restype = C.c_int
        lib.orc_spmv.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp]
        lib.orc_pcg.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_void_p, C.c_int, C.c_double, _dp,
                                C.POINTER(C.c_double), C.POINTER(C.c_int)]
        lib.orc_pcg.restype = C.c_int
        lib.orc_adjoint.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp,
                                    _ip]
        lib.orc_adjoint.restype = C.c_int
        lib.orc_linear_solve.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_int, C.c_double,
                                         C.c_int, _dp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        lib.orc_linear_solve.restype = C.c_int
        lib.orc_polar_vjp.argtypes = [_dp, _dp, _dp, _dp]
        lib.orc_polar_vjp.restype = C.c_int
        lib.orc_add_point_connection.argtypes = [C.c_void_p, _ip, _dp, _dp, C.c_double, C.c_double]
        lib.orc_point_connection_count.argtypes = [C.c_void_p]
        lib.orc_point_connection_count.restype = C.c_int
        lib.orc_set_spheres.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
        lib.orc_last_contact_count.argtypes = [C.c_void_p]
        lib.orc_last_contact_count.restype = C.c_int
        _oracle = lib
    return _oracle


def reference_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref_create.restype = C.c_void_p
        lib.ref_create.argtypes = [_fp, C.c_int, _ip, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.ref_counts.restype = C.c_int
        lib.ref_set_params.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, C.c_int,
                                       C.c_int]
        lib.ref_set_fixed.argtypes = [C.c_void_p, _bp]
        lib.ref_set_tet_active.argtypes = [C.c_void_p, _bp]
        lib.ref_set_young.argtypes = [C.c_void_p, _fp]
        lib.ref_set_fem_model.argtypes = [C.c_void_p, C.c_int]
        lib.ref_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_get_state.argtypes = [C.c_void_p, _fp, _fp]
        lib.ref_step.argtypes = [C.c_void_p, C.c_float]
        lib.ref_step.restype = C.c_int
        lib.ref_get_cg.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_int)]
        lib.ref_assemble.argtypes = [C.c_void_p, C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.ref_assemble.restype = C.c_int
        lib.ref_get_system.argtypes = [C.c_void_p, _ip, _ip, _ip, _fp, _fp, _fp]
        lib.ref_get_scatter.argtypes = [C.c_void_p, _ip]
        lib.ref_get_elements.argtypes = [C.c_void_p, _bp, _fp, _fp]
        lib.ref_get_rest.argtypes = [C.c_void_p, _fp, _fp, _fp]
        lib.ref_spmv.argtypes = [C.c_int, _ip, _ip, _fp, _fp, _fp]
        lib.ref_linear_solve.argtypes = [C.c_int, _ip, _ip, _fp, _fp, C.c_int, C.c_float, C.c_int, _fp,
                                         C.POINTER(C.c_float), C.POINTER(C.c_int)]
        lib.ref_linear_solve.restype = C.c_int
        lib.ref_add_point_connection.argtypes = [C.c_void_p, _ip, _fp, _fp, C.c_float, C.c_float]
        lib.ref_point_connection_count.argtypes = [C.c_void_p]
        lib.ref_point_connection_count.restype = C.c_int
        lib.ref_add_sphere.argtypes = [C.c_void_p] + [C.c_float] * 6
        lib.ref_add_sphere.restype = C.c_int
        lib.ref_sphere_settings.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.ref_move_sphere.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]
        lib.ref_sphere_overlaps.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _fp]
        lib.ref_sphere_overlaps.restype = C.c_int
        lib.ref_sphere_overlap_count.argtypes = [C.c_void_p, C.c_int]
        lib.ref_sphere_overlap_count.restype = C.c_int
        _ref = lib
    return _ref


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Oracle:
    """One oracle world over a single tet mesh (world node ids)."""

    def __init__(self, precision, rest, tets, young, poisson, damping, density):
        self.lib = oracle_lib()
        self.n = int(rest.shape[0])
        self.t = int(tets.shape[0])
        t = self.t

        def per_tet(v):
            a = np.asarray(v, dtype=np.float64)
            return _c(np.broadcast_to(a, (t,)), np.float64)

        self.h = self.lib.orc_create(precision, self.n, _c(rest, np.float64).reshape(-1), t,
                                     _c(tets, np.int32).reshape(-1), per_tet(young), per_tet(poisson),
                                     per_tet(damping), per_tet(density))

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, gravity=(0.0, 0.0, 0.0), substeps=1, tol=1e-4, max_iter=128, jacobi=True, implicit=True):
        self.lib.orc_set_params(self.h, gravity[0], gravity[1], gravity[2], substeps, tol, max_iter, int(jacobi),
                                int(implicit))

    def set_fixed(self, fixed):
        self.lib.orc_set_fixed(self.h, _c(fixed, np.uint8))

    def set_tet_active(self, active):
        self.lib.orc_set_tet_active(self.h, _c(active, np.uint8))

    def set_node_active(self, active):
        self.lib.orc_set_node_active(self.h, _c(active, np.uint8))

    def set_young(self, young):
        self.lib.orc_set_young(self.h, _c(young, np.float64))

    def set_fem_model(self, model):
        """0 Linear, 1 Corotational (default), 2 NeoHookean (FemModel, FEMModel.h:40-45)."""
        self.lib.orc_set_fem_model(self.h, int(model))

    def set_state(self, x=None, v=None):
        xa = None if x is None else _c(x, np.float64)
        va = None if v is None else _c(v, np.float64)
        self.lib.orc_set_state(self.h, None if xa is None else xa.ctypes.data,
                               None if va is None else va.ctypes.data)

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        self.lib.orc_get_state(self.h, x, v)
        return x, v

    def step(self, dt):
        return self.lib.orc_step(self.h, dt)

    def add_point_connection(self, nodes, weights, target, stiffness, damping):
        """A spring from the barycentric point of four nodes to a target for
        every substep of the next step (PHYSIK_AddPointConnection)."""
        self.lib.orc_add_point_connection(self.h, _c(nodes, np.int32), _c(weights, np.float64),
                                          _c(target, np.float64), float(stiffness), float(damping))

    def point_connection_count(self):
        return self.lib.orc_point_connection_count(self.h)

    def last_contact_count(self):
        """Sphere contacts generated by the last substep's assembly."""
        return self.lib.orc_last_contact_count(self.h)

    def set_spheres(self, centers=(), radii=(), stiffness=(), damping=()):
        """Collision spheres in creation order (PHYSIK_CreateCollisionSphereComponent
        + PHYSIK_SetCollisionSphereConnectionSettings); none removes them."""
        n = len(radii)
        c = _c(np.reshape(centers, (n, 3)) if n else np.zeros((1, 3)), np.float64)
        r, k, d = (_c(a if n else [0.0], np.float64) for a in (radii, stiffness, damping))
        self.lib.orc_set_spheres(self.h, n, c, r, k, d)

    def cg(self):
        it, res, conv = C.c_int(), C.c_double(), C.c_int()
        self.lib.orc_get_cg(self.h, C.byref(it), C.byref(res), C.byref(conv))
        return it.value, res.value, bool(conv.value)

    def assemble(self, dt):
        nd, nnz = C.c_int(), C.c_int()
        rc = self.lib.orc_assemble(self.h, dt, C.byref(nd), C.byref(nnz))
        nd, nnz = nd.value, nnz.value
        dyn = np.zeros(self.n, np.int32)
        rs = np.zeros(max(nd, 0) + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        vals = np.zeros(max(nnz, 1) * 9)
        rhs = np.zeros(max(nd, 1) * 3)
        mass = np.zeros(self.n)
        if nd > 0:
            self.lib.orc_get_system(self.h, dyn, rs, ci, vals, rhs, mass)
        return dict(rc=rc, nd=nd, nnz=nnz, dyn=dyn, row_start=rs, col=ci[:nnz], vals=vals[:9 * nnz].reshape(-1, 3, 3),
                    rhs=rhs[:3 * nd].reshape(-1, 3), mass=mass)

    def scatter(self):
        s = np.zeros(self.t * 16, np.int32)
        self.lib.orc_get_scatter(self.h, s)
        return s.reshape(self.t, 4, 4)

    def elements(self):
        valid = np.zeros(self.t, np.uint8)
        R = np.zeros(self.t * 9)
        f = np.zeros(self.t * 12)
        K = np.zeros(self.t * 144)
        self.lib.orc_get_elements(self.h, valid, R, f, K)
        return valid.astype(bool), R.reshape(-1, 3, 3), f.reshape(-1, 4, 3), K.reshape(-1, 4, 4, 3, 3)

    def adjoint(self, dt, steps, xbar, vbar=None, tol=1e-10, max_iter=10000):
        """Runs `steps` substeps from the current state and back-propagates the
        final-state adjoints. Returns (dL/dlogE, dL/dx0, dL/dv0, adjoint CG
        iterations per substep); the world is left at the final state."""
        vbar = np.zeros((self.n, 3)) if vbar is None else vbar
        gE = np.zeros(self.t)
        gx = np.zeros(self.n * 3)
        gv = np.zeros(self.n * 3)
        its = np.zeros(max(steps, 1), np.int32)
        rc = self.lib.orc_adjoint(self.h, dt, steps, _c(xbar, np.float64).reshape(-1),
                                  _c(vbar, np.float64).reshape(-1), tol, max_iter, gE, gx, gv, its)
        if rc != 0:
            raise RuntimeError(f"orc_adjoint failed with {rc}")
        return gE, gx.reshape(-1, 3), gv.reshape(-1, 3), its[:steps]

    def rest(self):
        vol = np.zeros(self.t)
        dminv = np.zeros(self.t * 9)
        grad = np.zeros(self.t * 12)
        self.lib.orc_get_rest(self.h, vol, dminv, grad)
        return vol, dminv.reshape(-1, 3, 3), grad.reshape(-1, 4, 3)


def oracle_polar(F, precision=64):
    R = np.zeros(9)
    it = oracle_lib().orc_polar(precision, _c(F, np.float64).reshape(-1), R)
    return R.reshape(3, 3), it


def oracle_polar_vjp(F, G):
    """(R, dL/dF, kind) of the oracle polar rotation for G = dL/dR."""
    R = np.zeros(9)
    Fb = np.zeros(9)
    kind = oracle_lib().orc_polar_vjp(_c(F, np.float64).reshape(-1), _c(G, np.float64).reshape(-1), R, Fb)
    return R.reshape(3, 3), Fb.reshape(3, 3), kind


def oracle_spmv(row_start, col, vals, x, precision=64):
    n = len(row_start) - 1
    y = np.zeros(n * 3)
    oracle_lib().orc_spmv(precision, n, len(col), _c(row_start, np.int32), _c(col, np.int32),
                          _c(vals, np.float64).reshape(-1), _c(x, np.float64).reshape(-1), y)
    return y.reshape(-1, 3)


def oracle_pcg(row_start, col, vals, b, minv, max_it, tol, precision=64):
    n = len(row_start) - 1
    x = np.zeros(n * 3)
    res = C.c_double()
    conv = C.c_int()
    mv = None if minv is None else _c(minv, np.float64).reshape(-1)
    it = oracle_lib().orc_pcg(precision, n, len(col), _c(row_start, np.int32), _c(col, np.int32),
                              _c(vals, np.float64).reshape(-1), _c(b, np.float64).reshape(-1),
                              None if mv is None else mv.ctypes.data, max_it, tol, x, C.byref(res), C.byref(conv))
    return x.reshape(-1, 3), it, res.value, bool(conv.value)


def oracle_linear_solve(row_start, col, vals, b, max_it, tol, jacobi=True, precision=64):
    """LinearSolver::Solve restated (block-Jacobi PCG on a symmetric-upper
    BSR): (x or None when the solution was cleared, iterations, residual,
    converged)."""
    n = len(row_start) - 1
    x = np.zeros(max(n, 1) * 3)
    res = C.c_double()
    conv = C.c_int()
    it = oracle_lib().orc_linear_solve(precision, n, len(col), _c(row_start, np.int32), _c(col, np.int32),
                                       _c(vals, np.float64).reshape(-1), _c(b, np.float64).reshape(-1), int(max_it),
                                       float(tol), int(jacobi), x, C.byref(res), C.byref(conv))
    sol = x[:3 * n].reshape(-1, 3) if it >= 0 else None
    return sol, (it if it >= 0 else -1 - it), res.value, bool(conv.value)


def ref_linear_solve(row_start, col, vals, b, max_it, tol, jacobi=True):
    """The reference's own GetCurrentLinearSolver().Solve (FP32)."""
    n = len(row_start) - 1
    x = np.zeros(max(n, 1) * 3, np.float32)
    res = C.c_float()
    conv = C.c_int()
    it = ref_lib().ref_linear_solve(n, _c(row_start, np.int32), _c(col, np.int32), _c(vals, np.float32).reshape(-1),
                                    _c(b, np.float32).reshape(-1), int(max_it), float(tol), int(jacobi), x,
                                    C.byref(res), C.byref(conv))
    sol = x[:3 * n].reshape(-1, 3) if it >= 0 else None
    return sol, (it if it >= 0 else -1 - it), res.value, bool(conv.value)


class Reference:
    """The compiled reference driven through its own C ABI (FP32)."""

    def __init__(self, rest, tets, density, young, poisson, damping):
        self.lib = ref_lib()
        self.h = self.lib.ref_create(_c(rest, np.float32).reshape(-1), int(rest.shape[0]),
                                     _c(tets, np.int32).reshape(-1), int(tets.shape[0]), density, young, poisson,
                                     damping)
        n, t = C.c_int(), C.c_int()
        self.lib.ref_counts(self.h, C.byref(n), C.byref(t))
        self.n, self.t = n.value, t.value

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, gravity=(0.0, 0.0, 0.0), substeps=1, tol=1e-4, max_iter=128, implicit=True):
        self.lib.ref_set_params(self.h, gravity[0], gravity[1], gravity[2], substeps, tol, max_iter, int(implicit))

    def set_fixed(self, fixed):
        self.lib.ref_set_fixed(self.h, _c(fixed, np.uint8))

    def set_tet_active(self, active):
        self.lib.ref_set_tet_active(self.h, _c(active, np.uint8))

    def set_young(self, young):
        self.lib.ref_set_young(self.h, _c(young, np.float32))

    def set_fem_model(self, model):
        self.lib.ref_set_fem_model(self.h, int(model))

    def set_state(self, x=None, v=None):
        xa = None if x is None else _c(x, np.float32)
        va = None if v is None else _c(v, np.float32)
        self.lib.ref_set_state(self.h, None if xa is None else xa.ctypes.data,
                               None if va is None else va.ctypes.data)

    def state(self):
        x = np.zeros((self.n, 3), np.float32)
        v = np.zeros((self.n, 3), np.float32)
        self.lib.ref_get_state(self.h, x, v)
        return x, v

    def step(self, dt):
        return self.lib.ref_step(self.h, dt)

    def add_point_connection(self, nodes, weights, target, stiffness, damping):
        self.lib.ref_add_point_connection(self.h, _c(nodes, np.int32), _c(weights, np.float32),
                                          _c(target, np.float32), float(stiffness), float(damping))

    def point_connection_count(self):
        return self.lib.ref_point_connection_count(self.h)

    def add_sphere(self, center, radius, stiffness=1000.0, damping=10.0):
        return self.lib.ref_add_sphere(self.h, *(float(c) for c in center), radius, stiffness, damping)

    def sphere_settings(self, sphere):
        k, d = C.c_float(), C.c_float()
        self.lib.ref_sphere_settings(self.h, sphere, C.byref(k), C.byref(d))
        return k.value, d.value

    def move_sphere(self, sphere, center):
        self.lib.ref_move_sphere(self.h, sphere, *(float(c) for c in center))

    def sphere_overlaps(self, sphere, max_overlaps=None):
        """(ints (n, 8): type, tet, nodes 0-3, node mask, node count;
        floats (n, 5): centre, radius, minimum node distance)."""
        if max_overlaps is None:
            max_overlaps = self.lib.ref_sphere_overlap_count(self.h, sphere)
        ints = np.zeros((max(1, max_overlaps), 8), np.int32)
        floats = np.zeros((max(1, max_overlaps), 5), np.float32)
        n = self.lib.ref_sphere_overlaps(self.h, sphere, max_overlaps, ints, floats)
        return ints[:n], floats[:n]

    def cg(self):
        it, res, conv = C.c_int(), C.c_float(), C.c_int()
        self.lib.ref_get_cg(self.h, C.byref(it), C.byref(res), C.byref(conv))
        return it.value, res.value, bool(conv.value)

    def assemble(self, dt):
        nd, nnz = C.c_int(), C.c_int()
        rc = self.lib.ref_assemble(self.h, dt, C.byref(nd), C.byref(nnz))
        nd, nnz = nd.value, nnz.value
        dyn = np.zeros(self.n, np.int32)
        rs = np.zeros(nd + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        vals = np.zeros(max(nnz, 1) * 9, np.float32)
        rhs = np.zeros(max(nd, 1) * 3, np.float32)
        mass = np.zeros(self.n, np.float32)
        self.lib.ref_get_system(self.h, dyn, rs, ci, vals, rhs, mass)
        return dict(rc=rc, nd=nd, nnz=nnz, dyn=dyn, row_start=rs, col=ci[:nnz], vals=vals[:9 * nnz].reshape(-1, 3, 3),
                    rhs=rhs[:3 * nd].reshape(-1, 3), mass=mass)

    def scatter(self):
        s = np.zeros(self.t * 16, np.int32)
        self.lib.ref_get_scatter(self.h, s)
        return s.reshape(self.t, 4, 4)

    def elements(self):
        valid = np.zeros(self.t, np.uint8)
        f = np.zeros(self.t * 12, np.float32)
        K = np.zeros(self.t * 144, np.float32)
        self.lib.ref_get_elements(self.h, valid, f, K)
        return valid.astype(bool), f.reshape(-1, 4, 3), K.reshape(-1, 4, 4, 3, 3)

    def rest(self):
        vol = np.zeros(self.t, np.float32)
        dminv = np.zeros(self.t * 9, np.float32)
        grad = np.zeros(self.t * 12, np.float32)
        self.lib.ref_get_rest(self.h, vol, dminv, grad)
        return vol, dminv.reshape(-1, 3, 3), grad.reshape(-1, 4, 3)


def ref_spmv(row_start, col, vals, x):
    n = len(row_start) - 1
    y = np.zeros(n * 3, np.float32)
    ref_lib().ref_spmv(n, _c(row_start, np.int32), _c(col, np.int32), _c(vals, np.float32).reshape(-1),
                       _c(x, np.float32).reshape(-1), y)
    return y.reshape(-1, 3)
