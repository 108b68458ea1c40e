"""Python handle on the GPU system (include/physik_amd.h).

``System`` owns a pk_system: one mesh (or a batch of bodies, one CG group per
body) resident on the GPU, stepped with pk_system_step. ``Pattern`` builds the
same dynamic numbering and BSR patterns on the host only (no GPU needed).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N


class _Desc:
    """Keeps the numpy arrays behind a MeshDesc alive."""

    def __init__(self, rest, tets, young, poisson, damping, density, tet_active=None, node_fixed=None,
                 node_active=None, node_group=None, extra_pairs=None):
        self.rest = N.as_f64(rest).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        t = self.tets.shape[0]

        def per_tet(v):
            return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (t,)))

        self.young, self.poisson = per_tet(young), per_tet(poisson)
        self.damping, self.density = per_tet(damping), per_tet(density)
        u8 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.uint8)  # noqa: E731
        self.tet_active, self.node_fixed, self.node_active = u8(tet_active), u8(node_fixed), u8(node_active)
        self.node_group = None if node_group is None else np.ascontiguousarray(node_group, dtype=np.int32)
        self.extra_pairs = (None if extra_pairs is None or len(extra_pairs) == 0
                            else np.ascontiguousarray(extra_pairs, dtype=np.int32).reshape(-1, 2))
        n_extra = 0 if self.extra_pairs is None else self.extra_pairs.shape[0]
        self.c = N.MeshDesc(self.rest.shape[0], N.ptr(self.rest), t, N.ptr(self.tets), N.ptr(self.young),
                            N.ptr(self.poisson), N.ptr(self.damping), N.ptr(self.density), N.ptr(self.tet_active),
                            N.ptr(self.node_fixed), N.ptr(self.node_active), N.ptr(self.node_group), n_extra,
                            N.ptr(self.extra_pairs))


class Pattern:
    def __init__(self, rest, tets, young=1.0, poisson=0.3, damping=0.0, density=1.0, **flags):
        lib = N.lib()
        d = _Desc(rest, tets, young, poisson, damping, density, **flags)
        h = C.c_void_p()
        N.check(lib.pk_pattern_create(C.byref(d.c), C.byref(h)), "pk_pattern_create")
        try:
            nd, nu, nf, nc = C.c_int(), C.c_longlong(), C.c_longlong(), C.c_longlong()
            lib.pk_pattern_sizes(h, C.byref(nd), C.byref(nu), C.byref(nf), C.byref(nc))
            self.nd, self.nnzb_upper, self.nnzb_full, self.contributions = nd.value, nu.value, nf.value, nc.value
            n, t = d.rest.shape[0], d.tets.shape[0]
            self.dyn = np.zeros(n, np.int32)
            self.upper_start = np.zeros(self.nd + 1, np.int32)
            self.upper_col = np.zeros(max(self.nnzb_upper, 1), np.int32)
            self.full_start = np.zeros(self.nd + 1, np.int32)
            self.full_col = np.zeros(max(self.nnzb_full, 1), np.int32)
            self.scatter = np.zeros(max(t * 16, 1), np.int32)
            lib.pk_pattern_copy(h, N.ptr(self.dyn), N.ptr(self.upper_start), N.ptr(self.upper_col),
                                N.ptr(self.full_start), N.ptr(self.full_col), N.ptr(self.scatter))
            self.upper_col = self.upper_col[:self.nnzb_upper]
            self.full_col = self.full_col[:self.nnzb_full]
            self.scatter = self.scatter[:t * 16].reshape(t, 4, 4)
        finally:
            lib.pk_pattern_destroy(h)


class System:
    """A mesh (or a batch of bodies) resident on the GPU."""

    def __init__(self, rest, tets, young, poisson, damping, density, precision=N.PK_FP64, **flags):
        self.lib = N.lib()
        self._d = _Desc(rest, tets, young, poisson, damping, density, **flags)
        h = C.c_void_p()
        N.check(self.lib.pk_system_create(C.byref(self._d.c), precision, C.byref(h)), "pk_system_create")
        self.h = h
        self.precision = precision
        n, t, g, nd = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        nu, nf = C.c_longlong(), C.c_longlong()
        self.lib.pk_system_sizes(h, C.byref(n), C.byref(t), C.byref(g), C.byref(nd), C.byref(nu), C.byref(nf))
        self.n, self.t, self.groups, self.nd = n.value, t.value, g.value, nd.value
        self.nnzb_upper, self.nnzb_full = nu.value, nf.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.pk_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, lanes_per_row=16, spmv_rows=128, vector_rows=1024, unroll=2):
        N.check(self.lib.pk_system_tune(self.h, lanes_per_row, spmv_rows, vector_rows, unroll), "pk_system_tune")

    def autotune(self, dt, cache=None, verbose=False):
        """Measure and apply the fastest CG launch shape (physik_amd.autotune)."""
        from .autotune import autotune
        return autotune(self, dt, cache=cache, verbose=verbose)

    def set_staging(self, mode=1):
        """SpMV staging of d in LDS: 0 off, 1 on (identical results)."""
        N.check(self.lib.pk_system_set_spmv_staging(self.h, int(mode)), "pk_system_set_spmv_staging")

    def get_staging(self):
        """(mode, rows a workgroup stages, segments in use)."""
        m, r, g = C.c_int(), C.c_int(), C.c_int()
        N.check(self.lib.pk_system_get_spmv_staging(self.h, C.byref(m), C.byref(r), C.byref(g)),
                "pk_system_get_spmv_staging")
        return m.value, r.value, bool(g.value)

    def stage_capacity(self):
        """Rows of d an SpMV workgroup may stage (0 with staging off)."""
        r = C.c_int()
        N.check(self.lib.pk_system_get_spmv_stage_capacity(self.h, C.byref(r)), "pk_system_get_spmv_stage_capacity")
        return r.value

    def slice_layout(self):
        """(slices, compact slices, elements) of sliced upper-only storage;
        zeros under full storage."""
        n, c, e = C.c_int(), C.c_int(), C.c_longlong()
        N.check(self.lib.pk_system_get_slice_layout(self.h, C.byref(n), C.byref(c), C.byref(e)),
                "pk_system_get_slice_layout")
        return n.value, c.value, e.value

    def get_tune(self):
        v = [C.c_int() for _ in range(4)]
        N.check(self.lib.pk_system_get_tune(self.h, *[C.byref(x) for x in v]), "pk_system_get_tune")
        return tuple(x.value for x in v)

    def set_state(self, x=None, v=None):
        xa = None if x is None else N.as_f64(x).reshape(-1)
        va = None if v is None else N.as_f64(v).reshape(-1)
        N.check(self.lib.pk_system_set_state(self.h, N.ptr(xa), N.ptr(va)), "pk_system_set_state")

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        N.check(self.lib.pk_system_get_state(self.h, N.ptr(x), N.ptr(v)), "pk_system_get_state")
        return x, v

    def set_state_device(self, x_ptr=None, v_ptr=None):
        """Node state from device memory: x_ptr, v_ptr are addresses of
        node_count x 3 float64 arrays on the system's device (or None)."""
        N.check(self.lib.pk_system_set_state_device(self.h, x_ptr, v_ptr), "pk_system_set_state_device")

    def get_state_device(self, x_ptr=None, v_ptr=None):
        """Node state into device memory (addresses as set_state_device)."""
        N.check(self.lib.pk_system_get_state_device(self.h, x_ptr, v_ptr), "pk_system_get_state_device")

    def tape_id(self):
        """Id of the recorded run backward() would reverse (0: none)."""
        v = C.c_longlong()
        N.check(self.lib.pk_system_tape_id(self.h, C.byref(v)), "pk_system_tape_id")
        return v.value

    def export_begin(self, positions=True, velocities=True):
        """Starts the float frame export (pk_system_export_begin); returns at once."""
        what = (N.PK_EXPORT_POSITIONS if positions else 0) | (N.PK_EXPORT_VELOCITIES if velocities else 0)
        N.check(self.lib.pk_system_export_begin(self.h, what), "pk_system_export_begin")

    def export_read(self, first=0, count=None, positions=True, velocities=True):
        """Waits for the frame export and returns float32 (count, 3) arrays
        (None for a quantity not asked for); raises on PK_ERR_STATE."""
        count = self.n - first if count is None else count
        x = np.zeros((count, 3), dtype=np.float32) if positions else None
        v = np.zeros((count, 3), dtype=np.float32) if velocities else None
        N.check(self.lib.pk_system_export_read(self.h, first, count, N.ptr(x), N.ptr(v)), "pk_system_export_read")
        return x, v

    def export_status(self, first=0, count=None, positions=True, velocities=True):
        """Status of pk_system_export_read for the same arguments (no raise)."""
        count = self.n - first if count is None else count
        x = np.zeros((max(count, 1), 3), dtype=np.float32) if positions else None
        v = np.zeros((max(count, 1), 3), dtype=np.float32) if velocities else None
        return self.lib.pk_system_export_read(self.h, first, count, N.ptr(x), N.ptr(v))

    def set_gravity(self, g):
        N.check(self.lib.pk_system_set_gravity(self.h, *[float(a) for a in g]), "pk_system_set_gravity")

    def set_solver(self, tol, max_iter, jacobi=True):
        N.check(self.lib.pk_system_set_solver(self.h, tol, max_iter, int(jacobi)), "pk_system_set_solver")

    def set_mode(self, implicit=True):
        N.check(self.lib.pk_system_set_mode(self.h, int(implicit)), "pk_system_set_mode")

    def set_fem_models(self, models):
        """FEM model per tet (or one for all): 0 linear, 1 corotational, 2 none."""
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(models, np.uint8), (self.t,)))
        N.check(self.lib.pk_system_set_fem_models(self.h, N.ptr(m)), "pk_system_set_fem_models")

    def set_young(self, young):
        y = np.ascontiguousarray(np.broadcast_to(np.asarray(young, np.float64), (self.t,)))
        N.check(self.lib.pk_system_set_young(self.h, y), "pk_system_set_young")

    def step(self, dt, substeps=1, frames=1):
        rc = self.lib.pk_system_step(self.h, dt, substeps, frames)
        if rc == N.PK_ERR_MASS:
            raise RuntimeError(N.last_error())
        N.check(rc, "pk_system_step")

    def results(self):
        g = self.groups
        it = np.zeros(g, np.int32)
        res = np.zeros(g)
        conv = np.zeros(g, np.int32)
        st = np.zeros(g, np.int32)
        N.check(self.lib.pk_system_group_results(self.h, N.ptr(it), N.ptr(res), N.ptr(conv), N.ptr(st)),
                "pk_system_group_results")
        return it, res, conv.astype(bool), st

    def profile(self, enable=True):
        N.check(self.lib.pk_system_profile(self.h, int(enable)), "pk_system_profile")

    def kernel_time(self, kernel):
        n, ms = C.c_longlong(), C.c_double()
        N.check(self.lib.pk_system_kernel_time(self.h, kernel, C.byref(n), C.byref(ms)), "pk_system_kernel_time")
        return n.value, ms.value

    def set_connections(self, nodes=None, weights=None, targets=None, stiffness=None, damping=None):
        """Point springs acting in every following substep (PHYSIK_AddPointConnection
        semantics); no arguments clears them. nodes (C, 4) ints, weights (C, 4),
        targets (C, 3), stiffness (C,), damping (C,)."""
        if nodes is None or len(nodes) == 0:
            N.check(self.lib.pk_system_set_connections(self.h, 0, None, None, None, None, None),
                    "pk_system_set_connections")
            return
        nd = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1, 4)
        c = nd.shape[0]
        w = N.as_f64(weights).reshape(c, 4)
        tg = N.as_f64(targets).reshape(c, 3)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(stiffness, np.float64), (c,)))
        dm = np.ascontiguousarray(np.broadcast_to(np.asarray(damping, np.float64), (c,)))
        N.check(self.lib.pk_system_set_connections(self.h, c, N.ptr(nd), N.ptr(w), N.ptr(tg), N.ptr(k), N.ptr(dm)),
                "pk_system_set_connections")

    def set_spheres(self, centers=None, radii=None, stiffness=1000.0, damping=10.0):
        """Collision spheres acting in every following substep (contacts from
        the tets' centroids, CollisionSphereComponent semantics); no arguments
        removes them. centers (S, 3), radii (S,), stiffness and damping (S,) or
        scalars."""
        if radii is None or len(np.atleast_1d(radii)) == 0:
            N.check(self.lib.pk_system_set_spheres(self.h, 0, None, None, None, None), "pk_system_set_spheres")
            return
        r = N.as_f64(np.atleast_1d(radii))
        n = r.shape[0]
        c = N.as_f64(centers).reshape(n, 3)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(stiffness, np.float64), (n,)))
        dm = np.ascontiguousarray(np.broadcast_to(np.asarray(damping, np.float64), (n,)))
        N.check(self.lib.pk_system_set_spheres(self.h, n, N.ptr(c), N.ptr(r), N.ptr(k), N.ptr(dm)),
                "pk_system_set_spheres")

    def set_tets_active(self, tets, active):
        """Switches tets on or off in place. Returns False, with nothing
        changed, when the switch would change the dynamic nodes (or turn on a
        tet not live at creation): the system must then be created again."""
        ids = np.ascontiguousarray(np.atleast_1d(tets), dtype=np.int32)
        flags = np.ascontiguousarray(np.broadcast_to(np.asarray(active, np.uint8), ids.shape))
        rc = self.lib.pk_system_set_tets_active(self.h, len(ids), N.ptr(ids), N.ptr(flags))
        if rc == N.PK_ERR_REBUILD:
            return False
        N.check(rc, "pk_system_set_tets_active")
        return True

    def set_matrix_store(self, max_bytes):
        """Bytes the reverse mode may keep for a checkpoint segment's matrices
        (default 8 GiB); 0 re-assembles every reverse substep instead."""
        N.check(self.lib.pk_system_set_matrix_store(self.h, int(max_bytes)), "pk_system_set_matrix_store")

    def set_graphs(self, mode=1):
        """HIP graphs (identical results to launching each kernel): 0 off;
        1 (or True) each solve's CG replayed four iterations per graph;
        2 every implicit forward substep replayed as one captured graph."""
        N.check(self.lib.pk_system_set_graphs(self.h, int(mode)), "pk_system_set_graphs")

    def spmv_work(self):
        """(full BSR blocks, upper-triangle blocks among them, dynamic rows,
        body-launches) the profiled SpMV launches multiplied; stopped bodies'
        rows are not counted."""
        v = [C.c_longlong() for _ in range(4)]
        N.check(self.lib.pk_system_spmv_work(self.h, *[C.byref(x) for x in v]), "pk_system_spmv_work")
        return tuple(x.value for x in v)

    def assemble(self, dt):
        N.check(self.lib.pk_system_assemble(self.h, dt), "pk_system_assemble")

    def matrix(self):
        vals = np.zeros((self.nnzb_full, 3, 3))
        rhs = np.zeros((self.nd, 3))
        minv = np.zeros((self.nd, 3, 3))
        N.check(self.lib.pk_system_get_matrix(self.h, N.ptr(vals), N.ptr(rhs), N.ptr(minv)), "pk_system_get_matrix")
        return vals, rhs, minv

    def elements(self):
        blocks = np.zeros((self.t, 10, 3, 3))
        forces = np.zeros((self.t, 4, 3))
        kv = np.zeros((self.t, 4, 3))
        N.check(self.lib.pk_system_get_elements(self.h, N.ptr(blocks), N.ptr(forces), N.ptr(kv)),
                "pk_system_get_elements")
        return blocks, forces, kv

    def rotations(self):
        """(R per tet (t, 3, 3), polar kind per tet: 0 Newton, 1 column 2
        negated, 2 Identity fallback, -1 no contribution)."""
        R = np.zeros((self.t, 3, 3))
        kind = np.zeros(self.t, np.int32)
        N.check(self.lib.pk_system_get_rotations(self.h, N.ptr(R), N.ptr(kind)), "pk_system_get_rotations")
        return R, kind

    def spmv(self, x):
        xa = N.as_f64(x).reshape(-1)
        y = np.zeros(self.nd * 3)
        N.check(self.lib.pk_system_spmv(self.h, xa, y), "pk_system_spmv")
        return y.reshape(-1, 3)

    def solution(self):
        dv = np.zeros(self.nd * 3)
        N.check(self.lib.pk_system_get_solution(self.h, dv), "pk_system_get_solution")
        return dv.reshape(-1, 3)

    def record(self, dt, steps, every=None):
        """Run `steps` implicit substeps of length dt from the current state,
        keeping device checkpoints every `every` substeps (default: all in one
        segment) for a later backward()."""
        self._steps = int(steps)
        N.check(self.lib.pk_system_record(self.h, dt, int(steps), int(every or steps)), "pk_system_record")

    def backward(self, xbar=None, vbar=None, tol=1e-6, max_iter=1024):
        """Adjoints of the recorded run's final state -> (dL/dlogE per tet,
        dL/dx0, dL/dv0, adjoint CG iterations per substep)."""
        xb = None if xbar is None else N.as_f64(xbar).reshape(-1)
        vb = None if vbar is None else N.as_f64(vbar).reshape(-1)
        gE = np.zeros(self.t)
        gx = np.zeros((self.n, 3))
        gv = np.zeros((self.n, 3))
        its = np.zeros(max(getattr(self, "_steps", 0), 1), np.int32)
        N.check(self.lib.pk_system_backward(self.h, N.ptr(xb), N.ptr(vb), tol, int(max_iter), N.ptr(gE), N.ptr(gx),
                                            N.ptr(gv), N.ptr(its)), "pk_system_backward")
        return gE, gx, gv, its[:getattr(self, "_steps", 0)]


    def backward_device(self, xbar_ptr=None, vbar_ptr=None, grad_log_young_ptr=None, grad_x0_ptr=None,
                        grad_v0_ptr=None, tol=1e-6, max_iter=1024):
        """backward() with device addresses for the seeds and the gradients;
        returns the adjoint CG iterations per substep."""
        its = np.zeros(max(getattr(self, "_steps", 0), 1), np.int32)
        N.check(self.lib.pk_system_backward_device(self.h, xbar_ptr, vbar_ptr, tol, int(max_iter), grad_log_young_ptr,
                                                   grad_x0_ptr, grad_v0_ptr, N.ptr(its)), "pk_system_backward_device")
        return its[:getattr(self, "_steps", 0)]


def device_count() -> int:
    return N.lib().pk_device_count()
