"""torch.autograd.Function over a recorded run of implicit steps and its adjoint.

    from physik_amd.autograd import ImplicitRollout
    x_S, v_S = ImplicitRollout.apply(system, log_young, x0, v0, dt, steps, every)
    loss = ((x_S[end] - target) ** 2).sum()
    loss.backward()          # log_young.grad, x0.grad, v0.grad

forward sets the per-tet Young's moduli exp(log_young) and the state (x0, v0)
of `system` (a physik_amd.system.System), runs pk_system_record (`steps`
implicit substeps of length dt, checkpoints every `every`) and returns the
final positions and velocities. backward runs pk_system_backward with the
incoming gradients of those as the adjoint seeds (checkpoint replay, adjoint
PCG on the assembled system, corotational adjoint with dR/dF) and returns
dL/dlog E per tet, dL/dx0 and dL/dv0. The arithmetic is the library's FP64
(or FP32-storage) GPU path.

Tensors on a GPU device cross the boundary as device pointers
(pk_system_set_state_device, pk_system_get_state_device,
pk_system_backward_device): the state and the gradients stay in device
memory, after a synchronisation of the current stream. Host tensors go
through host float64 arrays. The Young's moduli always pass through the
host, where the library derives the Lame parameters from them.

The system keeps one recorded run. backward raises when the system has
recorded another run since this forward (pk_system_tape_id) or when a later
change discarded the recording (pk_system_backward refuses it).
"""

from __future__ import annotations

import numpy as np
import torch


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


def _dev64(t):
    """A contiguous float64 copy (or view) of t on its own device."""
    return t.detach().to(torch.float64).contiguous()


class ImplicitRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, system, log_young, x0, v0, dt, steps, every=None, adj_tol=1e-8, adj_max_iter=2048):
        system.set_young(np.exp(_host(log_young)))
        on_gpu = x0.is_cuda and v0.is_cuda
        if on_gpu:
            xs, vs = _dev64(x0), _dev64(v0)
            torch.cuda.current_stream(xs.device).synchronize()
            system.set_state_device(xs.data_ptr(), vs.data_ptr())
        else:
            system.set_state(_host(x0), _host(v0))
        system.record(float(dt), int(steps), int(every) if every else None)
        ctx.system = system
        ctx.tape = system.tape_id()
        ctx.on_gpu = on_gpu
        ctx.adj = (float(adj_tol), int(adj_max_iter))
        ctx.set_materialize_grads(False)  # an unused output seeds nothing
        ctx.out = (log_young.device, log_young.dtype, x0.device, x0.dtype, v0.device, v0.dtype)
        if on_gpu:
            x = torch.empty((system.n, 3), dtype=torch.float64, device=x0.device)
            v = torch.empty((system.n, 3), dtype=torch.float64, device=v0.device)
            system.get_state_device(x.data_ptr(), v.data_ptr())
            return x.to(x0.dtype), v.to(v0.dtype)
        x, v = system.state()
        return torch.from_numpy(x).to(x0.device, x0.dtype), torch.from_numpy(v).to(v0.device, v0.dtype)

    @staticmethod
    def backward(ctx, grad_x, grad_v):
        s = ctx.system
        if s.tape_id() != ctx.tape:
            raise RuntimeError("ImplicitRollout.backward: the system has recorded another run (or changed) since "
                               "this forward; its gradients would belong to that run")
        tol, max_iter = ctx.adj
        dev_e, dt_e, dev_x, dt_x, dev_v, dt_v = ctx.out
        if ctx.on_gpu:
            dev = dev_x
            gx = None if grad_x is None else _dev64(grad_x.to(dev))
            gv = None if grad_v is None else _dev64(grad_v.to(dev))
            g_e = torch.empty(s.t, dtype=torch.float64, device=dev)
            g_x0 = torch.empty((s.n, 3), dtype=torch.float64, device=dev)
            g_v0 = torch.empty((s.n, 3), dtype=torch.float64, device=dev)
            torch.cuda.current_stream(dev).synchronize()
            s.backward_device(None if gx is None else gx.data_ptr(), None if gv is None else gv.data_ptr(),
                              g_e.data_ptr(), g_x0.data_ptr(), g_v0.data_ptr(), tol=tol, max_iter=max_iter)
            return (None, g_e.to(dev_e, dt_e), g_x0.to(dev_x, dt_x), g_v0.to(dev_v, dt_v),
                    None, None, None, None, None)
        xbar = None if grad_x is None else _host(grad_x)
        vbar = None if grad_v is None else _host(grad_v)
        g_log_young, g_x0, g_v0, _ = s.backward(xbar, vbar, tol=tol, max_iter=max_iter)
        return (None, torch.from_numpy(g_log_young).to(dev_e, dt_e), torch.from_numpy(g_x0).to(dev_x, dt_x),
                torch.from_numpy(g_v0).to(dev_v, dt_v), None, None, None, None, None)
