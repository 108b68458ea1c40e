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
(or FP32-storage) GPU path; tensors cross the boundary as host float64
arrays, and the gradients come back on the inputs' device.

The system keeps one recorded run: call backward before stepping, recording
or changing the system again (pk_system_backward refuses a run that a later
change invalidated).
"""

from __future__ import annotations

import numpy as np
import torch


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class ImplicitRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, system, log_young, x0, v0, dt, steps, every=None, adj_tol=1e-8, adj_max_iter=2048):
        system.set_young(np.exp(_host(log_young)))
        system.set_state(_host(x0), _host(v0))
        system.record(float(dt), int(steps), int(every) if every else None)
        x, v = system.state()
        ctx.system = system
        ctx.adj = (float(adj_tol), int(adj_max_iter))
        ctx.set_materialize_grads(False)  # an unused output seeds nothing
        ctx.out = (log_young.device, log_young.dtype, x0.device, x0.dtype, v0.device, v0.dtype)
        return torch.from_numpy(x).to(x0.device), torch.from_numpy(v).to(v0.device)

    @staticmethod
    def backward(ctx, grad_x, grad_v):
        xbar = None if grad_x is None else _host(grad_x)
        vbar = None if grad_v is None else _host(grad_v)
        tol, max_iter = ctx.adj
        g_log_young, g_x0, g_v0, _ = ctx.system.backward(xbar, vbar, tol=tol, max_iter=max_iter)
        dev_e, dt_e, dev_x, dt_x, dev_v, dt_v = ctx.out
        return (None, torch.from_numpy(g_log_young).to(dev_e, dt_e), torch.from_numpy(g_x0).to(dev_x, dt_x),
                torch.from_numpy(g_v0).to(dev_v, dt_v), None, None, None, None, None)
