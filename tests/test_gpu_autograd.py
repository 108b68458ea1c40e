"""The reverse mode behind torch.autograd (physik_amd.autograd.ImplicitRollout).

Bars: the gradients autograd returns are bit-identical to pk_system_backward
called directly with the same seeds (the seeds autograd forms for the loss
equal W.loss_and_seed's), and they agree with central differences of the
Function's own forward (FP64, tight CG) to 1e-5 relative on a few log E
entries and a v0 direction.
"""

import numpy as np
import pytest
import torch

from physik_amd import native as N
from physik_amd import workloads as W
from physik_amd.autograd import ImplicitRollout

from test_gpu_adjoint import small

pytestmark = pytest.mark.gpu


def rollout_loss(s, w, log_e, v0):
    x0 = torch.from_numpy(w.x0.copy())
    xS, vS = ImplicitRollout.apply(s, log_e, x0, v0, w.dt, w.steps, 3, 1e-12, 4000)
    nodes = torch.from_numpy(np.asarray(w.loss_nodes))
    return ((xS[nodes] - torch.from_numpy(w.target)) ** 2).sum(), xS


def test_autograd_gradients_equal_the_direct_backward():
    w = small(tol=1e-12)
    s = W.make_system(w, N.PK_FP64)
    log_e = torch.tensor(np.log(w.young), requires_grad=True)
    v0 = torch.tensor(w.v0.copy(), requires_grad=True)
    loss, xS = rollout_loss(s, w, log_e, v0)
    loss.backward()
    print(f"loss {float(loss.detach()):.6e}", flush=True)
    # the same run and seeds through pk_system_backward directly
    s.set_young(w.young)
    s.set_state(w.x0, w.v0)
    s.record(w.dt, w.steps, 3)
    x, _ = s.state()
    assert np.array_equal(x, xS.detach().numpy())
    L, xbar = W.loss_and_seed(w, x)
    assert abs(L - float(loss.detach())) <= 1e-12 * abs(L)
    gE, gx, gv, _ = s.backward(xbar, None, tol=1e-12, max_iter=4000)
    assert np.array_equal(log_e.grad.numpy(), gE)
    assert np.array_equal(v0.grad.numpy(), gv)
    s.close()


def test_autograd_gradients_match_central_differences():
    w = small(tol=1e-12)
    s = W.make_system(w, N.PK_FP64)
    base_e = np.log(w.young)
    log_e = torch.tensor(base_e, requires_grad=True)
    v0 = torch.tensor(w.v0.copy(), requires_grad=True)
    loss, _ = rollout_loss(s, w, log_e, v0)
    loss.backward()
    ge, gv = log_e.grad.numpy().copy(), v0.grad.numpy().copy()

    def f(e, v):
        with torch.no_grad():
            return float(rollout_loss(s, w, torch.from_numpy(e), torch.from_numpy(v))[0])

    h = 1e-5
    rng = np.random.default_rng(3)
    for t in rng.choice(len(base_e), 4, replace=False):
        ep, em = base_e.copy(), base_e.copy()
        ep[t] += h
        em[t] -= h
        fd = (f(ep, w.v0) - f(em, w.v0)) / (2 * h)
        print(f"log E[{t}]: adjoint {ge[t]:.9e} fd {fd:.9e}", flush=True)
        assert abs(ge[t] - fd) <= 1e-5 * max(abs(fd), np.abs(ge).max() * 1e-3)
    dirv = rng.normal(size=w.v0.shape)
    dirv[w.fixed == 1] = 0.0
    fd = (f(base_e, w.v0 + h * dirv) - f(base_e, w.v0 - h * dirv)) / (2 * h)
    ad = float(np.sum(gv * dirv))
    print(f"v0 direction: adjoint {ad:.9e} fd {fd:.9e}", flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    s.close()


def test_autograd_on_device_tensors_equals_the_host_path():
    # tensors on the GPU cross the boundary as device pointers: the outputs
    # and gradients must be bit-identical to the host-tensor path (a linear
    # loss, so the seeds torch forms are the same constants on both devices)
    w = small(tol=1e-12)
    s = W.make_system(w, N.PK_FP64)
    rng = np.random.default_rng(5)
    cx, cv = rng.normal(size=w.x0.shape), rng.normal(size=w.v0.shape)
    out = []
    for dev in ("cpu", "cuda"):
        log_e = torch.tensor(np.log(w.young), requires_grad=True, device=dev)
        v0 = torch.tensor(w.v0.copy(), requires_grad=True, device=dev)
        x0 = torch.tensor(w.x0.copy(), requires_grad=True, device=dev)
        xS, vS = ImplicitRollout.apply(s, log_e, x0, v0, w.dt, w.steps, 3, 1e-12, 4000)
        assert xS.device.type == dev and vS.device.type == dev
        loss = (xS * torch.from_numpy(cx).to(dev)).sum() + (vS * torch.from_numpy(cv).to(dev)).sum()
        loss.backward()
        out.append([t.detach().cpu().numpy() for t in (xS, vS, log_e.grad, x0.grad, v0.grad)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    s.close()


def test_backward_of_an_overwritten_recording_raises():
    # two rollouts on one system before the first backward: the first
    # backward would reverse the second run, so it must refuse
    w = small(tol=1e-10)
    s = W.make_system(w, N.PK_FP64)
    log_e = torch.tensor(np.log(w.young), requires_grad=True)
    v0 = torch.tensor(w.v0.copy(), requires_grad=True)
    loss1, _ = rollout_loss(s, w, log_e, v0)
    loss2, _ = rollout_loss(s, w, log_e, v0)
    with pytest.raises(RuntimeError, match="recorded another run"):
        loss1.backward()
    loss2.backward()  # the latest run still has its recording
    assert np.isfinite(log_e.grad.numpy()).all()
    s.close()
