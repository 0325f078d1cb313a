"""torch.autograd bridge to the device adjoint (include/gamadj.h).

``StepObjective(g, nsteps, ckpt_every, weights)`` turns a prepared
context into a differentiable function of the initial gas
state: calling it with a float64 tensor ``gas0`` of shape (Ni, Nj, Nk, 5)
puts ``gas0`` into the context's State, runs the set-up tail
(``gam_init_run``: boundary conditions, oState = State, dt = CalcDT), steps
``nsteps`` times recording checkpoints every ``ckpt_every`` steps, and
returns J = sum_c w_c P_c of the final State as a 0-d tensor.  Its backward
runs the device adjoint; because State and oState both start from
``gas0``, the gradient is the sum of the two adjoints, taken with respect
to the active cells (ghost cells follow through the boundary conditions;
the gradient's ghost entries are 0).  Each step keeps the dt of the forward
run.  torch handles only host tensors here: the state crosses PCIe once
per call, the steps and the adjoint run on the device.
"""
import numpy as np
import torch

from . import gamhip


class _StepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gas0, obj):
        J = obj._forward(gas0.detach().cpu().numpy())
        ctx.obj = obj
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gS, gO = ctx.obj.adj.backward()
        grad = torch.from_numpy(np.ascontiguousarray(gS + gO)) * gJ
        return grad, None


class StepObjective:
    def __init__(self, g, nsteps, ckpt_every, weights):
        self.g = g
        self.template = g.get_state(gamhip.GAM_STATE)
        self.t0 = self.template["time"]
        self.adj = gamhip.GamAdjoint(g, nsteps, ckpt_every)
        self.adj.set_weights(weights)

    def _forward(self, gas0):
        st = dict(self.template)
        st["Gas"] = np.asfortranarray(gas0, dtype=np.float64)
        st["time"] = self.t0
        self.g.put_state(st)
        t, dt, ts = self.g.get_time()
        self.g.set_time(self.t0, dt, ts)
        self.g.init_run()
        return self.adj.forward()

    def __call__(self, gas0):
        if gas0.dtype != torch.float64 or tuple(gas0.shape) != self.g.cshape + (5,):
            raise ValueError("gas0 must be a float64 tensor of shape (Ni, Nj, Nk, 5)")
        return _StepFn.apply(gas0, self)

    def close(self):
        self.adj.close()
