"""torch.autograd bridge to the device adjoint (include/gamadj.h).

``StepObjective(g, nsteps, ckpt_every, weights)`` turns a prepared
context into a differentiable function of the initial gas state and,
when the context has an inflow boundary, of the inflow time series:

    J = obj(gas0)                 # or obj(gas0, Q, B) with wind=True

``gas0`` is a float64 CPU tensor of shape (Ni, Nj, Nk, 5).  A call puts
``gas0`` into the context's State at the time
and step count the context had when the objective was made, runs the
set-up tail (``gam_init_run``: boundary conditions, oState = State,
dt = CalcDT), steps ``nsteps`` times recording checkpoints every
``ckpt_every`` steps, and returns J = sum_c w_c P_c of the final State as a
0-d float64 tensor.  Its backward runs the device
adjoint; because State and oState both start from ``gas0``, the gradient
is the sum of the two adjoints, taken with respect to the active cells
(ghost cells follow through the boundary conditions; the gradient's ghost
entries are 0).  By default each step keeps the dt of the forward run (the
gradient of the forward with that dt sequence held).  With
``dt_follows=True`` the adjoint also differentiates every dt: the set-up's
CalcDT of gas0 and the CalcDT after each step (gam_adj_set_dt_mode 3), so
the gradient is that of this forward, whose dt follows the state.

With ``wind=True`` the call takes the inflow series too, ``Q`` (n, 5)
primitives and ``B`` (n, 3), as given to ``gam_set_wind``, and the backward
returns dJ/dQ and dJ/dB (the Bx column's gradient is 0: Bx follows from
By and Bz; the sample index, a step function of time, is not
differentiated).

Tensors are host tensors: libgamhip runs on the HIP runtime of the ROCm it
was built with, and torch's own HIP runtime cannot see that GPU in the same
process, so the state crosses PCIe once per call (the steps and the adjoint
run on the device).  Callers on the library's runtime (C, C++, the Fortran
host) can pass device pointers to the same C ABI instead.

One adjoint recording serves one call.  Each call is numbered; a backward
whose call is not the last one recorded (J = obj(a) + obj(b), line
searches) records its own forward again before running the adjoint, so
every gradient belongs to its own input.

oState's ghost cells: the forward fills them with State's (the boundary
pass at State%time, as the reference's set-up does, init.F90:230-244), and
the adjoint is told so (gam_adj_set_ostate_bc_time), so it folds oState's
ghost adjoint through the boundary conditions at State%time too: the
gradient is that of this forward, inflow boundary included.
"""
import numpy as np
import torch

from . import gamhip


class _StepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gas0, Q, B, obj):
        J = obj._forward(gas0, Q, B)
        ctx.obj = obj
        ctx.gen = obj._gen
        ctx.save_for_backward(gas0, Q, B)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        obj = ctx.obj
        gas0, Q, B = ctx.saved_tensors
        if ctx.gen != obj._gen:        # another call was recorded since: record this one again
            obj._forward(gas0, Q, B)
        grad = obj._backward(gas0)
        gQ = gB = None
        if obj.wind:
            nq, nb = obj.adj.wind_grad(obj.nw)
            gQ = torch.from_numpy(np.ascontiguousarray(nq)) * gJ
            gB = torch.from_numpy(np.ascontiguousarray(nb)) * gJ
        return grad * gJ, gQ, gB, None


class StepObjective:
    def __init__(self, g, nsteps, ckpt_every, weights, wind=False, dt_follows=False):
        self.g = g
        self.template = g.get_state(gamhip.GAM_STATE)
        self.t0 = self.template["time"]
        self.ts0 = g.get_time()[2]
        self.adj = gamhip.GamAdjoint(g, nsteps, ckpt_every)
        self.adj.set_weights(weights)
        self.adj.set_ostate_bc_time(True)
        if dt_follows:
            self.adj.set_dt_mode(3)
        self.wind = bool(wind)
        self.windspec = None
        if self.wind:
            self.windspec = g.wind
            if self.windspec is None:
                raise ValueError("wind=True needs a context with an inflow series (set_wind)")
            self.nw = len(self.windspec["tW"])
        self._gen = 0

    def _put_gas(self, gas0):
        st = dict(self.template)
        st["time"] = self.t0
        st["Gas"] = np.asfortranarray(gas0.detach().numpy(), dtype=np.float64)
        self.g.put_state(st)

    def _forward(self, gas0, Q=None, B=None):
        if self.wind:
            w = dict(self.windspec)
            w["Q"] = np.asfortranarray(Q.detach().numpy(), dtype=np.float64)
            w["B"] = np.asfortranarray(B.detach().numpy(), dtype=np.float64)
            self.g.set_wind(w)
        self._put_gas(gas0)
        t, dt, ts = self.g.get_time()
        self.g.set_time(self.t0, dt, self.ts0)
        self.g.init_run()
        J = self.adj.forward()
        self._gen += 1
        return J

    def _backward(self, gas0):
        gS, gO = self.adj.backward()
        return torch.from_numpy(np.ascontiguousarray(gS + gO))

    def __call__(self, gas0, Q=None, B=None):
        if gas0.dtype != torch.float64 or tuple(gas0.shape) != self.g.cshape + (5,) or gas0.device.type != "cpu":
            raise ValueError("gas0 must be a float64 CPU tensor of shape (Ni, Nj, Nk, 5)")
        if self.wind:
            if Q is None or B is None or tuple(Q.shape) != (self.nw, 5) or tuple(B.shape) != (self.nw, 3):
                raise ValueError("wind=True: Q (n, 5) and B (n, 3) float64 CPU tensors are required")
        else:
            if Q is not None or B is not None:
                raise ValueError("the inflow series is an input only with wind=True")
            Q = B = torch.zeros(0, dtype=torch.float64)
        return _StepFn.apply(gas0, Q, B, self)

    def close(self):
        self.adj.close()
