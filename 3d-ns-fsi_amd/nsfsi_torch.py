"""PyTorch harness over the discrete adjoint (ns_adjoint_drag): the drag objective as a differentiable
function of the initial conserved state, the inlet velocity profile and the Reynolds number.

    J = DragObjective.apply(u0, inflow, re, config, nsteps, n0, n1, checkpoint, dts)

u0: (5, NK, NJ, NI) float64 tensor, inflow: inlet-plane float64 tensor (see Solver.inletPlaneShape),
re: 0-d float64 tensor. The forward runs the FP64 device solver; backward returns the gradients the
hand-written reverse sweep computed (no autograd tape: the solver is a black box to torch). dts: step
sizes by time level to run at (None: the stability limit of this run, frozen for the gradient).
"""
import ctypes as C

import numpy as np
import torch

import nsfsi


def _config_with_re(config, re):
    c = nsfsi.NSConfig()
    C.pointer(c)[0] = config
    c.Re = float(re)
    return c


def drag_and_gradients(config, u0, inflow, re, nsteps, n0, n1, checkpoint=20, dts=None):
    """One forward + reverse run; returns the dict of Solver.adjointDrag."""
    s = nsfsi.Solver(_config_with_re(config, re))
    try:
        s.initialize()
        s.setInitialField(np.asarray(u0, dtype=np.float64).reshape(-1))
        if inflow is not None:
            s.setInflowProfile(np.asarray(inflow, dtype=np.float64).reshape(-1))
        if dts is not None:
            s.setDtSchedule(dts)
        return s.adjointDrag(nsteps, n0, n1, checkpoint)
    finally:
        s.close()


class DragObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, inflow, re, config, nsteps, n0, n1, checkpoint=20, dts=None):
        r = drag_and_gradients(config, u0.detach().cpu().numpy(), inflow.detach().cpu().numpy(), re.item(), nsteps,
                               n0, n1, checkpoint, dts)
        dev, dt = u0.device, u0.dtype
        ctx.save_for_backward(torch.as_tensor(r["grad_u0"], dtype=dt, device=dev),
                              torch.as_tensor(r["grad_inflow"], dtype=dt, device=dev),
                              torch.tensor(r["grad_re"], dtype=dt, device=dev))
        ctx.dts = r["dts"]
        return torch.tensor(r["J"], dtype=dt, device=dev)

    @staticmethod
    def backward(ctx, gJ):
        gu, gq, gre = ctx.saved_tensors
        return gJ * gu, gJ * gq, gJ * gre, None, None, None, None, None, None
