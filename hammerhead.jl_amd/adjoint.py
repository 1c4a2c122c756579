"""The windowed correlation energy as a differentiable function.

``CorrelationEnergy`` is a ``torch.autograd.Function`` (the counterpart of
a ``ChainRulesCore.rrule`` on the Julia side): its forward evaluates E of a
node field on the GPU (``hh_energy``) and its backward calls the library's
hand-written adjoint (``hh_grad_energy``), which recomputes the warped
windows rather than keeping them from the forward. The reference has no
gradient path; the energy is defined in include/hammerhead_mi355x.h.

    E = correlation_energy(pu, pv, ctx, params, py, px)   # pu, pv: tensors
    E.backward()                                          # pu.grad, pv.grad
"""
from __future__ import annotations

import numpy as np
import torch


class CorrelationEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pu, pv, context, params, py, px):
        fctx.save_for_backward(pu, pv)
        fctx.args = (context, params, np.asarray(py, np.float32), np.asarray(px, np.float32))
        E = context.energy(params, py, px, pu.detach().cpu().numpy(), pv.detach().cpu().numpy())
        return torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(fctx, gE):
        pu, pv = fctx.saved_tensors
        context, params, py, px = fctx.args
        _, gu, gv = context.grad_energy(params, py, px, pu.detach().cpu().numpy(), pv.detach().cpu().numpy())
        s = float(gE)
        gu = torch.from_numpy(np.ascontiguousarray(gu)).to(dtype=pu.dtype, device=pu.device) * s
        gv = torch.from_numpy(np.ascontiguousarray(gv)).to(dtype=pv.dtype, device=pv.device) * s
        return gu, gv, None, None, None, None


def correlation_energy(pu, pv, context, params, py, px):
    """E(pu, pv) with an autograd backward through the GPU adjoint."""
    return CorrelationEnergy.apply(pu, pv, context, params, py, px)


def refine_final_pass(context, result, params, niter: int = 20):
    """Gradient-based refinement of a final-pass field: the PIVResult's
    (u, v) on its (y, x) grid is taken as the node field and refined by
    ``niter`` descent steps of the correlation energy over the windows of
    ``params``. Returns (u, v, energies)."""
    return context.refine_nodes(params, result.y, result.x, result.u, result.v, niter)
