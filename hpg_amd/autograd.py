"""Differentiable degridding: the degridder as a torch.autograd.Function.

Forward: model grid (complex128, shape (cube, mrow, y, x), on the GPU) ->
predicted visibilities (complex128, (n, npol)) through
``GridderState.degrid_forward`` (include/hpgx.h hpgx_degrid_forward).
Backward: the hand-written adjoint gridding kernel
(``GridderState.degrid_adjoint``, hpgx_degrid_adjoint) applied to the
incoming gradient.  For a real loss L, torch's complex gradient convention
(dL/dRe z + i dL/dIm z) makes the vector-Jacobian product of the complex-
linear degridding operator A equal to A^H applied to grad_output.

Intended for gradient checks and image-domain fitting; both directions run
through the C ABI (there is no torch fallback).
"""
import numpy as np
import torch

from . import Device, GridderState  # noqa: F401


class Degrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, state, records):
        state.set_model(Device.Serial, model.detach().contiguous())
        pred = state.degrid_forward(records)
        ctx.state = state
        ctx.records = records
        return pred

    @staticmethod
    def backward(ctx, grad_out):
        state = ctx.state
        state.reset_grid()
        state.degrid_adjoint(ctx.records, grad_out.contiguous())
        g = torch.from_numpy(np.ascontiguousarray(state.grid_values())).to(grad_out.device)
        return g, None, None


def degrid(model, state, records):
    """predicted visibilities = A(model), differentiable in model."""
    return Degrid.apply(model, state, records)
