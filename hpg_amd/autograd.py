"""Differentiable degridding: the degridder as a torch.autograd.Function.

Forward: model grid (complex128, shape (cube, mrow, y, x), on the GPU) ->
predicted visibilities (complex128, (n, npol)) through
``GridderState.degrid_forward`` (include/hpgx.h hpgx_degrid_forward).
Backward: the hand-written adjoint gridding kernel
(``GridderState.degrid_adjoint``, hpgx_degrid_adjoint) applied to the
incoming gradient.  For a real loss L, torch's complex gradient convention
(dL/dRe z + i dL/dIm z) makes the vector-Jacobian product of the complex-
linear degridding operator A equal to A^H applied to grad_output.

``ImageToModel`` is the image-domain leg in front of it: a real sky image
(cube, mrow, y, x) -> the model grid shift(FFT_sign(image)), computed by the
state's model passes (apply_model_fft, shift_model: include/hpgx.h
hpgx_apply_model_fft / hpgx_shift_model); its backward applies the adjoint
passes, the opposite shift and the unnormalized FFT of the opposite sign
(F_sign^H = F_-sign), and keeps the real part (the image is real).  So
``degrid(image_to_model(image, ...), ...)`` gives the gradient of a loss on
the predicted visibilities with respect to the image pixels.

Intended for gradient checks and image-domain fitting; both directions run
through the C ABI (there is no torch fallback).
"""
import numpy as np
import torch

from . import Device, FFTSign, GridderState, ShiftDirection  # noqa: F401


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


def _opposite_sign(sign):
    return FFTSign.POSITIVE if FFTSign(sign) == FFTSign.NEGATIVE else FFTSign.NEGATIVE


def _opposite_shift(direction):
    return (ShiftDirection.BACKWARD if ShiftDirection(direction) == ShiftDirection.FORWARD
            else ShiftDirection.FORWARD)


def _model_pass(state, values, sign, direction, fft_first):
    """the state's model set to `values` (complex128 GPU tensor), then its
    unnormalized FFT of `sign` and its shift in `direction` (FFT first when
    fft_first), read back as a complex128 tensor on values' device"""
    state.set_model(Device.Serial, values)
    if fft_first:
        state.apply_model_fft(1.0, sign, True)
        state.shift_model(direction)
    else:
        state.shift_model(direction)
        state.apply_model_fft(1.0, sign, True)
    return torch.from_numpy(np.ascontiguousarray(state.model_values())).to(values.device)


class ImageToModel(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, state, sign, direction):
        ctx.state, ctx.sign, ctx.direction = state, sign, direction
        return _model_pass(state, image.detach().to(torch.complex128).contiguous(), sign,
                           direction, True)

    @staticmethod
    def backward(ctx, grad_out):
        # (shift o F_sign)^H = F_-sign o shift^-1 on the incoming gradient
        g = _model_pass(ctx.state, grad_out.contiguous().to(torch.complex128),
                        _opposite_sign(ctx.sign), _opposite_shift(ctx.direction), False)
        return g.real.contiguous(), None, None, None


def image_to_model(image, state, sign=FFTSign.NEGATIVE, direction=ShiftDirection.FORWARD):
    """model = shift(FFT_sign(image)) through the state's model passes,
    differentiable in the (real) image"""
    return ImageToModel.apply(image, state, sign, direction)
