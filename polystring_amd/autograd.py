"""Thin torch.autograd binding of ln Q(wA, wB) of a Model_AB chain.

The forward pass sets the model's fields and runs the adjoint field gradient
through the C ABI (po_model_grad_w); the backward pass hands back
d(ln Q)/dw.  The GPU work runs in the library; torch only carries the values.
"""
import numpy as np
import torch


class LogQ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, wA, wB, model):
        model.set_fields(wA.detach().double().cpu().numpy(), wB.detach().double().cpu().numpy())
        gA, gB, Q = model.grad_w()
        M = gA.size
        ctx.save_for_backward(torch.from_numpy(gA / M).to(wA), torch.from_numpy(gB / M).to(wB))
        return wA.new_tensor(float(np.log(Q)))

    @staticmethod
    def backward(ctx, grad):
        gA, gB = ctx.saved_tensors
        return grad * gA, grad * gB, None


def log_q(wA, wB, model):
    """ln Q of `model` (a capi.ModelAB) at fields wA, wB, differentiable in both."""
    return LogQ.apply(wA, wB, model)
