"""The ptychographic forward model as a differentiable torch function.

``ptychographic_loss(obj, probe, positions, amplitudes)`` returns
L = sum_b sum (|fft2(P_shift_b * O_patch_b)| - A_b)^2 (float64), the
unnormalised error of the reference's GD loop
(py4DSTEM/process/phase/ptychographic_methods.py:1576-1635), and its
backward runs the hand-written adjoint kernels of libpty_hip.so
(``pty_loss_gradients``): the object gradient is -2 S^2 times the object term
of ``_gradient_descent_adjoint`` (ptychographic_methods.py:1878-1884), the
probe gradient includes the adjoint of the Fourier sub-pixel shift.
Gradients follow torch's convention for real losses of complex inputs,
dL/dRe + i dL/dIm.
"""

from __future__ import annotations

import torch

from . import _lib
from . import kernels as K
from ._lib import ptr, stream_handle


def _loss_call(obj, probe, positions, amplitudes_k, order, chunk, scale, want_obj, want_probe):
    S = probe.shape[-1]
    K.check_frame(S, lds_only=True)
    Px, Py = obj.shape
    N = order.numel()
    ws_bytes = _lib.load().pty_loss_workspace_bytes(S, Px, Py, N, chunk)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=obj.device)
    loss = torch.empty(1, dtype=torch.float64, device=obj.device)
    go = torch.empty_like(obj) if want_obj else None
    gp = torch.empty_like(probe) if want_probe else None
    if positions.shape[0] != amplitudes_k.shape[0]:
        raise ValueError("positions and amplitudes must describe the same scan")
    _lib.call("pty_loss_gradients", ptr(obj), ptr(probe), Px, Py, ptr(positions),
              positions.shape[0], ptr(amplitudes_k), ptr(order), N, int(chunk), S, float(scale),
              ptr(loss), ptr(go), ptr(gp), ptr(ws), ws.numel(), stream_handle())
    return loss[0], go, gp


class PtychographicLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj, probe, positions, amplitudes_k, order, chunk):
        obj_c, probe_c = obj.detach().contiguous(), probe.detach().contiguous()
        loss, _, _ = _loss_call(obj_c, probe_c, positions, amplitudes_k, order, chunk, 1.0,
                                False, False)
        ctx.save_for_backward(obj_c, probe_c, positions, amplitudes_k, order)
        ctx.chunk = chunk
        return loss

    @staticmethod
    def backward(ctx, grad_out):
        obj, probe, positions, amplitudes_k, order = ctx.saved_tensors
        _, go, gp = _loss_call(obj, probe, positions, amplitudes_k, order, ctx.chunk, 1.0,
                               ctx.needs_input_grad[0], ctx.needs_input_grad[1])
        # the incoming gradient stays on the device (no host synchronisation)
        g = grad_out.to(torch.float32)
        return (go * g if go is not None else None), (gp * g if gp is not None else None), \
            None, None, None, None


def ptychographic_loss(obj, probe, positions, amplitudes_k, order=None, chunk=4096):
    """Differentiable loss on the current ROCm device.

    obj: complex64 (Px, Py); probe: complex64 (S, S), corner-centred;
    positions: float32 (N, 2) object pixels; amplitudes_k: float32 (N, S, S)
    in kernel layout (``kernels.to_kernel_layout``); order: int32 slots to
    include (default: all positions).
    """
    if order is None:
        order = torch.arange(positions.shape[0], dtype=torch.int32, device=positions.device)
    return PtychographicLoss.apply(obj, probe, positions, amplitudes_k, order, int(chunk))
