"""Tensor-level wrappers over libpty_hip.so.

Every function takes and returns torch tensors on the current ROCm device and
launches on the current torch stream; shapes and dtypes are checked here so
that the C-ABI only ever sees consistent pointers.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import PtyGdParams, PtyProblem, c_void_p, ptr, stream_handle

SUPPORTED_FRAMES = (32, 64, 128, 256)
# frames transformed inside one work group with the frame in LDS; 256 px
# frames run gradient descent through line passes in global memory
LDS_FRAMES = (32, 64, 128)


def _req(t, dtype, shape=None, name="tensor"):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA/ROCm tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


# radix split N1 x N2 of each frame size (N1 along a row pass, N2 lanes per
# line); (S, 1) makes the 256 px kernel layout the natural order
FRAME_RADIX = {32: (8, 4), 64: (8, 8), 128: (16, 8), 256: (256, 1)}


def check_frame(S, lds_only=False):
    if S not in SUPPORTED_FRAMES:
        raise NotImplementedError(f"frame size {S} not supported (compiled for {SUPPORTED_FRAMES})")
    if lds_only and S not in LDS_FRAMES:
        raise NotImplementedError(f"frame size {S}: this operation takes frames of {LDS_FRAMES} px "
                                  "(256 px frames run gradient descent only)")


def kernel_index(S):
    """Natural frequency index held at each kernel-layout position."""
    check_frame(S)
    N1, N2 = FRAME_RADIX[S]
    p = np.arange(S)
    return p // N2 + N1 * (p % N2)


def fft2(x, inverse=False):
    """Batched 2-D FFT of (..., S, S) complex64 frames (numpy semantics)."""
    S = x.shape[-1]
    check_frame(S)
    _req(x, torch.complex64, name="x")
    if x.shape[-2] != S:
        raise ValueError("frames must be square")
    out = torch.empty_like(x)
    batch = x.numel() // (S * S)
    _lib.call("pty_fft2_c64", ptr(x), ptr(out), batch, S, int(bool(inverse)), stream_handle())
    return out


def to_kernel_layout(frames, inverse=False):
    """Reorder real (..., S, S) float32 frames to (or from) kernel layout."""
    S = frames.shape[-1]
    check_frame(S)
    _req(frames, torch.float32, name="frames")
    out = torch.empty_like(frames)
    _lib.call("pty_permute_frames", ptr(frames), ptr(out), frames.numel() // (S * S), S,
              1 if inverse else 0, stream_handle())
    return out


def probe_spectrum(probe):
    S = probe.shape[-1]
    check_frame(S)
    _req(probe, torch.complex64, (S, S), "probe")
    out = torch.empty_like(probe)
    _lib.call("pty_probe_spectrum", ptr(probe), ptr(out), S, stream_handle())
    return out


def forward_adjoint(probe_hat, obj, positions, amplitudes_k, order, with_far=False):
    """Fused forward pass for the slots of ``order``; returns
    (P_shift, chi, err[, far]) with P_shift/chi natural (B, S, S) and far
    in the staging layout (far_to_kernel_layout converts it)."""
    S = probe_hat.shape[-1]
    check_frame(S)
    _req(probe_hat, torch.complex64, (S, S), "probe_hat")
    _req(obj, torch.complex64, name="object")
    N = positions.shape[0]
    _req(positions, torch.float32, (N, 2), "positions")
    _req(amplitudes_k, torch.float32, (N, S, S), "amplitudes")
    _req(order, torch.int32, name="order")
    B = order.numel()
    Px, Py = obj.shape
    ps = torch.empty((B, S, S), dtype=torch.complex64, device=obj.device)
    chi = torch.empty_like(ps)
    err = torch.empty(B, dtype=torch.float64, device=obj.device)
    far = torch.empty_like(ps) if with_far else None
    _lib.call("pty_forward_adjoint_batch", ptr(probe_hat), ptr(obj), Px, Py, ptr(positions), N,
              ptr(amplitudes_k), ptr(order), B, S, ptr(ps), ptr(chi), ptr(err), ptr(far),
              stream_handle())
    return (ps, chi, err, far) if with_far else (ps, chi, err)


def num_tiles(Px, Py):
    """Number of object tiles of the deterministic scatter (tilemax length)."""
    return int(_lib.load().pty_scatter_num_tiles(int(Px), int(Py)))


def scatter(ps, chi, positions, order, object_shape):
    """(scatter(conj(ps) chi), scatter(|ps|^2), per-tile max)."""
    B, S, _ = ps.shape
    Px, Py = object_shape
    dev = ps.device
    grad = torch.empty((Px, Py), dtype=torch.complex64, device=dev)
    spsum = torch.empty((Px, Py), dtype=torch.float32, device=dev)
    tmax = torch.empty(num_tiles(Px, Py), dtype=torch.float32, device=dev)
    _lib.call("pty_scatter_patches", ptr(ps), ptr(chi), ptr(positions), positions.shape[0],
              ptr(order), B, S, Px, Py, ptr(grad), ptr(spsum), ptr(tmax), stream_handle())
    return grad, spsum, tmax


def object_update(obj, grad, spsum, tilemax, step_size, normalization_min):
    Px, Py = obj.shape
    _lib.call("pty_object_update", ptr(obj), ptr(grad), ptr(spsum), ptr(tilemax), tilemax.numel(),
              Px, Py, float(step_size), float(normalization_min), stream_handle())
    return obj


def probe_update(probe, chi, obj, positions, order, step_size, normalization_min):
    S = probe.shape[-1]
    B = order.numel()
    nbytes = _lib.load().pty_probe_partials_bytes(S, B)
    part = torch.empty(nbytes, dtype=torch.uint8, device=probe.device)
    Px, Py = obj.shape
    _lib.call("pty_probe_update", ptr(probe), ptr(chi), ptr(obj), Px, Py, ptr(positions),
              positions.shape[0], ptr(order), B, S, float(step_size), float(normalization_min),
              ptr(part), stream_handle())
    return probe


def gaussian_kernel1d(sigma, truncate=4.0):
    """The float64 kernel scipy.ndimage.gaussian_filter uses for one axis
    (radius int(truncate * sigma + 0.5), exp(-x^2 / (2 sigma^2)) normalised to
    unit sum), evaluated with NumPy as scipy does; returns (weights, radius)."""
    sigma = float(sigma)
    radius = int(truncate * sigma + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return phi / phi.sum(), radius


def object_gaussian(obj, sigma_px, pure_phase_object=False):
    """In-place gaussian smoothing of a complex64 object (sigma in pixels), as
    scipy.ndimage.gaussian_filter(obj, sigma_px) (or of its phase)."""
    Px, Py = obj.shape
    _req(obj, torch.complex64, name="object")
    if not float(sigma_px) > 1e-15:  # scipy leaves such axes unfiltered
        return obj
    w, r = gaussian_kernel1d(sigma_px)
    wd = torch.from_numpy(w).to(obj.device)
    scratch = torch.empty(Px * Py * 8, dtype=torch.uint8, device=obj.device)
    _lib.call("pty_object_gaussian", ptr(obj), Px, Py, ptr(wd), r, int(bool(pure_phase_object)),
              ptr(scratch), stream_handle())
    return obj


def object_butterworth(obj, q_lowpass, q_highpass, butterworth_order, sampling):
    """In-place Butterworth filter of a complex64 object (cut-offs in 1/A,
    None or 0 = that factor off), as the reference's
    _object_butterworth_constraint."""
    Px, Py = obj.shape
    _req(obj, torch.complex64, name="object")
    scratch = torch.empty(Px * Py * 8, dtype=torch.uint8, device=obj.device)
    part = torch.empty(_lib.load().pty_constraint_partials_bytes(), dtype=torch.uint8,
                       device=obj.device)
    _lib.call("pty_object_butterworth", ptr(obj), Px, Py, float(q_lowpass or 0.0),
              float(q_highpass or 0.0), float(butterworth_order), float(sampling[0]),
              float(sampling[1]), ptr(scratch), ptr(part), stream_handle())
    return obj


def object_constraints(obj, mask_inverse=None, shrinkage_rad=0.0, pure_phase_object=False):
    Px, Py = obj.shape
    part = torch.empty(_lib.load().pty_constraint_partials_bytes(), dtype=torch.uint8,
                       device=obj.device)
    if mask_inverse is not None:
        _req(mask_inverse, torch.uint8, (Px, Py), "mask")
    _lib.call("pty_object_constraints", ptr(obj), Px, Py, ptr(mask_inverse), float(shrinkage_rad),
              int(bool(pure_phase_object)), ptr(part), stream_handle())
    return obj


def probe_com_constraint(probe):
    _lib.call("pty_probe_com_constraint", ptr(probe), probe.shape[-1], stream_handle())
    return probe


def make_problem(obj, probe, positions, amplitudes_k, mask_inverse, mean_intensity,
                 fourier_mask_k=None):
    """pty_problem for device tensors; a float32 object is a potential
    (PTY_OBJECT_POTENTIAL), a complex64 one a transmission."""
    S = probe.shape[-1]
    Px, Py = obj.shape
    if obj.dtype not in (torch.complex64, torch.float32):
        raise TypeError(f"object must be complex64 or float32 (potential), not {obj.dtype}")
    return PtyProblem(S, Px, Py, positions.shape[0], c_void_p(obj.data_ptr()),
                      c_void_p(probe.data_ptr()), c_void_p(positions.data_ptr()),
                      c_void_p(amplitudes_k.data_ptr()),
                      c_void_p(mask_inverse.data_ptr()) if mask_inverse is not None else None,
                      float(mean_intensity),
                      c_void_p(fourier_mask_k.data_ptr()) if fourier_mask_k is not None else None,
                      1 if obj.dtype == torch.float32 else 0)


def gd_workspace_bytes(problem, params):
    import ctypes
    return int(_lib.load().pty_gd_workspace_bytes(ctypes.byref(problem), ctypes.byref(params)))


def gd_run(problem, params, orders, errors, workspace):
    import ctypes
    _lib.call("pty_gd_run", ctypes.byref(problem), ctypes.byref(params), ptr(orders), ptr(errors),
              ptr(workspace), workspace.numel(), stream_handle())


def projection_params(exit_waves, a, b, c, exit_valid=False):
    """pty_projection_params for device exit waves [N, S, S] (complex64)."""
    from ._lib import PtyProjectionParams
    _req(exit_waves, torch.complex64, name="exit_waves")
    return PtyProjectionParams(float(a), float(b), float(c), c_void_p(exit_waves.data_ptr()),
                               int(bool(exit_valid)))


def projection_workspace_bytes(problem, params):
    import ctypes
    return int(_lib.load().pty_projection_workspace_bytes(ctypes.byref(problem),
                                                          ctypes.byref(params)))


def projection_run(problem, params, proj_params, order, errors, workspace):
    import ctypes
    _lib.call("pty_projection_run", ctypes.byref(problem), ctypes.byref(params),
              ctypes.byref(proj_params), ptr(order), ptr(errors), ptr(workspace),
              workspace.numel(), stream_handle())


def projection_positions_workspace_bytes(problem, params):
    import ctypes
    return int(_lib.load().pty_projection_positions_workspace_bytes(ctypes.byref(problem),
                                                                    ctypes.byref(params)))


def projection_run_positions(problem, params, proj_params, position_params, order, errors,
                             workspace):
    import ctypes
    _lib.call("pty_projection_run_positions", ctypes.byref(problem), ctypes.byref(params),
              ctypes.byref(proj_params), ctypes.byref(position_params), ptr(order), ptr(errors),
              ptr(workspace), workspace.numel(), stream_handle())


def projection_batch(probe_hat, obj, positions, amplitudes_k, order, proj_params):
    """Exit-wave step of the projection sets for the slots of ``order``
    (exit waves of proj_params advanced in place); returns (P_shift, err)."""
    import ctypes
    S = probe_hat.shape[-1]
    check_frame(S, lds_only=True)
    B = order.numel()
    Px, Py = obj.shape
    ps = torch.empty((B, S, S), dtype=torch.complex64, device=obj.device)
    err = torch.empty(B, dtype=torch.float64, device=obj.device)
    _lib.call("pty_projection_batch", ptr(probe_hat), ptr(obj), Px, Py, ptr(positions),
              positions.shape[0], ptr(amplitudes_k), ptr(order), B, S, ctypes.byref(proj_params),
              ptr(ps), ptr(err), stream_handle())
    return ps, err


def mixed_workspace_bytes(problem, params, num_probes):
    import ctypes
    return int(_lib.load().pty_mixed_workspace_bytes(ctypes.byref(problem), ctypes.byref(params),
                                                     int(num_probes)))


def mixed_run(problem, params, num_probes, orders, errors, workspace, position_params=None):
    import ctypes
    if position_params is None:
        _lib.call("pty_mixed_run", ctypes.byref(problem), ctypes.byref(params), int(num_probes),
                  ptr(orders), ptr(errors), ptr(workspace), workspace.numel(), stream_handle())
    else:
        _lib.call("pty_mixed_run_positions", ctypes.byref(problem), ctypes.byref(params),
                  int(num_probes), ctypes.byref(position_params), ptr(orders), ptr(errors),
                  ptr(workspace), workspace.numel(), stream_handle())


def mixed_positions_workspace_bytes(problem, params, num_probes):
    import ctypes
    return int(_lib.load().pty_mixed_positions_workspace_bytes(
        ctypes.byref(problem), ctypes.byref(params), int(num_probes)))


def mixed_forward_adjoint(probe_hats, obj, positions, amplitudes_k, order):
    """Mixed-state forward pass for the slots of ``order`` with K probe
    spectra [K, S, S] (kernel layout); returns (P_shift [K, B, S, S],
    chi [K, B, S, S], err [B], far [B, K, S, S] staging layout)."""
    Kp, S = probe_hats.shape[0], probe_hats.shape[-1]
    check_frame(S, lds_only=True)
    _req(probe_hats, torch.complex64, (Kp, S, S), "probe_hats")
    B = order.numel()
    Px, Py = obj.shape
    ps = torch.empty((Kp, B, S, S), dtype=torch.complex64, device=obj.device)
    chi = torch.empty_like(ps)
    far = torch.empty((B, Kp, S, S), dtype=torch.complex64, device=obj.device)
    err = torch.empty(B, dtype=torch.float64, device=obj.device)
    _lib.call("pty_mixed_forward_adjoint_batch", ptr(probe_hats), Kp, ptr(obj), Px, Py,
              ptr(positions), positions.shape[0], ptr(amplitudes_k), ptr(order), B, S, ptr(ps),
              ptr(chi), ptr(err), ptr(far), stream_handle())
    return ps, chi, err, far


def probe_com_modes(probe):
    """Shift every mode of probe [K, S, S] (in place) by mode 0's centre of mass."""
    Kp, S = probe.shape[0], probe.shape[-1]
    check_frame(S, lds_only=True)
    shift = torch.empty(2, dtype=torch.float32, device=probe.device)
    _lib.call("pty_probe_com_modes", ptr(probe), Kp, S, ptr(shift), stream_handle())


def far_to_kernel_layout(far):
    """(B, S, S) far field from the staging layout of forward_adjoint to the
    kernel layout (element j of row-pass thread t -> row t // (S/N1),
    column N1 * (t % (S/N1)) + j)."""
    S = far.shape[-1]
    if S == 256:  # staged in natural order
        return far.contiguous()
    N1 = FRAME_RADIX[S][0]
    x = far.reshape(far.shape[0], N1, S, S // N1)  # [b][j][row][q], t = row * (S/N1) + q
    return x.permute(0, 2, 3, 1).reshape(far.shape[0], S, S).contiguous()


def position_params(positions, positions_initial, step_size, max_update_px=None,
                    max_total_px=None, center_of_mass=False, initial_com=(0.0, 0.0),
                    singular=None):
    """pty_position_params for device positions [N, 2] (updated in place);
    limits in object pixels, None for no limit."""
    from ._lib import PtyPositionParams
    pp = PtyPositionParams()
    pp.positions = c_void_p(positions.data_ptr())
    pp.positions_initial = c_void_p(positions_initial.data_ptr())
    pp.step_size = float(step_size)
    pp.clip_update = int(max_update_px is not None)
    pp.max_update = float(max_update_px) if max_update_px is not None else 0.0
    pp.clip_total = int(max_total_px is not None)
    pp.max_total = float(max_total_px) if max_total_px is not None else 0.0
    pp.center_of_mass = int(bool(center_of_mass))
    pp.initial_com[0] = float(initial_com[0])
    pp.initial_com[1] = float(initial_com[1])
    pp.singular = c_void_p(singular.data_ptr()) if singular is not None else None
    return pp


def gd_positions_workspace_bytes(problem, params):
    import ctypes
    return int(_lib.load().pty_gd_positions_workspace_bytes(ctypes.byref(problem),
                                                            ctypes.byref(params)))


def gd_run_positions(problem, params, position_params, orders, errors, workspace):
    import ctypes
    _lib.call("pty_gd_run_positions", ctypes.byref(problem), ctypes.byref(params),
              ctypes.byref(position_params), ptr(orders), ptr(errors), ptr(workspace),
              workspace.numel(), stream_handle())


def position_correction(position_params, ps, far, obj, amplitudes_k, order, sums=None):
    """One batch of position correction (pty_position_correction_batch)."""
    import ctypes
    S = ps.shape[-1]
    Px, Py = obj.shape
    _lib.call("pty_position_correction_batch", ctypes.byref(position_params), ptr(ps), ptr(far),
              ptr(obj), Px, Py, ptr(amplitudes_k), amplitudes_k.shape[0], ptr(order),
              order.numel(), S, ptr(sums), stream_handle())


def positions_com(positions, initial_com):
    import ctypes
    com = (ctypes.c_float * 2)(float(initial_com[0]), float(initial_com[1]))
    _lib.call("pty_positions_com", ptr(positions), positions.shape[0], com, stream_handle())


__all__ = ["fft2", "to_kernel_layout", "probe_spectrum", "forward_adjoint", "scatter",
           "object_update", "probe_update", "object_constraints", "probe_com_constraint",
           "make_problem", "gd_workspace_bytes", "gd_run", "PtyProblem", "PtyGdParams",
           "kernel_index", "num_tiles", "position_params", "gd_positions_workspace_bytes",
           "gd_run_positions", "position_correction", "positions_com", "projection_params",
           "projection_workspace_bytes", "projection_run", "projection_batch",
           "far_to_kernel_layout", "mixed_workspace_bytes", "mixed_run", "mixed_forward_adjoint",
           "probe_com_modes"]


def pattern_com(frames, mask=None):
    """Per-pattern centre of mass (row, column) of float32 [N, S, S] frames on
    the device; returns two float32 NumPy arrays."""
    _req(frames, torch.float32, name="frames")
    N, S, _ = frames.shape
    m = None
    if mask is not None:
        m = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.float32)).to(frames.device)
    cx = torch.empty(N, dtype=torch.float32, device=frames.device)
    cy = torch.empty_like(cx)
    _lib.call("pty_pattern_com", ptr(frames), N, S, ptr(m), ptr(cx), ptr(cy), stream_handle())
    return cx.cpu().numpy(), cy.cpu().numpy()


def shift_amplitudes(frames, shift_x, shift_y, select=None):
    """Kernel-layout amplitudes of the bilinearly shifted patterns and the
    float64 per-pattern sums (see pty_shift_amplitudes)."""
    check_frame(frames.shape[-1])
    _req(frames, torch.float32, name="frames")
    n_in, S, _ = frames.shape
    dev = frames.device
    sx = torch.from_numpy(np.ascontiguousarray(shift_x, dtype=np.float32)).to(dev)
    sy = torch.from_numpy(np.ascontiguousarray(shift_y, dtype=np.float32)).to(dev)
    if sx.numel() != n_in or sy.numel() != n_in:
        raise ValueError("one shift per input pattern is required")
    sel = None if select is None else torch.from_numpy(np.ascontiguousarray(select, dtype=np.int32)).to(dev)
    n_out = n_in if sel is None else sel.numel()
    amps = torch.empty((n_out, S, S), dtype=torch.float32, device=dev)
    sums = torch.empty(n_out, dtype=torch.float64, device=dev)
    _lib.call("pty_shift_amplitudes", ptr(frames), n_in, S, ptr(sel), n_out, ptr(sx), ptr(sy), ptr(amps),
              ptr(sums), stream_handle())
    return amps, sums
