"""ctypes binding of libpty_hip.so (declarations in include/pty_hip.h).

torch is imported first so that the library's libamdhip64.so.7 dependency
resolves to the HIP runtime torch already loaded (same SONAME), which lets
the kernels run on torch's streams and device allocations.  There is no
fallback: if the library or a GPU is missing, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede loading the HIP library)

from ._build import LIB_PATH

c_int, c_float, c_size_t, c_void_p, c_char_p = (ctypes.c_int, ctypes.c_float, ctypes.c_size_t,
                                                ctypes.c_void_p, ctypes.c_char_p)
c_double_p = ctypes.POINTER(ctypes.c_double)


class PtyProblem(ctypes.Structure):
    _fields_ = [("S", c_int), ("Px", c_int), ("Py", c_int), ("N", c_int),
                ("object", c_void_p), ("probe", c_void_p), ("positions", c_void_p),
                ("amplitudes", c_void_p), ("fov_mask_inverse", c_void_p),
                ("mean_intensity", c_float), ("fourier_mask", c_void_p),
                ("object_type", c_int)]


class PtyGdParams(ctypes.Structure):
    _fields_ = [("num_iter", c_int), ("max_batch", c_int), ("fix_probe", c_int),
                ("fix_probe_com", c_int), ("pure_phase_object", c_int),
                ("step_size", c_float), ("normalization_min", c_float),
                ("shrinkage_rad", c_float), ("stage", c_int),
                ("fix_probe_aperture", c_int), ("probe_aperture", c_void_p),
                ("constrain_probe_amplitude", c_int), ("probe_amplitude_radius", ctypes.c_double),
                ("probe_amplitude_width", ctypes.c_double),
                ("gaussian_weights", c_void_p), ("gaussian_radius", c_int),
                ("butterworth", c_int), ("q_lowpass", ctypes.c_double),
                ("q_highpass", ctypes.c_double), ("butterworth_order", ctypes.c_double),
                ("sampling", ctypes.c_double * 2), ("object_positivity", c_int),
                ("pipeline", c_int)]


class PtyPositionParams(ctypes.Structure):
    _fields_ = [("positions", c_void_p), ("positions_initial", c_void_p), ("step_size", c_float),
                ("clip_update", c_int), ("max_update", ctypes.c_double), ("clip_total", c_int),
                ("max_total", ctypes.c_double), ("center_of_mass", c_int),
                ("initial_com", c_float * 2), ("singular", c_void_p)]


class PtyProjectionParams(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double),
                ("exit_waves", c_void_p), ("exit_valid", c_int)]


class PtyConvDesc(ctypes.Structure):
    _fields_ = [("n", c_int), ("Hi", c_int), ("Wi", c_int), ("Ci", c_int), ("Ho", c_int),
                ("Wo", c_int), ("Co", c_int), ("stride", c_int), ("dil", c_int)]


class PtyMultiProblem(ctypes.Structure):
    _fields_ = [("problem", PtyProblem), ("params", PtyGdParams), ("orders", c_void_p),
                ("errors", c_void_p)]


_SIGNATURES = {
    "pty_version": (c_int, []),
    "pty_last_error": (c_char_p, []),
    "pty_fft2_c64": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "pty_permute_frames": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "pty_probe_spectrum": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "pty_forward_adjoint_batch": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                                          c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p]),
    "pty_scatter_num_tiles": (c_int, [c_int, c_int]),
    "pty_scatter_patches": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int,
                                    c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_object_update": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                  c_float, c_float, c_void_p]),
    "pty_probe_partials_bytes": (c_size_t, [c_int, c_int]),
    "pty_probe_update": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                                 c_void_p, c_int, c_int, c_float, c_float, c_void_p, c_void_p]),
    "pty_constraint_partials_bytes": (c_size_t, []),
    "pty_object_constraints": (c_int, [c_void_p, c_int, c_int, c_void_p, c_float, c_int,
                                       c_void_p, c_void_p]),
    "pty_probe_com_constraint": (c_int, [c_void_p, c_int, c_void_p]),
    "pty_object_gaussian": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_void_p,
                                    c_void_p]),
    "pty_resample_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "pty_resample_frames": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                    c_void_p, c_void_p, c_void_p]),
    "pty_object_butterworth": (c_int, [c_void_p, c_int, c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       c_void_p, c_void_p, c_void_p]),
    "pty_loss_workspace_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    "pty_loss_gradients": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p,
                                   c_void_p, c_int, c_int, c_int, ctypes.c_double, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_simulate_intensities": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                         c_void_p, c_void_p]),
    "pty_pattern_com": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_shift_amplitudes": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]),
    "pty_multi_workspace_bytes": (c_size_t, [c_void_p, c_int]),
    "pty_multi_run": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_size_t, c_void_p]),
    "pty_conv_weight_words": (c_size_t, [ctypes.POINTER(PtyConvDesc), c_int]),
    "pty_conv_forward": (c_int, [ctypes.POINTER(PtyConvDesc), c_void_p, c_void_p, c_void_p, c_int,
                                 c_void_p, c_void_p, c_void_p]),
    "pty_conv_backward_data": (c_int, [ctypes.POINTER(PtyConvDesc), c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p]),
    "pty_conv_backward_weight_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyConvDesc)]),
    "pty_conv_backward_weight": (c_int, [ctypes.POINTER(PtyConvDesc), c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_conv_adam": (c_int, [ctypes.POINTER(PtyConvDesc), c_float, c_float, c_float, c_float, c_int,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_conv_pack_weights": (c_int, [ctypes.POINTER(PtyConvDesc), c_void_p, c_void_p, c_void_p,
                                      c_void_p]),
    "pty_net_inputs_workspace_bytes": (c_size_t, [c_int]),
    "pty_net_inputs": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                               c_size_t, c_void_p]),
    "pty_net_loss_workspace_bytes": (c_size_t, [c_int]),
    "pty_net_loss": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_size_t,
                             c_void_p]),
    "pty_net_patches": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "pty_net_finish": (c_int, [c_void_p, c_void_p, c_int, c_int, c_float, c_void_p, c_void_p]),
    "pty_gd_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams)]),
    "pty_gd_run": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams), c_void_p,
                           c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_gd_run_noalloc": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams),
                                   c_void_p, c_void_p, c_void_p]),
    "pty_gd_positions_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyProblem),
                                                    ctypes.POINTER(PtyGdParams)]),
    "pty_gd_run_positions": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams),
                                     ctypes.POINTER(PtyPositionParams), c_void_p, c_void_p,
                                     c_void_p, c_size_t, c_void_p]),
    "pty_position_correction_batch": (c_int, [ctypes.POINTER(PtyPositionParams), c_void_p,
                                              c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                                              c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "pty_positions_com": (c_int, [c_void_p, c_int, ctypes.POINTER(c_float), c_void_p]),
    "pty_mixed_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams),
                                             c_int]),
    "pty_mixed_run": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams), c_int,
                              c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_mixed_positions_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyProblem),
                                                       ctypes.POINTER(PtyGdParams), c_int]),
    "pty_mixed_run_positions": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams), c_int,
                                        ctypes.POINTER(PtyPositionParams), c_void_p, c_void_p,
                                        c_void_p, c_size_t, c_void_p]),
    "pty_mixed_forward_adjoint_batch": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p,
                                                c_int, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_probe_com_modes": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "pty_projection_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyProblem),
                                                  ctypes.POINTER(PtyGdParams)]),
    "pty_projection_run": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams),
                                   ctypes.POINTER(PtyProjectionParams), c_void_p, c_void_p,
                                   c_void_p, c_size_t, c_void_p]),
    "pty_projection_positions_workspace_bytes": (c_size_t, [ctypes.POINTER(PtyProblem),
                                                            ctypes.POINTER(PtyGdParams)]),
    "pty_projection_run_positions": (c_int, [ctypes.POINTER(PtyProblem), ctypes.POINTER(PtyGdParams),
                                             ctypes.POINTER(PtyProjectionParams),
                                             ctypes.POINTER(PtyPositionParams), c_void_p, c_void_p,
                                             c_void_p, c_size_t, c_void_p]),
    "pty_projection_batch": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p,
                                     c_void_p, c_int, c_int, ctypes.POINTER(PtyProjectionParams),
                                     c_void_p, c_void_p, c_void_p]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load(path=LIB_PATH):
    """Load (once) and return the ctypes library; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m py4dstem_amd._build` "
            "(or __graft_entry__.build()) first. There is no CPU fallback.")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().pty_last_error().decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("py4dstem_amd needs a ROCm GPU (MI355X); none is visible. "
                           "There is no CPU fallback.")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def ptr(t):
    """Device pointer of a tensor (or None)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return c_void_p(t.data_ptr())


def call(name, *args):
    check(getattr(load(), name)(*args), name)
