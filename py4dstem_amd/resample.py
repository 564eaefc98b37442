"""Resampling and padding of the diffraction patterns on the GPU.

py4DSTEM resamples the datacube before anything else when
``preprocess(diffraction_intensities_shape=..., reshaping_method=...)`` is
given, and zero-pads it for ``padded_diffraction_intensities_shape``
(phase_base_class.py:223-420, with py4DSTEM/preprocess/preprocess.py:155-215
for 'bin', :626-716 for 'fourier' and 'bilinear', :719-790 for the padding,
and py4DSTEM/process/utils/utils.py:412-640 ``fourier_resample``).  Every one
of these operations is linear along each detector axis and acts on the two
axes independently, so a pattern X becomes Re(Mx X My^T).  The per-axis
matrices are built here by applying the reference's one-dimensional
operation to unit vectors (NumPy FFTs for 'fourier', scipy's order-1 zoom for
'bilinear', block sums for 'bin', an embedding for the padding), and the
patterns are mapped on the GPU by ``pty_resample_frames``.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import ptr, stream_handle

METHODS = ("bilinear", "fourier", "bin")


def _fourier_axis(q, s):
    """One axis of ``fourier_resample`` (utils.py:412-640): fft, keep the
    first ceil(m/2) and last floor(m/2) frequencies (m the smaller size) in
    the same places, inverse fft at the new size.  Complex (s, q)."""
    m = min(q, s)
    lo, hi = (m + 1) // 2, m // 2
    spec = np.fft.fft(np.eye(q), axis=0)  # column j: spectrum of unit vector j
    crop = np.zeros((s, q), dtype=np.complex128)
    crop[:lo] = spec[:lo]
    if hi:
        crop[s - hi:] = spec[q - hi:]
    return np.fft.ifft(crop, axis=0)


def _bilinear_axis(q, factor):
    """One axis of the order-1 zoom the reference applies to each pattern
    (preprocess.py:692-707: scipy.ndimage.zoom, mode='nearest',
    grid_mode=True), divided by the factor (conserve_array_sums)."""
    from scipy.ndimage import zoom
    z = zoom(np.eye(q), (factor, 1.0), order=1, mode="nearest", grid_mode=True)
    return z / factor


def _bin_axis(q, b):
    """Block sums of b pixels, trailing q % b pixels dropped (preprocess.py:
    155-215)."""
    s = q // b
    m = np.zeros((s, q))
    for o in range(s):
        m[o, o * b:(o + 1) * b] = 1.0
    return m


def _pad_axis(q, s):
    """Zero padding to s pixels, (s - q) // 2 before (preprocess.py:719-790)."""
    before = (s - q) // 2
    m = np.zeros((s, q))
    m[before + np.arange(q), np.arange(q)] = 1.0
    return m


def axis_maps(shape, diffraction_intensities_shape=None, reshaping_method="bilinear",
              padded_diffraction_intensities_shape=None):
    """(Mx, My, q_pixel_factor): per-axis maps for patterns of `shape` and
    the factor the reciprocal pixel size is multiplied by.  Raises the
    reference's ValueErrors (phase_base_class.py:290-300, 405-409;
    preprocess.py:764-765)."""
    Qx, Qy = (int(v) for v in shape)
    Mx, My = np.eye(Qx), np.eye(Qy)
    factor = 1.0
    if diffraction_intensities_shape is not None:
        Sx, Sy = (int(v) for v in diffraction_intensities_shape)
        rx, ry = Sx / Qx, Sy / Qy
        if rx != ry:
            raise ValueError("Datacube calibration can only handle uniform Q-sampling.")
        if reshaping_method == "bin":
            b = int(1 / rx)
            if b < 1:
                raise ValueError(f"Calculated binning factor {b} is less than 1.")
            Mx, My = _bin_axis(Qx, b), _bin_axis(Qy, b)
            factor = float(b)
        elif reshaping_method == "fourier":
            Mx, My = _fourier_axis(Qx, Sx), _fourier_axis(Qy, Sy)
            factor = 1.0 / rx
        elif reshaping_method == "bilinear":
            Mx, My = _bilinear_axis(Qx, rx), _bilinear_axis(Qy, rx)
            factor = 1.0 / rx
        else:
            raise ValueError("reshaping_method needs to be one of 'bilinear', 'fourier', or "
                             f"'bin', not {reshaping_method}.")
    if padded_diffraction_intensities_shape is not None:
        Px, Py = (int(v) for v in padded_diffraction_intensities_shape)
        if Px < Mx.shape[0] or Py < My.shape[0]:
            raise ValueError(f"'output_size' must be at-least as large as "
                             f"{(Mx.shape[0], My.shape[0])}.")
        Mx = _pad_axis(Mx.shape[0], Px) @ Mx
        My = _pad_axis(My.shape[0], Py) @ My
    return Mx, My, factor


def resample_frames(frames_dev, Mx, My, chunk_bytes=256 << 20):
    """Re(Mx X My^T) for every pattern of the float32 device tensor
    (N, Qx, Qy); returns a new float32 device tensor (N, Sx, Sy)."""
    N, Qx, Qy = frames_dev.shape
    Sx, Sy = Mx.shape[0], My.shape[0]
    if Mx.shape[1] != Qx or My.shape[1] != Qy:
        raise ValueError(f"maps {Mx.shape}, {My.shape} do not fit patterns {(Qx, Qy)}")
    dev = frames_dev.device
    mx = torch.from_numpy(np.ascontiguousarray(Mx, dtype=np.complex64)).to(dev)
    my = torch.from_numpy(np.ascontiguousarray(My, dtype=np.complex64)).to(dev)
    out = torch.empty((N, Sx, Sy), dtype=torch.float32, device=dev)
    per = max(1, min(65535, chunk_bytes // (Sx * Qy * 8)))
    lib = _lib.load()
    ws = torch.empty(int(lib.pty_resample_workspace_bytes(min(per, N), Qy, Sx)), dtype=torch.uint8,
                     device=dev)
    src = frames_dev.contiguous()
    for b0 in range(0, N, per):
        nb = min(per, N - b0)
        _lib.call("pty_resample_frames", ptr(src[b0:b0 + nb]), nb, Qx, Qy, ptr(mx), ptr(my), Sx, Sy,
                  ptr(out[b0:b0 + nb]), ptr(ws), stream_handle())
    return out
