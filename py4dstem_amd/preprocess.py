"""One-time host-side set-up for the reconstruction (``preprocess()``).

Runs once per dataset in NumPy/SciPy; the per-iteration work lives in the
HIP library.  Each helper cites the reference code whose behaviour it keeps
(py4DSTEM 0.14.18, ``phase/`` = ``py4DSTEM/process/phase/``).
"""

from __future__ import annotations

import warnings

import numpy as np

from .probe import electron_wavelength_angstrom


# ---------------------------------------------------------------------------
# calibrations  (phase/phase_base_class.py:424-611, 2197-2212)
# ---------------------------------------------------------------------------

def extract_calibrations(datacube, energy, require_calibrations=True,
                         force_scan_sampling=None, force_angular_sampling=None,
                         force_reciprocal_sampling=None, verbose=False):
    """Return (scan_sampling, angular_sampling, reciprocal_sampling) tuples.

    Units and error behaviour follow phase_base_class.py:489-609: real space
    in 'A' or 'nm', reciprocal space in 'A^-1' or 'mrad'; 'pixels' raises
    ``ValueError`` when calibrations are required.
    """
    cal = datacube.calibration
    lam = electron_wavelength_angstrom(energy) if energy is not None else None

    if force_scan_sampling is not None:
        scan = (force_scan_sampling,) * 2
    else:
        units = cal.get_R_pixel_units()
        if units == "pixels":
            if require_calibrations:
                raise ValueError("Real-space calibrations must be given in 'A'")
            if verbose:
                warnings.warn("Iterative reconstruction will not be quantitative unless you "
                              "specify real-space calibrations in 'A'", UserWarning)
            scan = (1.0, 1.0)
        elif units == "A":
            scan = (cal.get_R_pixel_size(),) * 2
        elif units == "nm":
            scan = (cal.get_R_pixel_size() * 10,) * 2
        else:
            raise ValueError(f"Real-space calibrations must be given in 'A', not {units}")

    angular = reciprocal = None
    if force_angular_sampling is not None or force_reciprocal_sampling is not None:
        if force_angular_sampling is not None and force_reciprocal_sampling is not None:
            raise ValueError("Only one of angular or reciprocal calibration can be forced.")
        if force_angular_sampling is not None:
            angular = (force_angular_sampling,) * 2
            if lam is not None:
                reciprocal = (force_angular_sampling / lam / 1e3,) * 2
        else:
            reciprocal = (force_reciprocal_sampling,) * 2
            if lam is not None:
                angular = (force_reciprocal_sampling * lam * 1e3,) * 2
    else:
        units = cal.get_Q_pixel_units()
        if units == "pixels":
            if require_calibrations:
                raise ValueError("Reciprocal-space calibrations must be given in in 'A^-1' or 'mrad'")
            angular = reciprocal = (1.0, 1.0)
        elif units == "A^-1":
            reciprocal = (cal.get_Q_pixel_size(),) * 2
            if lam is not None:
                angular = (reciprocal[0] * lam * 1e3,) * 2
        elif units == "mrad":
            angular = (cal.get_Q_pixel_size(),) * 2
            if lam is not None:
                reciprocal = (angular[0] / lam / 1e3,) * 2
        else:
            raise ValueError("Reciprocal-space calibrations must be given in 'A^-1' or 'mrad', "
                             f"not {units}")
    return scan, angular, reciprocal


def real_space_sampling(energy, angular_sampling, frame_shape):
    """Real-space pixel size in Angstrom (phase_base_class.py:2202-2212)."""
    lam = electron_wavelength_angstrom(energy)
    return tuple(lam * 1e3 / dk / n for dk, n in zip(angular_sampling, frame_shape))


# ---------------------------------------------------------------------------
# centre of mass and origin fit  (phase_base_class.py:613-767;
# process/calibration/origin.py:76-190; process/fit/fit.py:38-128)
# ---------------------------------------------------------------------------

FIT_FUNCTIONS = ("plane", "parabola", "bezier_two", "constant")


def _origin_basis(name, x, y):
    """Basis functions of the origin models (py4DSTEM/process/fit/fit.py:
    127-162): the fitted surface is their linear combination."""
    if name == "plane":
        return [x, y, np.ones_like(x)]
    if name == "parabola":
        return [np.ones_like(x), x, x ** 2, y, y ** 2, x * y]
    # quadratic Bernstein polynomials in x times those in y
    bx = [(1 - x) ** 2, 2 * (1 - x) * x, x ** 2]
    by = [(1 - y) ** 2, 2 * (1 - y) * y, y ** 2]
    return [bx[i] * by[j] for i in range(3) for j in range(3)]


def fit_origin(com, mask=None, fit_function="plane"):
    """Fitted origin map of a centre-of-mass map (py4DSTEM/process/calibration/
    origin.py:76-190, non-robust), coordinates normalised to [0, 1].

    'constant' is the mean of the whole map (the mask is not used, as in the
    reference).  The other models are linear in their parameters; the
    reference fits them with ``scipy.optimize.curve_fit``, whose optimum is the
    linear least-squares solution computed here.
    """
    if fit_function not in FIT_FUNCTIONS:
        raise AssertionError(f"fit_function must be one of {FIT_FUNCTIONS}, got {fit_function!r}")
    com = np.asarray(com)
    if fit_function == "constant":
        return np.mean(com) * np.ones_like(com)
    shape = com.shape
    ry, rx = np.meshgrid(np.linspace(0, 1, shape[1]), np.linspace(0, 1, shape[0]))
    sel = np.ones(shape, dtype=bool) if mask is None else mask
    basis = _origin_basis(fit_function, rx, ry)
    design = np.column_stack([b[sel] for b in basis])
    coef = np.linalg.lstsq(design, np.asarray(com[sel], dtype=np.float64), rcond=None)[0]
    return sum(c * b for c, b in zip(coef, basis))



def _curl_div(cx, cy, maximize_divergence):
    """Mean |curl| or |div| of rotated CoM fields, leading axis = angle."""
    if maximize_divergence:
        gxx = cx[..., 2:, 1:-1] - cx[..., :-2, 1:-1]
        gyy = cy[..., 1:-1, 2:] - cy[..., 1:-1, :-2]
        return np.mean(np.abs(gxx + gyy), axis=(-2, -1))
    gxy = cx[..., 1:-1, 2:] - cx[..., 1:-1, :-2]
    gyx = cy[..., 2:, 1:-1] - cy[..., :-2, 1:-1]
    return np.mean(np.abs(gyx - gxy), axis=(-2, -1))


def _rotate_com(nx, ny, theta, transpose):
    if transpose:
        return (np.cos(theta) * ny - np.sin(theta) * nx, np.sin(theta) * ny + np.cos(theta) * nx)
    return (np.cos(theta) * nx - np.sin(theta) * ny, np.sin(theta) * nx + np.cos(theta) * ny)


def solve_com_rotation(com_norm_x, com_norm_y, rotation_angles_deg=None,
                       maximize_divergence=False, force_com_rotation=None,
                       force_com_transpose=None, verbose=False):
    """Relative scan/detector rotation and transpose by curl minimisation
    (or divergence maximisation) of the CoM field.

    Follows the four branches of phase_base_class.py:769-1345 (plots
    omitted).  Returns (rotation_rad, transpose, com_x, com_y).
    """
    if rotation_angles_deg is None:
        rotation_angles_deg = np.arange(-89.0, 90.0, 1.0)
    better = np.argmax if maximize_divergence else np.argmin

    if force_com_rotation is not None:
        rot = np.deg2rad(force_com_rotation)
        if verbose:
            warnings.warn(f"Best fit rotation forced to {force_com_rotation:.0f} degrees.",
                          UserWarning)
        if force_com_transpose is not None:
            transpose = force_com_transpose
            if verbose:
                warnings.warn(f"Transpose of intensities forced to {force_com_transpose}.",
                              UserWarning)
        else:
            plain = _curl_div(*_rotate_com(com_norm_x, com_norm_y, rot, False),
                              maximize_divergence)
            trans = _curl_div(*_rotate_com(com_norm_x, com_norm_y, rot, True),
                              maximize_divergence)
            transpose = bool(trans > plain) if maximize_divergence else bool(trans < plain)
            if verbose and transpose:
                warnings.warn("Diffraction intensities should be transposed.", UserWarning)
    else:
        angles_deg = np.asarray(rotation_angles_deg, dtype=np.float32)
        angles = np.deg2rad(angles_deg)[:, None, None]
        if force_com_transpose is not None:
            transpose = force_com_transpose
            score = _curl_div(*_rotate_com(com_norm_x[None], com_norm_y[None], angles, transpose),
                              maximize_divergence)
            k = int(better(score))
        else:
            s0 = _curl_div(*_rotate_com(com_norm_x[None], com_norm_y[None], angles, False),
                           maximize_divergence)
            s1 = _curl_div(*_rotate_com(com_norm_x[None], com_norm_y[None], angles, True),
                           maximize_divergence)
            k0, k1 = int(better(s0)), int(better(s1))
            if maximize_divergence:
                transpose = not (s0[k0] >= s1[k1])
            else:
                transpose = not (s0[k0] <= s1[k1])
            k = k1 if transpose else k0
        rot = np.squeeze(angles)[k] if angles.size > 1 else np.squeeze(angles)
        if verbose:
            warnings.warn(f"Best fit rotation = {angles_deg[k]:.0f} degrees.", UserWarning)
            if force_com_transpose is None and transpose:
                warnings.warn("Diffraction intensities should be transposed.", UserWarning)
    com_x, com_y = _rotate_com(com_norm_x, com_norm_y, rot, transpose)
    return rot, transpose, com_x, com_y


# ---------------------------------------------------------------------------
# scan positions  (phase_base_class.py:1811-1903; phase/utils.py:675-731)
# ---------------------------------------------------------------------------

def _rotate_about(positions, angle, origin):
    """Rotation part of ``AffineTransform(angle)(x, origin)`` (float32 matrix)."""
    c, s = np.cos(angle), np.sin(angle)
    matrix = (np.eye(2, dtype=np.float32) @ np.eye(2, dtype=np.float32)
              @ np.array([[c, -s], [s, c]], dtype=np.float32))
    origin = np.asarray(origin, dtype=np.float32)
    return ((positions - origin) @ matrix) + (np.array((0.0, 0.0)) + origin)


def scan_positions_in_pixels(positions_ang, grid_scan_shape, scan_sampling, rotation_rad,
                             transpose, sampling, roi_shape, object_padding_px=None,
                             positions_mask=None, positions_offset_ang=None):
    """Probe positions in object pixels (float64) and the padding used."""
    if positions_ang is None:
        nx, ny = grid_scan_shape
        x = np.arange(nx) * scan_sampling[0]
        y = np.arange(ny) * scan_sampling[1]
        x, y = np.meshgrid(x, y, indexing="ij")
        if positions_offset_ang is not None:
            x += positions_offset_ang[0]
            y += positions_offset_ang[1]
        if positions_mask is not None:
            x, y = x[positions_mask], y[positions_mask]
        positions = np.stack((x.ravel(), y.ravel()), axis=-1)
    else:
        positions = np.array(positions_ang, dtype=np.float64, copy=True)
    if rotation_rad is not None:
        positions = _rotate_about(positions, rotation_rad, positions.mean(0))
    if transpose:
        positions = np.flip(positions, 1)
        sampling = sampling[::-1]
    positions = positions - np.min(positions, axis=0).clip(-np.inf, 0)
    positions[:, 0] /= sampling[0]
    positions[:, 1] /= sampling[1]
    roi_shape = np.asarray(roi_shape)
    if object_padding_px is None:
        half = roi_shape / 2
        object_padding_px = (half, half)
    elif np.isscalar(object_padding_px[0]):
        object_padding_px = ((object_padding_px[0],) * 2, (object_padding_px[1],) * 2)
    positions[:, 0] += object_padding_px[0][0]
    positions[:, 1] += object_padding_px[1][0]
    return positions, object_padding_px


def initial_object_shape(positions_px, object_padding_px, roi_shape, object_fov_ang=None,
                         sampling=None):
    """Default object size (ptychographic_methods.py:40-79)."""
    if object_fov_ang is not None:
        return tuple(int(v) for v in np.ceil(np.array(object_fov_ang) / np.array(sampling)))
    pad_x = object_padding_px[0][1]
    pad_y = object_padding_px[1][1]
    p, q = np.round(np.max(positions_px, axis=0))
    p = int(np.max([np.round(p + pad_x), roi_shape[0]]))
    q = int(np.max([np.round(q + pad_y), roi_shape[1]]))
    return (p, q)


def center_positions(positions_px32, object_shape):
    """Shift float32 positions so their mean is the object centre
    (singleslice_ptychography.py:462-478)."""
    com = positions_px32.mean(0)
    positions_px32 -= com - np.array(object_shape) / 2
    return positions_px32


def center_of_mass(ar):
    """Intensity centre of mass in pixel coordinates (rows, columns) of a
    2-D array (py4DSTEM/process/utils/utils.py:166-187, not corner-centred):
    integer index grids times the array, summed, over the array's own sum."""
    ar = np.asarray(ar)
    nx, ny = ar.shape
    ry, rx = np.meshgrid(np.arange(ny), np.arange(nx))
    total = np.sum(ar)
    return np.sum(rx * ar) / total, np.sum(ry * ar) / total


def bilinear_shift(ar, xshift, yshift):
    """Periodic bilinear shift by (xshift, yshift) pixels
    (py4DSTEM/preprocess/utils.py:59-121 with bilinear=True): the four
    integer rolls around the shift weighted by the fractional parts."""
    ar = np.asarray(ar)
    xf = int(np.floor(xshift))
    yf = int(np.floor(yshift))
    wx = xshift - xf
    wy = yshift - yf
    return (np.roll(ar, (xf, yf), axis=(0, 1)) * ((1 - wx) * (1 - wy))
            + np.roll(ar, (xf + 1, yf), axis=(0, 1)) * (wx * (1 - wy))
            + np.roll(ar, (xf, yf + 1), axis=(0, 1)) * ((1 - wx) * wy)
            + np.roll(ar, (xf + 1, yf + 1), axis=(0, 1)) * (wx * wy))


class ScaledQCalibration:
    """A datacube's calibration with the reciprocal pixel size multiplied by
    `factor`: the calibration the reference's resampling leaves on the
    datacube (preprocess.py:209-213, 660-663, 708-710), without modifying the
    caller's object."""

    def __init__(self, datacube, factor):
        self._cal = datacube.calibration
        self._factor = float(factor)
        self.calibration = self

    def get_Q_pixel_size(self):
        return self._cal.get_Q_pixel_size() * self._factor

    def __getattr__(self, name):
        return getattr(self._cal, name)
