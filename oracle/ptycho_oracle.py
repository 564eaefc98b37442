"""NumPy restatement of py4DSTEM's single-slice gradient-descent ptychography.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference: py4DSTEM 0.14.18.  Abbreviations in citations:
``phase/`` = ``py4DSTEM/process/phase/``.

Precision modes
---------------
``mode="c64"`` is the NumPy >= 2 behaviour (``np.fft`` keeps complex64), which
is what a reference run on this image's NumPy 2.2 computes.  ``mode="c128"``
reproduces NumPy < 2 (the version pinned by the reference's ``setup.py:26``),
where every ``np.fft`` call upcasts to complex128 while the object stays
complex64.  The two modes differ only in ``_fft2`` / ``_ifft2`` below; every
other dtype follows from NumPy's own promotion rules, as in the reference.
"""

from __future__ import annotations

import math
import warnings

import numpy as np
import scipy.ndimage

# Which check in tests/ pins each function (see the package docstring).
PINNING = {
    "electron_wavelength_angstrom": "closed form: lambda(300 keV) = 0.019687 A",
    "ComplexProbeOracle": "closed-form aperture/rolloff and defocus chi; unit norm",
    "fourier_translation_operator": "integer shift == np.roll; shifts compose",
    "fft_shift": "integer shift == np.roll; shifts compose; Parseval",
    "subdivide_into_batches": "hand-derived vectors and sum/ordering invariants",
    "patch_indices": "brute-force loops on tiny grids, round-half-even",
    "scatter_bincount": "brute-force python loop scatter, float64 order",
    "get_shifted_ar": "integer shift == np.roll; bilinear linearity",
    "get_CoM": "closed form on single-pixel and symmetric images",
    "fit_plane": "exact recovery of a noiseless plane",
    "fit_origin_2d": "exact recovery of noiseless parabola / Bezier surfaces; equals "
                     "scipy curve_fit on the reference's model functions",
    "forward": "Parseval, projection fixed point, A=0 => chi=-psi, angle(0)=0; detector mask: "
               "ones = no mask, masked pixels carry no update and count A^2 in the error",
    "adjoint": "noise-free ground truth is a fixed point; FP64 finite differences",
    "loss_gradients": "FP64 central finite differences (object and probe pixels)",
    "object_constraints": "mean phase outside FOV is 0 after shrinkage; |O|<=1",
    "probe_com_constraint": "probe CoM ~ 0 after recentring",
    "probe_aperture_constraint": "output Fourier amplitude proportional to the aperture, "
                                 "Fourier phase and intensity kept",
    "probe_amplitude_constraint": "intensity kept; erf edge value 1/2 at the radius; identity "
                                  "inside for a narrow edge",
    "position_partials": "brute-force DFT loops on tiny frames with np.roll'ed objects; probe "
                         "modes: a zero second mode reproduces the single-mode partials, two "
                         "modes give the sum of each mode's partials",
    "object_constraints (gaussian)": "scipy.ndimage.gaussian_filter (the reference's call); "
                                     "the GPU's summation order reproduces it bit for bit",
    "object_constraints (potential)": "shrinkage leaves the masked mean at -shrinkage_rad; "
                                      "positivity is clip(0)",
    "object_butterworth_constraint": "constant plus plane wave: the wave scaled by the envelope "
                                     "at its frequency (1/2 at the cut-off); zero cut-offs are "
                                     "the identity",
    "forward/adjoint (potential)": "V and exp(iV) give the same exit waves; the one-position "
                                   "potential step is Re(-i conj(exp(iV)) dO_complex)",
    "fourier_resample_4d / resample_datacube": "band-limited pattern resampled to its exact "
                                               "samples x (Q/S)^2; binning and padding closed "
                                               "forms",
    "position_correction": "exact recovery of a planted least-squares step; clipping "
                           "closed forms; sign/magnitude on a sub-pixel displaced scan",
    "positions_constraints": "scan centre of mass restored exactly; exact affine maps "
                             "recovered (with outliers) by the RANSAC fit",
    "affine_ransac": "exact recovery of planted affine maps with 10% outliers",
    "method_parameters": "hand-derived (a, b, c) table and range errors",
    "projection_forward": "AP output meets the modulus constraint; DM = e + P_M(2 psi - e) - psi; "
                          "a consistent exit wave is a fixed point for every (a, b, c)",
    "projection_adjoint": "noise-free truth with normalization_min=0 is reproduced",
    "mixed_probe_modes": "equal moduli; mode-to-mode phase is a linear ramp",
    "probe_orthogonalization": "orthonormal-up-to-scale output, sorted, intensity and span "
                               "preserved",
    "OracleMixedstatePtychography.forward/adjoint": "K=1 equals the single-mode path; data "
                                                    "from the mixed model is a fixed point; FP64 "
                                                    "finite differences of the mixed loss",
    "bilinear_resample": "zoom by 1 returns the input values; the float64 divisor promotes "
                         "complex64 to complex128 (NumPy >= 2 scalar rules)",
    "solve_com_relative_rotation": "curl-free (gradient) CoM fields rotated by a known angle, "
                                   "plain and transposed: the search returns that angle and "
                                   "transpose with zero curl; a radial field rotated by 90 "
                                   "degrees for the divergence branch; the forced-rotation "
                                   "transpose decision on the same fields",
    "reconstruct": "zero object: first error = sum A^2 / (mean_I N) = 1; object c*truth: "
                   "first error = (1 - c)^2 (pins the error normalisation); noise-free truth "
                   "with fixed probe and |O| = 1, zero phase outside the field: every error "
                   "~0 and the object passes the constraint sequence unchanged; one "
                   "iteration equals the separately pinned forward/adjoint/constraint "
                   "functions composed by hand",
    "adjoint (step normalisation)": "one position: normalization_min = 0 gives step*chi/P "
                                    "and step*chi/O_patch, = 1 gives step*conj(P)chi/max|P|^2 "
                                    "and step*conj(O)chi/max|O|^2; several positions at "
                                    "normalization_min = 1: the object step is "
                                    "-step/(2 S^2 max overlap) times the FP64 finite-difference "
                                    "gradient of the loss",
}


# ---------------------------------------------------------------------------
# Numerical utilities
# ---------------------------------------------------------------------------

def electron_wavelength_angstrom(E_eV):
    """Relativistic electron wavelength in Angstrom.

    Restates py4DSTEM/process/utils/utils.py:90-97 (same constants and
    evaluation order).
    """
    m = 9.109383 * 10**-31
    e = 1.602177 * 10**-19
    c = 299792458
    h = 6.62607 * 10**-34
    return (
        h / math.sqrt(2 * m * e * E_eV) / math.sqrt(1 + e * E_eV / 2 / m / c**2) * 10**10
    )


POLAR_SYMBOLS = (
    "C10", "C12", "phi12",
    "C21", "phi21", "C23", "phi23",
    "C30", "C32", "phi32", "C34", "phi34",
    "C41", "phi41", "C43", "phi43", "C45", "phi45",
    "C50", "C52", "phi52", "C54", "phi54", "C56", "phi56",
)  # phase/utils.py:29-35

POLAR_ALIASES = {
    "defocus": "C10", "astigmatism": "C12", "astigmatism_angle": "phi12",
    "coma": "C21", "coma_angle": "phi21", "Cs": "C30", "C5": "C50",
}  # phase/utils.py:38-43


def set_polar_parameters(target: dict, parameters: dict):
    """Alias handling with the defocus sign flip.

    Restates phase/phase_base_class.py:1788-1809 and phase/utils.py:141-160.
    """
    for symbol, value in parameters.items():
        if symbol in target:
            target[symbol] = value
        elif symbol == "defocus":
            target[POLAR_ALIASES[symbol]] = -value
        elif symbol in POLAR_ALIASES:
            target[POLAR_ALIASES[symbol]] = value
        else:
            raise ValueError("{} not a recognized parameter".format(symbol))
    return target


def spatial_frequencies(gpts, sampling):
    """phase/utils.py:444-462."""
    return tuple(np.fft.fftfreq(n, d).astype(np.float32) for n, d in zip(gpts, sampling))


class ComplexProbeOracle:
    """Restatement of ``ComplexProbe`` (phase/utils.py:50-426), CPU only.

    Envelopes (focal/angular/gaussian spread) default to zero in the reference
    and are not used by ``SingleslicePtychography``; they are omitted.
    """

    def __init__(self, energy, gpts, sampling, semiangle_cutoff=np.inf, rolloff=2.0,
                 vacuum_probe_intensity=None, parameters=None, mode="c64"):
        self._energy = energy
        self._wavelength = electron_wavelength_angstrom(energy)
        self._gpts = tuple(int(g) for g in gpts)
        self._sampling = tuple(sampling)
        self._semiangle_cutoff = semiangle_cutoff
        self._rolloff = rolloff
        self._vacuum_probe_intensity = vacuum_probe_intensity
        self._phase_shift = 0.0
        self._mode = mode
        self._parameters = dict(zip(POLAR_SYMBOLS, [0.0] * len(POLAR_SYMBOLS)))
        set_polar_parameters(self._parameters, dict(parameters or {}))

    def evaluate_aperture(self, alpha):
        """phase/utils.py:162-191."""
        semiangle_cutoff = self._semiangle_cutoff / 1000
        if self._vacuum_probe_intensity is not None:
            v = np.asarray(self._vacuum_probe_intensity, dtype=np.float32)
            return np.sqrt(np.maximum(v, 0))
        if self._semiangle_cutoff == np.inf:
            return np.ones_like(alpha)
        if self._rolloff > 0.0:
            rolloff = self._rolloff / 1000.0
            array = 0.5 * (1 + np.cos(np.pi * (alpha - semiangle_cutoff + rolloff) / rolloff))
            array[alpha > semiangle_cutoff] = 0.0
            array = np.where(alpha > semiangle_cutoff - rolloff, array,
                             np.ones_like(alpha, dtype=np.float32))
        else:
            array = np.array(alpha < semiangle_cutoff).astype(np.float32)
        return array

    def evaluate_chi(self, alpha, phi):
        """phase/utils.py:290-368 (including its ``phi41`` duplicate in the
        fourth-order guard tuple, which only affects when the term is added)."""
        p = self._parameters
        alpha2 = alpha**2
        alpha = np.array(alpha)
        array = np.zeros(alpha.shape, dtype=np.float32)
        if any(p[s] != 0.0 for s in ("C10", "C12", "phi12")):
            array += 1 / 2 * alpha2 * (p["C10"] + p["C12"] * np.cos(2 * (phi - p["phi12"])))
        if any(p[s] != 0.0 for s in ("C21", "phi21", "C23", "phi23")):
            array += 1 / 3 * alpha2 * alpha * (
                p["C21"] * np.cos(phi - p["phi21"]) + p["C23"] * np.cos(3 * (phi - p["phi23"]))
            )
        if any(p[s] != 0.0 for s in ("C30", "C32", "phi32", "C34", "phi34")):
            array += 1 / 4 * alpha2**2 * (
                p["C30"] + p["C32"] * np.cos(2 * (phi - p["phi32"]))
                + p["C34"] * np.cos(4 * (phi - p["phi34"]))
            )
        if any(p[s] != 0.0 for s in ("C41", "phi41", "C43", "phi43", "C45", "phi41")):
            array += 1 / 5 * alpha2**2 * alpha * (
                p["C41"] * np.cos(phi - p["phi41"]) + p["C43"] * np.cos(3 * (phi - p["phi43"]))
                + p["C45"] * np.cos(5 * (phi - p["phi45"]))
            )
        if any(p[s] != 0.0 for s in ("C50", "C52", "phi52", "C54", "phi54", "C56", "phi56")):
            array += 1 / 6 * alpha2**3 * (
                p["C50"] + p["C52"] * np.cos(2 * (phi - p["phi52"]))
                + p["C54"] * np.cos(4 * (phi - p["phi54"]))
                + p["C56"] * np.cos(6 * (phi - p["phi56"]))
            )
        return 2 * np.pi / self._wavelength * array + self._phase_shift

    def get_scattering_angles(self):
        """phase/utils.py:401-418."""
        kx, ky = spatial_frequencies(self._gpts, self._sampling)
        x = kx * self._wavelength
        y = ky * self._wavelength
        alpha = np.sqrt(x[:, None] ** 2 + y[None, :] ** 2)
        phi = np.arctan2(y[None, :], x[:, None])
        return alpha, phi

    def evaluate_ctf(self):
        """phase/utils.py:370-399."""
        alpha, phi = self.get_scattering_angles()
        array = np.exp(-1.0j * self.evaluate_chi(alpha, phi))
        if self._semiangle_cutoff < np.inf or self._vacuum_probe_intensity is not None:
            array *= self.evaluate_aperture(alpha)
        return array

    def build(self):
        """Corner-centred probe, unit L2 norm (phase/utils.py:420-426)."""
        array = _ifft2(self.evaluate_ctf(), self._mode)
        array = array / np.sqrt((np.abs(array) ** 2).sum())
        self._array = array
        return self


def _fft2(x, mode):
    if mode == "c128":
        return np.fft.fft2(np.asarray(x, dtype=np.complex128))
    return np.fft.fft2(x)


def _ifft2(x, mode):
    if mode == "c128":
        return np.fft.ifft2(np.asarray(x, dtype=np.complex128))
    return np.fft.ifft2(x)


def fourier_translation_operator(positions, shape):
    """Phase ramp(s) exp(-2 pi i (kx x + ky y)) (phase/utils.py:468-505)."""
    positions_shape = positions.shape
    if len(positions_shape) == 1:
        positions = positions[None]
    kx, ky = spatial_frequencies(shape, (1.0, 1.0))
    positions = np.asarray(positions, dtype=np.float32)
    x = positions[:, 0].ravel()[:, None, None]
    y = positions[:, 1].ravel()[:, None, None]
    result = np.exp(-2.0j * np.pi * kx[None, :, None] * x) * np.exp(
        -2.0j * np.pi * ky[None, None, :] * y
    )
    if len(positions_shape) == 1:
        return result[0]
    return result


def fft_shift(array, positions, mode="c64"):
    """Fourier shift of ``array`` by ``positions`` (phase/utils.py:508-533)."""
    op = fourier_translation_operator(positions, array.shape[-2:])
    fourier_array = _fft2(array, mode)
    if len(op.shape) == 3 and len(fourier_array.shape) == 3:
        shifted = fourier_array[None] * op[:, None]
    else:
        shifted = fourier_array * op
    return _ifft2(shifted, mode)


def subdivide_into_batches(num_items, num_batches=None, max_batch=None):
    """Near-equal split, larger batches first (phase/utils.py:539-579)."""
    if (num_batches is not None) & (max_batch is not None):
        raise RuntimeError()
    if num_batches is None:
        if max_batch is not None:
            num_batches = (num_items + (-num_items % max_batch)) // max_batch
        else:
            raise RuntimeError()
    if num_items < num_batches:
        raise RuntimeError("num_batches may not be larger than num_items")
    elif num_items % num_batches == 0:
        return [num_items // num_batches] * num_batches
    else:
        v = []
        zp = num_batches - (num_items % num_batches)
        pp = num_items // num_batches
        for i in range(num_batches):
            if i >= zp:
                v = [pp + 1] + v
            else:
                v = [pp] + v
        return v


def generate_batches(num_items, num_batches=None, max_batch=None, start=0):
    """phase/utils.py:582-589."""
    for batch in subdivide_into_batches(num_items, num_batches, max_batch):
        end = start + batch
        yield start, end
        start = end


def affine_rotate(positions, angle, origin):
    """``AffineTransform(angle=angle)(positions, origin)`` with scale 1, no
    shear, no translation (phase/utils.py:675-698, 726-731)."""
    cosx = np.cos(angle)
    sinx = np.sin(angle)
    rot = np.array([[+cosx, -sinx], [+sinx, +cosx]], dtype="float32")
    scale = np.array([[1.0, 0.0], [0.0, 1.0]], dtype="float32")
    shear = np.array([[1.0, 0.0], [0.0, 1.0]], dtype="float32")
    tf_matrix = scale @ shear @ rot
    origin = np.asarray(origin, dtype=np.float32)
    tf_translation = np.array((0.0, 0.0)) + origin
    return ((positions - origin) @ tf_matrix) + tf_translation


def get_CoM(ar, corner_centered=False):
    """Centre of mass (py4DSTEM/process/utils/utils.py:166-187)."""
    ar = np.asarray(ar)
    nx, ny = ar.shape
    if corner_centered:
        ry, rx = np.meshgrid(np.fft.fftfreq(ny, 1 / ny), np.fft.fftfreq(nx, 1 / nx))
    else:
        ry, rx = np.meshgrid(np.arange(ny), np.arange(nx))
    tot = np.sum(ar)
    return np.sum(rx * ar) / tot, np.sum(ry * ar) / tot


def get_shifted_ar(ar, xshift, yshift):
    """Bilinear, periodic shift (py4DSTEM/preprocess/utils.py:59-121 with
    ``periodic=True, bilinear=True``)."""
    ar = np.asarray(ar)
    xF = np.floor(xshift).astype(int).item()
    yF = np.floor(yshift).astype(int).item()
    wx = xshift - xF
    wy = yshift - yF
    return (
        np.roll(ar, (xF, yF), axis=(0, 1)) * ((1 - wx) * (1 - wy))
        + np.roll(ar, (xF + 1, yF), axis=(0, 1)) * ((wx) * (1 - wy))
        + np.roll(ar, (xF, yF + 1), axis=(0, 1)) * ((1 - wx) * (wy))
        + np.roll(ar, (xF + 1, yF + 1), axis=(0, 1)) * ((wx) * (wy))
    )


def fourier_resample_4d(data, output_size):
    """``fourier_resample`` of a 4-D stack with conserve_array_sums=True
    (py4DSTEM/process/utils/utils.py:412-640): per pattern fft2, the four
    low-frequency quadrants copied into a zero complex64 array of the new
    size, ifft2, real part as float32."""
    Qx, Qy = data.shape[-2:]
    Sx, Sy = output_size

    def sl(i, o):
        if o == i:
            return [(slice(None), slice(None))] * 2
        m = min(i, o)
        a, b = (m + 1) // 2, m // 2
        return [(slice(0, a), slice(0, a)), (slice(o - b, o), slice(i - b, i))]

    xs, ys = sl(Qx, Sx), sl(Qy, Sy)
    out = np.zeros(data.shape[:2] + (Sx, Sy), np.float32)
    for rx in range(data.shape[0]):
        for ry in range(data.shape[1]):
            spec = np.fft.fft2(data[rx, ry])
            res = np.zeros((Sx, Sy), dtype=np.complex64)
            for xo, xi in xs:
                for yo, yi in ys:
                    res[xo, yo] = spec[xi, yi]
            out[rx, ry] = np.real(np.fft.ifft2(res)).astype(np.float32)
    return out


def resample_datacube(data, diffraction_intensities_shape=None, reshaping_method="bilinear",
                      padded_diffraction_intensities_shape=None):
    """phase_base_class.py:223-420 on the pattern stack: 'bin'
    (py4DSTEM/preprocess/preprocess.py:155-215), 'fourier' (:636-663),
    'bilinear' (:665-711: scipy zoom per pattern, order 1, mode 'nearest',
    grid_mode, divided by the factor product), then zero padding (:719-790).
    Returns (data, factor on the reciprocal pixel size)."""
    factor = 1.0
    data = np.asarray(data)
    if diffraction_intensities_shape is not None:
        Qx, Qy = data.shape[-2:]
        Sx, Sy = diffraction_intensities_shape
        rx, ry = Sx / Qx, Sy / Qy
        if rx != ry:
            raise ValueError("Datacube calibration can only handle uniform Q-sampling.")
        if reshaping_method == "bin":
            b = int(1 / rx)
            d = data[:, :, : Qx - Qx % b, : Qy - Qy % b]
            data = d.reshape(d.shape[0], d.shape[1], Qx // b, b, Qy // b, b).sum(
                axis=(3, 5)).astype(data.dtype)
            factor = float(b)
        elif reshaping_method == "fourier":
            data = fourier_resample_4d(data, (Sx, Sy))
            factor = 1.0 / rx
        elif reshaping_method == "bilinear":
            from scipy.ndimage import zoom
            r = np.array([rx, rx])
            out = np.zeros(data.shape[:2] + tuple(np.round(r * np.array(data.shape[-2:])).astype(int)))
            for i in range(data.shape[0]):
                for j in range(data.shape[1]):
                    out[i, j] = zoom(data[i, j].astype(np.float32), r, order=1, mode="nearest",
                                     grid_mode=True)
            data = out / r.prod()
            factor = 1.0 / rx
        else:
            raise ValueError(f"reshaping_method needs to be one of 'bilinear', 'fourier', or 'bin', "
                             f"not {reshaping_method}.")
    if padded_diffraction_intensities_shape is not None:
        Qx, Qy = data.shape[-2:]
        Sx, Sy = padded_diffraction_intensities_shape
        px, py = Sx - Qx, Sy - Qy
        data = np.pad(data, ((0, 0), (0, 0), (px // 2, px // 2 + px % 2),
                             (py // 2, py // 2 + py % 2)), mode="constant")
    return data, factor


def fit_origin_2d(data, mask=None, fitfunction="plane"):
    """``fit_origin`` (py4DSTEM/process/calibration/origin.py:76-190,
    non-robust) with the models of py4DSTEM/process/fit/fit.py:127-162.
    'constant' is the mean of the whole map (no mask, as the reference); the
    other models are linear in their parameters, so ``curve_fit``'s optimum
    is the least-squares solution, computed here with ``lstsq``."""
    assert fitfunction in ("plane", "parabola", "bezier_two", "constant")
    if fitfunction == "constant":
        return np.mean(data) * np.ones_like(data)
    if fitfunction == "plane":
        return fit_plane(data, mask)
    shape = data.shape
    x, y = np.linspace(0, 1, shape[0]), np.linspace(0, 1, shape[1])
    ry, rx = np.meshgrid(y, x)
    if mask is None:
        mask = np.ones(shape, dtype=bool)
    if fitfunction == "parabola":
        cols = [np.ones_like(rx), rx, rx ** 2, ry, ry ** 2, rx * ry]
    else:
        bx = [(1 - rx) ** 2, 2 * (1 - rx) * rx, rx ** 2]
        by = [(1 - ry) ** 2, 2 * (1 - ry) * ry, ry ** 2]
        cols = [bx[i] * by[j] for i in range(3) for j in range(3)]
    A = np.stack([c[mask] for c in cols], axis=1)
    popt, *_ = np.linalg.lstsq(A, np.asarray(data[mask], dtype=np.float64), rcond=None)
    return sum(p * c for p, c in zip(popt, cols))


def fit_plane(data, mask=None):
    """Least-squares plane ``mx*x + my*y + b`` on coordinates normalised to
    [0, 1] (py4DSTEM/process/fit/fit.py:38-124, 127-128, non-robust).

    The reference minimises the same linear least-squares objective with
    ``scipy.optimize.curve_fit``; for a model linear in its parameters the
    minimiser is the normal-equation solution computed here with ``lstsq``.
    """
    shape = data.shape
    x, y = np.linspace(0, 1, shape[0]), np.linspace(0, 1, shape[1])
    ry, rx = np.meshgrid(y, x)
    if mask is None:
        mask = np.ones(shape, dtype=bool)
    A = np.stack([rx[mask], ry[mask], np.ones(int(mask.sum()))], axis=1)
    popt, *_ = np.linalg.lstsq(A, np.asarray(data[mask], dtype=np.float64), rcond=None)
    return popt[0] * rx + popt[1] * ry + popt[2]


def patch_indices(positions_px, roi_shape, obj_shape):
    """Corner-centred periodic patch indices (phase/phase_base_class.py:1970-2003)."""
    x0 = np.round(positions_px[:, 0]).astype("int")
    y0 = np.round(positions_px[:, 1]).astype("int")
    x_ind = np.fft.fftfreq(roi_shape[0], d=1 / roi_shape[0]).astype("int")
    y_ind = np.fft.fftfreq(roi_shape[1], d=1 / roi_shape[1]).astype("int")
    rows = (x0[:, None, None] + x_ind[None, :, None]) % obj_shape[0]
    cols = (y0[:, None, None] + y_ind[None, None, :]) % obj_shape[1]
    return rows, cols


def _scatter_base(patches, positions_px, roi_shape, obj_shape):
    """``np.bincount`` scatter-add with float64 accumulation
    (phase/phase_base_class.py:1905-1940)."""
    x0 = np.round(positions_px[:, 0]).astype("int")
    y0 = np.round(positions_px[:, 1]).astype("int")
    x_ind = np.fft.fftfreq(roi_shape[0], d=1 / roi_shape[0]).astype("int")
    y_ind = np.fft.fftfreq(roi_shape[1], d=1 / roi_shape[1]).astype("int")
    flat_weights = patches.ravel()
    indices = ((y0[:, None, None] + y_ind[None, None, :]) % obj_shape[1]) + (
        (x0[:, None, None] + x_ind[None, :, None]) % obj_shape[0]
    ) * obj_shape[1]
    counts = np.bincount(indices.ravel(), weights=flat_weights,
                         minlength=int(np.prod(obj_shape)))
    return np.reshape(counts, obj_shape).astype(np.float32)


def scatter_bincount(patches, positions_px, roi_shape, obj_shape):
    """Real/imag split scatter (phase/phase_base_class.py:1942-1968)."""
    if np.iscomplexobj(patches):
        real = _scatter_base(patches.real, positions_px, roi_shape, obj_shape)
        imag = _scatter_base(patches.imag, positions_px, roi_shape, obj_shape)
        return real + 1.0j * imag
    return _scatter_base(patches, positions_px, roi_shape, obj_shape)


# ---------------------------------------------------------------------------
# Minimal data model used by the oracle (duck-typed like py4DSTEM's DataCube)
# ---------------------------------------------------------------------------

# ---------------------------------------------------------------------------
# Global affine model of the scan positions (phase/utils.py:595-825)
# ---------------------------------------------------------------------------
# A transform is the tuple (scale0, scale1, shear1, angle, t0, t1); the
# identity keeps Python floats, a fitted one float32 translations, as the
# reference's AffineTransform does (the dtype of the result follows from it).
AFFINE_IDENTITY = (1.0, 1.0, 0.0, 0.0, 0.0, 0.0)


def affine_matrix(tf):
    """scale @ shear @ rotation, each a float32 2x2 (utils.py:673-702)."""
    scale0, scale1, shear1, angle = tf[:4]
    c, s = np.cos(angle), np.sin(angle)
    scale = np.array([[scale0, 0.0], [0.0, scale1]], dtype="float32")
    shear = np.array([[1.0, 0.0], [shear1, 1.0]], dtype="float32")
    rot = np.array([[+c, -s], [+s, +c]], dtype="float32")
    return scale @ shear @ rot


def affine_apply(tf, x, origin=(0, 0)):
    """utils.py:726-731: ((x - origin) @ M) + (t + origin)."""
    origin = np.asarray(origin, dtype=np.float32)
    M = affine_matrix(tf).astype(np.float32)
    return ((x - origin) @ M) + (np.array((tf[4], tf[5])) + origin)


def affine_from_matrix(T):
    """Graphics Gems II 7.1 decomposition of a 3x2 (or 2x2) least-squares
    matrix (utils.py:639-671); a degenerate row gives the identity."""
    R = T[:2, :2].copy()
    scale0 = np.linalg.norm(R[0])
    if scale0 <= 0:
        return AFFINE_IDENTITY
    R[0] /= scale0
    shear1 = R[0] @ R[1]
    R[1] -= shear1 * R[0]
    scale1 = np.linalg.norm(R[1])
    if scale1 <= 0:
        return AFFINE_IDENTITY
    R[1] /= scale1
    shear1 /= scale1
    angle = np.arccos(R[0, 0])
    t0, t1 = T[2] if T.shape[0] > 2 else (0.0, 0.0)
    return (float(scale0), float(scale1), float(shear1), float(angle), t0, t1)


def affine_fit(positions0, positions1, origin=(0, 0), translation_allowed=True):
    """Least-squares affine map positions0 -> positions1 and its residual
    norm (utils.py:743-770)."""
    origin = np.asarray(origin, dtype=np.float32)
    try:
        if translation_allowed:
            a = np.pad(positions0 - origin, ((0, 0), (0, 1)), constant_values=1)
        else:
            a = positions0 - origin
        b = positions1 - origin
        aT = a.conj().swapaxes(-1, -2)
        x = np.linalg.inv(aT @ a) @ aT @ b
        tf = affine_from_matrix(x)
    except np.linalg.LinAlgError:
        tf = AFFINE_IDENTITY
    error = np.linalg.norm(affine_apply(tf, positions0, origin) - positions1)
    return tf, error


def affine_ransac(positions0, positions1, origin=(0, 0), translation_allowed=True,
                  min_sample=64, max_error=16, min_consensus=0.75, max_iter=20):
    """RANSAC over `max_iter` random subsets drawn from the global NumPy
    generator (utils.py:773-825); returns (best transform, its residual)."""
    best_fitness = np.inf
    transform = AFFINE_IDENTITY
    for subset in np.random.choice(a=len(positions0), size=(max_iter, min_sample), replace=True):
        subset = np.unique(subset)
        candidate, _ = affine_fit(positions0[subset], positions1[subset], origin,
                                  translation_allowed)
        position_error = np.linalg.norm(affine_apply(candidate, positions0, origin) - positions1,
                                        axis=-1)
        inliers = position_error <= max_error
        if np.sum(inliers) / len(inliers) >= min_consensus:
            candidate, fitness = affine_fit(positions0[inliers], positions1[inliers], origin,
                                            translation_allowed)
            if fitness < best_fitness:
                best_fitness = fitness
                transform = candidate
    return transform, best_fitness


def bilinear_resample(array, output_size):
    """``bilinear_resample`` (phase/utils.py:2345-2428) with the arguments of
    its calls in the forward projection: vectorised order-1 ``zoom`` with
    mode 'grid-wrap' and grid_mode=True over the last two axes, then division
    by the product of the scale factors as a float64 array (which promotes a
    complex64 input to complex128 under NumPy >= 2)."""
    size = np.array(array.shape)
    input_size = size[-2:].copy()
    output_size = np.array(output_size)
    scale_output = tuple(output_size / input_size)
    scale_output = (1,) * (size.size - input_size.size) + scale_output
    array = scipy.ndimage.zoom(array, scale_output, order=1, mode="grid-wrap", grid_mode=True)
    return array / np.array(scale_output).prod()


def _rotated_com(nx, ny, theta, transposed):
    """CoM field rotated by theta, from the transposed intensities when
    ``transposed`` (phase_base_class.py:892-899, 907-914)."""
    if transposed:
        return (np.cos(theta) * ny - np.sin(theta) * nx,
                np.sin(theta) * ny + np.cos(theta) * nx)
    return (np.cos(theta) * nx - np.sin(theta) * ny,
            np.sin(theta) * nx + np.cos(theta) * ny)


def _mean_abs_curl(cx, cy):
    """phase_base_class.py:904-906: mean |d cy/dx - d cx/dy| with central
    differences on the interior (last two axes)."""
    gxy = cx[..., 1:-1, 2:] - cx[..., 1:-1, :-2]
    gyx = cy[..., 2:, 1:-1] - cy[..., :-2, 1:-1]
    return np.mean(np.abs(gyx - gxy), axis=(-2, -1))


def _mean_abs_div(cx, cy):
    """phase_base_class.py:900-903: mean |d cx/dx + d cy/dy| on the interior."""
    gxx = cx[..., 2:, 1:-1] - cx[..., :-2, 1:-1]
    gyy = cy[..., 1:-1, 2:] - cy[..., 1:-1, :-2]
    return np.mean(np.abs(gxx + gyy), axis=(-2, -1))


def solve_com_relative_rotation(com_normalized_x, com_normalized_y, rotation_angles_deg=None,
                                maximize_divergence=False, force_com_rotation=None,
                                force_com_transpose=None):
    """``_solve_for_center_of_mass_relative_rotation`` (phase_base_class.py:
    769-1345, plots and summaries omitted).  Four branches: rotation and
    transpose forced (:868-882); rotation forced, transpose decided by
    comparing the two curls/divergences at that angle (:884-938); transpose
    forced, rotation the argmin curl / argmax divergence over the angles
    (:940-1067); both searched, the transposed field winning only when
    strictly better (:1069-1190).  Angles are float32 degrees
    (np.arange(-89, 90, 1) by default).  Returns (rotation_rad, transpose,
    com_x, com_y) with the corrected CoM of :1265-1283."""
    nx, ny = com_normalized_x, com_normalized_y
    score = _mean_abs_div if maximize_divergence else _mean_abs_curl
    if rotation_angles_deg is None:
        rotation_angles_deg = np.arange(-89.0, 90.0, 1.0)
    if force_com_rotation is not None:
        rotation = np.deg2rad(force_com_rotation)
        if force_com_transpose is not None:
            transpose = force_com_transpose
        else:
            plain = score(*_rotated_com(nx, ny, rotation, False))
            flipped = score(*_rotated_com(nx, ny, rotation, True))
            transpose = flipped > plain if maximize_divergence else flipped < plain
    else:
        angles_deg = np.asarray(rotation_angles_deg, dtype=np.float32)
        angles_rad = np.deg2rad(angles_deg)[:, None, None]
        flat_angles = np.squeeze(angles_rad)
        pick = np.argmax if maximize_divergence else np.argmin
        if force_com_transpose is not None:
            transpose = force_com_transpose
            s = score(*_rotated_com(nx[None], ny[None], angles_rad, transpose))
            rotation = flat_angles[pick(s)]
        else:
            s_plain = score(*_rotated_com(nx[None], ny[None], angles_rad, False))
            s_flip = score(*_rotated_com(nx[None], ny[None], angles_rad, True))
            i, j = pick(s_plain), pick(s_flip)
            plain_wins = (s_plain[i] >= s_flip[j]) if maximize_divergence else (
                s_plain[i] <= s_flip[j])
            transpose = not plain_wins
            rotation = flat_angles[j] if transpose else flat_angles[i]
    com_x, com_y = _rotated_com(nx, ny, rotation, transpose)
    return rotation, bool(transpose), com_x, com_y


class _Cal:
    def __init__(self, R_size, R_units, Q_size, Q_units):
        self._r = (R_size, R_units)
        self._q = (Q_size, Q_units)

    def get_R_pixel_size(self):
        return self._r[0]

    def get_R_pixel_units(self):
        return self._r[1]

    def get_Q_pixel_size(self):
        return self._q[0]

    def get_Q_pixel_units(self):
        return self._q[1]


# ---------------------------------------------------------------------------
# The reconstruction class
# ---------------------------------------------------------------------------

class OracleSingleslicePtychography:
    """Restatement of ``SingleslicePtychography`` for the GD path.

    Follows phase/singleslice_ptychography.py:118-996 and the mixin and base
    class methods it calls.  Only the in-scope options are implemented:
    complex object, square ROI equal to the detector, no resampling/padding,
    no detector Fourier mask, fixed positions, and the default constraints
    (shrinkage against the FOV mask, amplitude threshold, probe CoM).
    """

    def __init__(self, energy, datacube=None, semiangle_cutoff=None,
                 semiangle_cutoff_pixels=None, rolloff=2.0,
                 vacuum_probe_intensity=None, polar_parameters=None,
                 initial_object_guess=None, initial_probe_guess=None,
                 initial_scan_positions=None, object_padding_px=None,
                 verbose=False, mode="c64", object_type="complex", **kwargs):
        # singleslice_ptychography.py:142-189
        if object_type != "potential" and object_type != "complex":
            raise ValueError(f"object_type must be either 'potential' or 'complex', not {object_type}")
        self._object_type = object_type
        for key in kwargs:
            if key not in POLAR_SYMBOLS and key not in POLAR_ALIASES:
                raise ValueError("{} not a recognized parameter".format(key))
        self._polar_parameters = dict(zip(POLAR_SYMBOLS, [0.0] * len(POLAR_SYMBOLS)))
        polar_parameters = dict(polar_parameters or {})
        polar_parameters.update(kwargs)
        set_polar_parameters(self._polar_parameters, polar_parameters)
        self._datacube = datacube
        self._object = initial_object_guess
        self._probe = initial_probe_guess
        self._vacuum_probe_intensity = vacuum_probe_intensity
        self._scan_positions = initial_scan_positions
        self._energy = energy
        self._semiangle_cutoff = semiangle_cutoff
        self._semiangle_cutoff_pixels = semiangle_cutoff_pixels
        self._rolloff = rolloff
        self._object_padding_px = object_padding_px
        self._verbose = verbose
        self._mode = mode

    # -- calibrations -------------------------------------------------------
    @property
    def sampling(self):
        """phase/phase_base_class.py:2202-2212."""
        return tuple(
            electron_wavelength_angstrom(self._energy) * 1e3 / dk / n
            for dk, n in zip(self._angular_sampling, self._amplitudes_shape)
        )

    def _extract_intensities_and_calibrations(self, datacube, data=None, q_factor=1.0):
        """phase/phase_base_class.py:424-611 (calibrated branch); `data` and
        `q_factor`: the resampled patterns and the factor on the reciprocal
        pixel size the resampling leaves in the calibration."""
        intensities = np.asarray(datacube.data if data is None else data, dtype=np.float32)
        self._grid_scan_shape = intensities.shape[:2]
        cal = datacube.calibration
        r_units = cal.get_R_pixel_units()
        if r_units == "A":
            self._scan_sampling = (cal.get_R_pixel_size(),) * 2
        elif r_units == "nm":
            self._scan_sampling = (cal.get_R_pixel_size() * 10,) * 2
        else:
            raise ValueError(f"Real-space calibrations must be given in 'A', not {r_units}")
        q_units = cal.get_Q_pixel_units()
        lam = electron_wavelength_angstrom(self._energy)
        if q_units == "A^-1":
            rs = cal.get_Q_pixel_size() * q_factor
            self._reciprocal_sampling = (rs,) * 2
            self._angular_sampling = (rs * lam * 1e3,) * 2
        elif q_units == "mrad":
            a = cal.get_Q_pixel_size() * q_factor
            self._angular_sampling = (a,) * 2
            self._reciprocal_sampling = (a / lam / 1e3,) * 2
        else:
            raise ValueError("Reciprocal-space calibrations must be given in 'A^-1' or 'mrad'")
        return intensities

    # -- preprocess ---------------------------------------------------------
    def preprocess(self, force_com_rotation=None, force_com_transpose=None,
                   force_com_shifts=None, fit_function="plane",
                   object_fov_mask=None, center_positions_in_fov=True,
                   max_batch_size=None, diffraction_intensities_shape=None,
                   reshaping_method="bilinear", padded_diffraction_intensities_shape=None,
                   **kwargs):
        """phase/singleslice_ptychography.py:191-604 (no plotting)."""
        mode = self._mode
        dc = self._datacube
        Rshape = dc.data.shape[:2]
        if force_com_shifts is not None and np.isscalar(force_com_shifts[0]):
            # phase_base_class.py:277-282
            force_com_shifts = (np.ones(Rshape) * force_com_shifts[0],
                                np.ones(Rshape) * force_com_shifts[1])
        data, q_factor = None, 1.0
        if diffraction_intensities_shape is not None or padded_diffraction_intensities_shape is not None:
            data, q_factor = resample_datacube(dc.data, diffraction_intensities_shape, reshaping_method,
                                               padded_diffraction_intensities_shape)
            if diffraction_intensities_shape is not None and force_com_shifts is not None:
                rf = diffraction_intensities_shape[0] / dc.data.shape[-2]  # :290-294
                force_com_shifts = (force_com_shifts[0] * rf, force_com_shifts[1] * rf)
        intensities = self._extract_intensities_and_calibrations(dc, data, q_factor)
        if self._semiangle_cutoff_pixels:
            self._semiangle_cutoff = self._semiangle_cutoff_pixels * self._angular_sampling[0]

        # CoM (phase_base_class.py:613-767, vectorised branch)
        kx = np.arange(intensities.shape[-2], dtype=np.float32)
        ky = np.arange(intensities.shape[-1], dtype=np.float32)
        kya, kxa = np.meshgrid(ky, kx)
        isum = np.sum(intensities, axis=(-2, -1))
        com_mx = np.sum(intensities * kxa[None, None], axis=(-2, -1)) / isum
        com_my = np.sum(intensities * kya[None, None], axis=(-2, -1)) / isum
        if force_com_shifts is None:
            if fit_function is not None:
                finite = np.isfinite(com_mx)
                com_fx = np.asarray(fit_origin_2d(com_mx, finite, fit_function), dtype=np.float32)
                com_fy = np.asarray(fit_origin_2d(com_my, finite, fit_function), dtype=np.float32)
            else:
                com_fx = np.asarray(com_mx, dtype=np.float32)
                com_fy = np.asarray(com_my, dtype=np.float32)
        else:
            com_fx = np.asarray(force_com_shifts[0], dtype=np.float32)
            com_fy = np.asarray(force_com_shifts[1], dtype=np.float32)
        self._com_fitted_x, self._com_fitted_y = com_fx, com_fy

        # normalised CoM (phase_base_class.py:749-757) and the rotation /
        # transpose search (phase_base_class.py:769-1345)
        self._com_normalized_x = np.nan_to_num(com_mx - com_fx) * self._reciprocal_sampling[0]
        self._com_normalized_y = np.nan_to_num(com_my - com_fy) * self._reciprocal_sampling[1]
        (self._rotation_best_rad, self._rotation_best_transpose, self._com_x,
         self._com_y) = solve_com_relative_rotation(
            self._com_normalized_x, self._com_normalized_y,
            kwargs.get("rotation_angles_deg"), kwargs.get("maximize_divergence", False),
            force_com_rotation, force_com_transpose)

        # normalisation (phase_base_class.py:1347-1460)
        mean_intensity = 0
        diff = intensities.copy()
        for rx in range(diff.shape[0]):
            for ry in range(diff.shape[1]):
                I = get_shifted_ar(diff[rx, ry], -com_fx[rx, ry], -com_fy[rx, ry])
                mean_intensity += np.sum(I)
                diff[rx, ry] = np.sqrt(np.maximum(I, 0))
        qx, qy = diff.shape[-2:]
        diff = diff.reshape((-1, qx, qy))
        mean_intensity /= diff.shape[0]
        self._amplitudes = diff
        self._mean_diffraction_intensity = mean_intensity
        self._num_diffraction_patterns = diff.shape[0]
        self._amplitudes_shape = np.array(diff.shape[-2:])
        # singleslice_ptychography.py:431-436: a given region of interest turns
        # on the far-field resampling in the forward projection; only the
        # patterns' own shape is restated here
        roi = kwargs.get("region_of_interest_shape")
        self._resample_exit_waves = roi is not None
        if roi is not None and tuple(int(v) for v in np.ravel(roi)) != tuple(diff.shape[-2:]):
            raise NotImplementedError("oracle: region_of_interest_shape other than the patterns'")
        self._region_of_interest_shape = np.array(diff.shape[-2:])

        # positions (phase_base_class.py:1811-1903)
        self._positions_px, self._object_padding_px = self._positions_in_pixels(
            self._scan_positions, self._object_padding_px)

        # object (ptychographic_methods.py:40-79)
        if self._object is None:
            pad_x = self._object_padding_px[0][1]
            pad_y = self._object_padding_px[1][1]
            p, q = np.round(np.max(self._positions_px, axis=0))
            p = np.max([np.round(p + pad_x), self._region_of_interest_shape[0]]).astype("int")
            q = np.max([np.round(q + pad_y), self._region_of_interest_shape[1]]).astype("int")
            self._object = (np.zeros((p, q), dtype=np.float32) if self._object_type == "potential"
                            else np.ones((p, q), dtype=np.complex64))
        else:
            self._object = np.asarray(self._object, dtype=np.float32
                                      if self._object_type == "potential" else np.complex64)
        self._object_initial = self._object.copy()
        self._object_shape = self._object.shape

        # singleslice_ptychography.py:462-478
        self._positions_px = np.asarray(self._positions_px, dtype=np.float32)
        com = self._positions_px.mean(0)
        if center_positions_in_fov:
            self._positions_px -= com - np.array(self._object_shape) / 2
        self._positions_px_initial = self._positions_px.copy()

        # probe (ptychographic_methods.py:1011-1102)
        if self._probe is None:
            vacuum = None
            if self._vacuum_probe_intensity is not None:  # :1033-1064 (no crop, no resampling)
                vacuum = np.asarray(self._vacuum_probe_intensity, dtype=np.float32)
                x0, y0 = get_CoM(vacuum)
                vacuum = get_shifted_ar(vacuum, -x0, -y0)
                if vacuum.shape != tuple(self._region_of_interest_shape):
                    raise NotImplementedError("vacuum probe resampling")
                self._semiangle_cutoff = np.inf
            probe = ComplexProbeOracle(
                energy=self._energy, gpts=self._region_of_interest_shape,
                sampling=self.sampling, semiangle_cutoff=self._semiangle_cutoff,
                rolloff=self._rolloff, vacuum_probe_intensity=vacuum,
                parameters=self._polar_parameters, mode=mode,
            ).build()._array
            probe_intensity = np.sum(np.abs(_fft2(probe, mode)) ** 2)
            probe *= np.sqrt(self._mean_diffraction_intensity / probe_intensity)
        else:
            probe = np.asarray(self._probe, dtype=np.complex64)
        self._probe = probe
        self._probe_initial = probe.copy()
        self._probe_initial_aperture = np.abs(_fft2(probe, mode))

        # FOV mask (singleslice_ptychography.py:504-537)
        if object_fov_mask is None:
            N = self._num_diffraction_patterns
            if max_batch_size is None:
                max_batch_size = N
            overlap = np.zeros(self._object_shape, dtype=np.float32)
            for start, end in generate_batches(N, max_batch=max_batch_size):
                pos = self._positions_px[start:end]
                frac = pos - np.round(pos)
                sp = fft_shift(self._probe, frac, mode)
                overlap += scatter_bincount(np.abs(sp) ** 2, pos,
                                            self._region_of_interest_shape,
                                            self._object_shape)
            self._probe_overlap = overlap
            blurred = scipy.ndimage.gaussian_filter(overlap, 1.0)
            self._object_fov_mask = blurred > 0.25 * blurred.max()
        elif object_fov_mask is True:
            self._object_fov_mask = np.full(self._object_shape, True)
        else:
            self._object_fov_mask = np.asarray(object_fov_mask)
        self._object_fov_mask_inverse = np.invert(self._object_fov_mask)
        return self

    def _positions_in_pixels(self, positions, object_padding_px):
        """phase/phase_base_class.py:1811-1903."""
        sampling = self.sampling
        if positions is None:
            nx, ny = self._grid_scan_shape
            sx, sy = self._scan_sampling
            x = np.arange(nx) * sx
            y = np.arange(ny) * sy
            x, y = np.meshgrid(x, y, indexing="ij")
            positions = np.stack((x.ravel(), y.ravel()), axis=-1)
        else:
            positions = np.array(positions, dtype=np.float64)
        if self._rotation_best_rad is not None:
            positions = affine_rotate(positions, self._rotation_best_rad, positions.mean(0))
        if self._rotation_best_transpose:
            positions = np.flip(positions, 1)
            sampling = sampling[::-1]
        positions -= np.min(positions, axis=0).clip(-np.inf, 0)
        positions[:, 0] /= sampling[0]
        positions[:, 1] /= sampling[1]
        if object_padding_px is None:
            float_padding = self._region_of_interest_shape / 2
            object_padding_px = (float_padding, float_padding)
        elif np.isscalar(object_padding_px[0]):
            object_padding_px = ((object_padding_px[0],) * 2, (object_padding_px[1],) * 2)
        positions[:, 0] += object_padding_px[0][0]
        positions[:, 1] += object_padding_px[1][0]
        return positions, object_padding_px

    # -- forward / adjoint --------------------------------------------------
    def forward(self, obj, probe, positions_px, amplitudes, fourier_mask=None):
        """``_forward`` for GD (ptychographic_methods.py:1482-1527, 1576-1635,
        1736-1816), with the optional detector mask of :1603-1616.  Returns
        (shifted_probes, object_patches, exit_waves, error)."""
        mode = self._mode
        frac = positions_px - np.round(positions_px)
        rows, cols = patch_indices(positions_px, self._region_of_interest_shape, obj.shape)
        shifted_probes = fft_shift(probe, frac, mode)
        object_patches = obj[rows, cols]
        if getattr(self, "_object_type", "complex") == "potential":  # :1522-1523
            object_patches = np.exp(1j * object_patches)
        overlap = shifted_probes * object_patches
        fourier_overlap = _fft2(overlap, mode)
        resample = getattr(self, "_resample_exit_waves", False)
        if resample:  # ptychographic_methods.py:1603-1611
            fourier_overlap = bilinear_resample(fourier_overlap, self._amplitudes_shape)
        if fourier_mask is not None:
            fourier_overlap *= fourier_mask
        farfield = np.abs(fourier_overlap)
        error = np.sum(np.abs(amplitudes - farfield) ** 2)
        modified = amplitudes * np.exp(1j * np.angle(fourier_overlap))
        modified = modified - fourier_overlap
        if fourier_mask is not None:
            modified *= fourier_mask
        if resample:  # ptychographic_methods.py:1620-1629
            modified = bilinear_resample(modified, self._region_of_interest_shape)
        exit_waves = _ifft2(modified, mode)
        return shifted_probes, object_patches, exit_waves, error

    def adjoint(self, obj, probe, object_patches, shifted_probes, positions_px,
                exit_waves, step_size, normalization_min, fix_probe):
        """``_gradient_descent_adjoint`` (ptychographic_methods.py:1818-1908),
        complex or potential object.  Updates in place, as the reference's
        ``+=`` does."""
        roi = self._region_of_interest_shape
        shape = obj.shape
        probe_normalization = scatter_bincount(np.abs(shifted_probes) ** 2,
                                               positions_px, roi, shape)
        probe_normalization = 1 / np.sqrt(
            1e-16
            + ((1 - normalization_min) * probe_normalization) ** 2
            + (normalization_min * np.max(probe_normalization)) ** 2
        )
        if getattr(self, "_object_type", "complex") == "potential":  # :1868-1880
            obj += step_size * (
                scatter_bincount(np.real(-1j * np.conj(object_patches) * np.conj(shifted_probes)
                                         * exit_waves), positions_px, roi, shape)
                * probe_normalization
            )
        else:
            obj += step_size * (
                scatter_bincount(np.conj(shifted_probes) * exit_waves, positions_px, roi, shape)
                * probe_normalization
            )
        if not fix_probe:
            object_normalization = np.sum(np.abs(object_patches) ** 2, axis=0)
            object_normalization = 1 / np.sqrt(
                1e-16
                + ((1 - normalization_min) * object_normalization) ** 2
                + (normalization_min * np.max(object_normalization)) ** 2
            )
            probe += step_size * (
                np.sum(np.conj(object_patches) * exit_waves, axis=0) * object_normalization
            )
        return obj, probe

    def loss_gradients(self, obj, probe, positions_px, amplitudes):
        """Gradients of L = sum_b sum (|fft2(P_shift_b * O_b)| - A_b)^2 as
        dL/dRe + i dL/dIm.

        Object: -2 S^2 scatter(conj(P_shift) chi), i.e. the un-normalised,
        un-stepped object term of _gradient_descent_adjoint
        (ptychographic_methods.py:1878-1884).  Probe: the chain rule through
        the Fourier sub-pixel shift P_shift = ifft2(fft2(P) T_b)
        (phase/utils.py:508-533), -2 S^2 sum_b ifft2(fft2(conj(O_b) chi_b)
        conj(T_b)); it equals the reference's probe term
        (ptychographic_methods.py:1900-1906) only when every fraction is 0.
        Returns (loss, grad_object, grad_probe).
        """
        mode = self._mode
        roi = self._region_of_interest_shape
        S2 = int(roi[0]) * int(roi[1])
        sp, op, chi, err = self.forward(obj, probe, positions_px, amplitudes)
        g_obj = -2.0 * S2 * scatter_bincount(np.conj(sp) * chi, positions_px, roi, obj.shape)
        frac = positions_px - np.round(positions_px)
        # per-position adjoint shift ifft2(fft2(g_b) conj(T_b)); fft_shift would
        # broadcast a 3-D array against 3-D positions as an outer product
        T = fourier_translation_operator(frac, tuple(roi))
        back = _ifft2(_fft2(np.conj(op) * chi, mode) * np.conj(T), mode)
        g_probe = -2.0 * S2 * np.sum(back, axis=0)
        return err, g_obj, g_probe

    # -- projection sets ----------------------------------------------------
    @staticmethod
    def method_parameters(reconstruction_method, reconstruction_parameter=1.0,
                          reconstruction_parameter_a=None, reconstruction_parameter_b=None,
                          reconstruction_parameter_c=None):
        """phase_base_class.py:2005-2108: (use_projection_scheme, a, b, c)."""
        m = reconstruction_method
        r = reconstruction_parameter
        if m == "generalized-projections":
            if (reconstruction_parameter_a is None or reconstruction_parameter_b is None
                    or reconstruction_parameter_c is None):
                raise ValueError("reconstruction_parameter_a/b/c must all be specified "
                                 "when using reconstruction_method='generalized-projections'.")
            return (True, reconstruction_parameter_a, reconstruction_parameter_b,
                    reconstruction_parameter_c)
        if m in ("DM_AP", "difference-map_alternating-projections"):
            if r < 0.0 or r > 1.0:
                raise ValueError("reconstruction_parameter must be between 0-1.")
            return True, -r, 1, 1 + r
        if m in ("RAAR", "relaxed-averaged-alternating-reflections"):
            if r < 0.0 or r > 1.0:
                raise ValueError("reconstruction_parameter must be between 0-1.")
            return True, 1 - 2 * r, r, 2
        if m in ("RRR", "relax-reflect-reflect"):
            if r < 0.0 or r > 2.0:
                raise ValueError("reconstruction_parameter must be between 0-2.")
            return True, -r, r, 2
        if m in ("SUPERFLIP", "charge-flipping"):
            return True, 0, 1, 2
        if m in ("GD", "gradient-descent"):
            return False, None, None, None
        raise ValueError(f"reconstruction_method {m!r} not recognized")

    def projection_forward(self, obj, probe, positions_px, amplitudes, exit_waves, a, b, c):
        """``_forward`` with use_projection_scheme (ptychographic_methods.py:
        1736-1816 -> 1637-1734).  Returns (shifted_probes, object_patches,
        exit_waves, error)."""
        mode = self._mode
        frac = positions_px - np.round(positions_px)
        rows, cols = patch_indices(positions_px, self._region_of_interest_shape, obj.shape)
        shifted_probes = fft_shift(probe, frac, mode)
        object_patches = obj[rows, cols]
        overlap = shifted_probes * object_patches
        projection_x = 1 - a - b
        projection_y = 1 - c
        if exit_waves is None:
            exit_waves = overlap.copy()
        factor_to_be_projected = c * overlap + projection_y * exit_waves
        fourier_projected_factor = _fft2(factor_to_be_projected, mode)
        farfield_amplitudes = np.abs(fourier_projected_factor)
        error = np.sum(np.abs(amplitudes - farfield_amplitudes) ** 2)
        fourier_projected_factor = amplitudes * np.exp(1j * np.angle(fourier_projected_factor))
        projected_factor = _ifft2(fourier_projected_factor, mode)
        exit_waves = projection_x * exit_waves + a * overlap + b * projected_factor
        return shifted_probes, object_patches, exit_waves, error

    def projection_adjoint(self, obj, probe, object_patches, shifted_probes, positions_px,
                           exit_waves, normalization_min, fix_probe):
        """``_projection_sets_adjoint`` for a complex object
        (ptychographic_methods.py:1910-1998): object and probe replaced by
        their normalised back-projections."""
        roi = self._region_of_interest_shape
        shape = obj.shape
        probe_normalization = scatter_bincount(np.abs(shifted_probes) ** 2, positions_px, roi,
                                               shape)
        probe_normalization = 1 / np.sqrt(
            1e-16
            + ((1 - normalization_min) * probe_normalization) ** 2
            + (normalization_min * np.max(probe_normalization)) ** 2
        )
        obj = (scatter_bincount(np.conj(shifted_probes) * exit_waves, positions_px, roi, shape)
               * probe_normalization)
        if not fix_probe:
            object_normalization = np.sum(np.abs(object_patches) ** 2, axis=0)
            object_normalization = 1 / np.sqrt(
                1e-16
                + ((1 - normalization_min) * object_normalization) ** 2
                + (normalization_min * np.max(object_normalization)) ** 2
            )
            probe = np.sum(np.conj(object_patches) * exit_waves, axis=0) * object_normalization
        return obj, probe

    # -- position correction ------------------------------------------------
    def position_partials(self, obj, positions_px, shifted_probes, overlap, amplitudes):
        """Intensity residual and the two finite-difference intensity partials
        of ``_position_correction`` (ptychographic_methods.py:2117-2186):
        F = fft2(overlap), d = A^2 - |F|^2 and, for the object patch moved one
        pixel along each axis, 2 Re((F - F_shifted) conj(F)).  With probe
        modes ([B, K, S, S] shifted probes and overlaps) the estimated
        intensity is sqrt(sum_k |F_k|^2)^2 (:2736-2740) and the partials are
        summed over the modes (:2187-2189).  Returns (d, partial_x,
        partial_y), each [B, S*S]."""
        mode = self._mode
        mixed = shifted_probes.ndim == 4
        rows, cols = patch_indices(positions_px, self._region_of_interest_shape, obj.shape)
        overlap_fft = _fft2(overlap, mode)
        overlap_fft_conj = np.conj(overlap_fft)
        if mixed:
            estimated_intensity = np.sqrt(np.sum(np.abs(overlap_fft) ** 2, axis=1)) ** 2
        else:
            estimated_intensity = np.abs(overlap_fft) ** 2
        measured_intensity = amplitudes ** 2
        flat_shape = (measured_intensity.shape[0], -1)
        difference_intensity = (measured_intensity - estimated_intensity).reshape(flat_shape)
        patch_dx = obj[(rows + 1) % obj.shape[0], cols]
        patch_dy = obj[rows, (cols + 1) % obj.shape[1]]
        if mixed:
            patch_dx, patch_dy = patch_dx[:, None], patch_dy[:, None]
        overlap_dx = shifted_probes * patch_dx
        overlap_dy = shifted_probes * patch_dy
        overlap_dx_fft = overlap_fft - _fft2(overlap_dx, mode)
        overlap_dy_fft = overlap_fft - _fft2(overlap_dy, mode)
        partial_dx = 2 * np.real(overlap_dx_fft * overlap_fft_conj)
        partial_dy = 2 * np.real(overlap_dy_fft * overlap_fft_conj)
        if mixed:
            partial_dx = partial_dx.sum(1)
            partial_dy = partial_dy.sum(1)
        return (difference_intensity, partial_dx.reshape(flat_shape),
                partial_dy.reshape(flat_shape))

    def position_correction(self, obj, positions_px, shifted_probes, overlap, amplitudes,
                            positions_px_initial, positions_step_size,
                            max_position_update_distance=None, max_position_total_distance=None):
        """``_position_correction`` (ptychographic_methods.py:2071-2225) for a
        single-mode probe without resampling: per position the least-squares
        step u = (A^T A)^-1 A^T d with A = [partial_x, partial_y], scaled by
        the step size, optionally clipped, subtracted from the positions.
        `positions_px` is the batch's copy and is updated in place."""
        diff, pdx, pdy = self.position_partials(obj, positions_px, shifted_probes, overlap,
                                                amplitudes)
        coefficients_matrix = np.dstack((pdx, pdy))
        coefficients_matrix_T = coefficients_matrix.conj().swapaxes(-1, -2)
        positions_update = (np.linalg.inv(coefficients_matrix_T @ coefficients_matrix)
                            @ coefficients_matrix_T @ diff[..., None])
        positions_update = positions_update[..., 0] * positions_step_size
        sampling = self.sampling
        if max_position_update_distance is not None:
            max_position_update_distance /= np.sqrt(sampling[0] ** 2 + sampling[1] ** 2)
            update_norms = np.linalg.norm(positions_update, axis=1)
            outlier_ind = update_norms > max_position_update_distance
            positions_update[outlier_ind] /= (update_norms[outlier_ind, None]
                                              / max_position_update_distance)
        if max_position_total_distance is not None:
            max_position_total_distance /= np.sqrt(sampling[0] ** 2 + sampling[1] ** 2)
            deltas = (positions_px - positions_px_initial) - positions_update
            dsts = np.linalg.norm(deltas, axis=1)
            outlier_ind = dsts > max_position_total_distance
            positions_update[outlier_ind] = 0
        positions_px -= positions_update
        return positions_px

    # -- constraints --------------------------------------------------------
    @staticmethod
    def object_butterworth_constraint(obj, q_lowpass, q_highpass, butterworth_order, sampling):
        """ptychographic_constraints.py:153-201 (complex object): NumPy dtypes
        as the reference's, so the float64 envelope promotes the object to
        complex128 from here on."""
        qx = np.fft.fftfreq(obj.shape[-2], sampling[0])
        qy = np.fft.fftfreq(obj.shape[-1], sampling[1])
        qya, qxa = np.meshgrid(qy, qx)
        qra = np.sqrt(qxa ** 2 + qya ** 2)
        env = np.ones_like(qra)
        if q_highpass:
            env *= 1 - 1 / (1 + (qra / q_highpass) ** (2 * butterworth_order))
        if q_lowpass:
            env *= 1 / (1 + (qra / q_lowpass) ** (2 * butterworth_order))
        mean = np.mean(obj, axis=(-2, -1), keepdims=True)
        obj = obj - mean
        obj = np.fft.ifft2(np.fft.fft2(obj) * env)
        return obj + mean

    @staticmethod
    def object_constraints(obj, object_mask, shrinkage_rad=0.0, pure_phase_object=False,
                           gaussian_sigma_px=None, butterworth=None, object_type="complex",
                           object_positivity=True):
        """ptychographic_constraints.py:29-118, 411-464; the gaussian
        smoothing of :119-151 when gaussian_sigma_px is given (the reference's
        scipy.ndimage.gaussian_filter call with its defaults).  A potential
        object is smoothed directly, shrinks by subtraction and is clipped at
        zero (object_positivity) instead of the amplitude threshold."""
        potential = object_type == "potential"
        if gaussian_sigma_px is not None:
            import scipy.ndimage
            if not pure_phase_object or potential:
                obj = scipy.ndimage.gaussian_filter(obj, gaussian_sigma_px)
            else:
                phase = scipy.ndimage.gaussian_filter(np.angle(obj), gaussian_sigma_px)
                obj = np.exp(1.0j * phase)
        if butterworth is not None:  # (q_lowpass, q_highpass, order, sampling)
            obj = OracleSingleslicePtychography.object_butterworth_constraint(obj, *butterworth)
            if potential:
                obj = np.real(obj)
        if potential:
            if shrinkage_rad > 0.0 or object_mask is not None:
                if object_mask is not None:
                    shrinkage_rad += obj[..., object_mask].mean()
                obj -= shrinkage_rad
            return obj.clip(0.0) if object_positivity else obj
        if shrinkage_rad > 0.0 or object_mask is not None:
            phase = np.angle(obj)
            amp = np.abs(obj)
            if object_mask is not None:
                shrinkage_rad += phase[..., object_mask].mean()
            phase -= shrinkage_rad
            obj = amp * np.exp(1.0j * phase)
        phase = np.angle(obj)
        amplitude = 1.0 if pure_phase_object else np.minimum(np.abs(obj), 1.0)
        return amplitude * np.exp(1.0j * phase)

    def probe_com_constraint(self, probe):
        """ptychographic_constraints.py:894-918."""
        intensity = np.abs(probe) ** 2
        x0, y0 = get_CoM(intensity, corner_centered=True)
        return fft_shift(probe, -np.array([x0, y0]), self._mode)

    def probe_extra_constraints(self, probe):
        """Probe constraints of subclasses (none for a single probe)."""
        return probe

    def probe_aperture_constraint(self, probe, initial_aperture):
        """ptychographic_constraints.py:1004-1033: Fourier amplitude replaced
        by the initial aperture, intensity kept."""
        current_probe_sum = np.sum(np.abs(probe) ** 2)
        phase = np.angle(_fft2(probe, self._mode))
        updated = _ifft2(np.exp(1j * phase) * initial_aperture, self._mode)
        updated_sum = np.sum(np.abs(updated) ** 2)
        return updated * np.sqrt(current_probe_sum / updated_sum)

    @staticmethod
    def probe_amplitude_constraint(probe, relative_radius, relative_width):
        """ptychographic_constraints.py:920-957: real-space top-hat with an
        erf edge, intensity kept."""
        from scipy.special import erf
        current_probe_sum = np.sum(np.abs(probe) ** 2)
        X = np.fft.fftfreq(probe.shape[0])[:, None]
        Y = np.fft.fftfreq(probe.shape[1])[None]
        r = np.hypot(X, Y) - relative_radius
        sigma = np.sqrt(np.pi) / relative_width
        tophat_mask = 0.5 * (1 - erf(sigma * r / (1 - r ** 2)))
        updated = probe * tophat_mask
        updated_sum = np.sum(np.abs(updated) ** 2)
        return updated * np.sqrt(current_probe_sum / updated_sum)

    @staticmethod
    def positions_constraints(current_positions, initial_positions, fix_positions,
                              fix_positions_com, global_affine_transformation):
        """ptychographic_constraints.py:1292-1377.  `fix_positions_com` is the
        flag as the reconstruct loop passes it (its own flag and not
        fix_positions, singleslice_ptychography.py:956): the centre-of-mass
        step runs when that flag is False."""
        if not fix_positions:
            if not fix_positions_com:
                current_positions -= current_positions.mean(0) - initial_positions.mean(0)
            if global_affine_transformation:
                origin = initial_positions.mean(0)
                tf, _ = affine_ransac(initial_positions, current_positions, origin,
                                      min_sample=initial_positions.shape[0] // 10)
                current_positions = affine_apply(tf, initial_positions, origin)
        return current_positions

    # -- reconstruct --------------------------------------------------------
    def reconstruct(self, num_iter=8, max_batch_size=None, seed_random=None,
                    step_size=0.5, normalization_min=1, fix_probe=False,
                    fix_probe_com=True, fix_potential_baseline=True,
                    shrinkage_rad=0.0, pure_phase_object=False,
                    store_iterations=False, reset=None, positions_step_size=0.5,
                    fix_positions=True, fix_positions_com=True,
                    max_position_update_distance=None, max_position_total_distance=None,
                    global_affine_transformation=False, reconstruction_method="gradient-descent",
                    reconstruction_parameter=1.0, reconstruction_parameter_a=None,
                    reconstruction_parameter_b=None, reconstruction_parameter_c=None,
                    detector_fourier_mask=None, fix_probe_aperture=False,
                    constrain_probe_amplitude=False,
                    constrain_probe_amplitude_relative_radius=0.5,
                    constrain_probe_amplitude_relative_width=0.05, gaussian_filter_sigma=None,
                    gaussian_filter=True, butterworth_filter=True, q_lowpass=None, q_highpass=None,
                    butterworth_order=2, object_positivity=True, **kwargs):
        """phase/singleslice_ptychography.py:606-996 (gradient descent and the
        projection sets)."""
        potential = getattr(self, "_object_type", "complex") == "potential"
        if potential and (reconstruction_method not in ("GD", "gradient-descent")
                          or not fix_positions):
            raise NotImplementedError("potential objects: gradient descent with fixed positions")
        pure_phase_object = pure_phase_object and not potential  # :976
        # ptychographic_constraints.py:142, flag as singleslice_ptychography.py:959
        gaussian_sigma_px = (float(gaussian_filter_sigma) / float(self.sampling[0])
                             if gaussian_filter and gaussian_filter_sigma is not None else None)
        butterworth = ((q_lowpass, q_highpass, butterworth_order, self.sampling)
                       if butterworth_filter and (q_lowpass is not None or q_highpass is not None)
                       else None)
        use_projection, pa, pb, pc = self.method_parameters(
            reconstruction_method, reconstruction_parameter, reconstruction_parameter_a,
            reconstruction_parameter_b, reconstruction_parameter_c)
        if use_projection and max_batch_size is not None:
            raise ValueError("Stochastic object/probe updating is inconsistent with 'DM_AP', "
                             "'RAAR', 'RRR', and 'SUPERFLIP'. Use reconstruction_method='GD' or "
                             "set max_batch_size=None.")
        if use_projection and detector_fourier_mask is not None:
            raise NotImplementedError()  # ptychographic_methods.py:1686-1687
        # ptychographic_methods.py:252-294
        if store_iterations and (not hasattr(self, "object_iterations") or reset):
            self.object_iterations = []
            self.probe_iterations = []
        if reset is True:
            self.error_iterations = []
            self._object = self._object_initial.copy()
            self._probe = self._probe_initial.copy()
            self._positions_px = self._positions_px_initial.copy()
            self._exit_waves = None
        elif reset is None:
            if hasattr(self, "error"):
                warnings.warn("Continuing reconstruction from previous result. "
                              "Use reset=True for a fresh start.", UserWarning)
            else:
                self.error_iterations = []
                self._exit_waves = None
        else:
            if not hasattr(self, "error"):
                self.error_iterations = []
                self._exit_waves = None
        if not hasattr(self, "_exit_waves"):
            self._exit_waves = None

        N = self._num_diffraction_patterns
        shuffled_indices = np.arange(N)
        if max_batch_size is not None:
            np.random.seed(seed_random)
        else:
            max_batch_size = N
        object_mask = (self._object_fov_mask_inverse
                       if fix_potential_baseline and self._object_fov_mask_inverse.sum() > 0
                       else None)
        error = 0.0
        for _ in range(num_iter):
            error = 0.0
            if not use_projection:
                np.random.shuffle(shuffled_indices)
            for start, end in generate_batches(N, max_batch=max_batch_size):
                idx = shuffled_indices[start:end]
                pos = self._positions_px[idx]
                if use_projection:
                    sp, op, self._exit_waves, err = self.projection_forward(
                        self._object, self._probe, pos, self._amplitudes[idx], self._exit_waves,
                        pa, pb, pc)
                    self._object, self._probe = self.projection_adjoint(
                        self._object, self._probe, op, sp, pos, self._exit_waves,
                        normalization_min, fix_probe)
                    ew = None
                else:
                    sp, op, ew, err = self.forward(self._object, self._probe, pos,
                                                   self._amplitudes[idx], detector_fourier_mask)
                    self._object, self._probe = self.adjoint(
                        self._object, self._probe, op, sp, pos, ew, step_size,
                        normalization_min, fix_probe)
                if not fix_positions:  # singleslice_ptychography.py:913-926
                    overlap = sp * (op[:, None] if sp.ndim == 4 else op)
                    self._positions_px[idx] = self.position_correction(
                        self._object, pos, sp, overlap, self._amplitudes[idx],
                        self._positions_px_initial[idx], positions_step_size,
                        max_position_update_distance, max_position_total_distance)
                error += err
            error /= self._mean_diffraction_intensity * N
            self._object = self.object_constraints(self._object, object_mask,
                                                   shrinkage_rad, pure_phase_object,
                                                   gaussian_sigma_px, butterworth,
                                                   getattr(self, "_object_type", "complex"),
                                                   object_positivity)
            if fix_probe_com and not fix_probe:
                self._probe = self.probe_com_constraint(self._probe)
            # ptychographic_constraints.py:1091-1147 (single probe) and
            # 1218-1286 (modes: the aperture on mode 0, the amplitude on every mode)
            if fix_probe_aperture and not fix_probe:
                if self._probe.ndim == 3:
                    self._probe[0] = self.probe_aperture_constraint(self._probe[0],
                                                                    self._probe_initial_aperture[0])
                else:
                    self._probe = self.probe_aperture_constraint(self._probe,
                                                                 self._probe_initial_aperture)
            if constrain_probe_amplitude and not fix_probe:
                if self._probe.ndim == 3:
                    for k in range(self._probe.shape[0]):
                        self._probe[k] = self.probe_amplitude_constraint(
                            self._probe[k], constrain_probe_amplitude_relative_radius,
                            constrain_probe_amplitude_relative_width)
                else:
                    self._probe = self.probe_amplitude_constraint(
                        self._probe, constrain_probe_amplitude_relative_radius,
                        constrain_probe_amplitude_relative_width)
            self._probe = self.probe_extra_constraints(self._probe)
            self._positions_px = self.positions_constraints(
                self._positions_px, self._positions_px_initial, fix_positions,
                fix_positions_com and not fix_positions, global_affine_transformation)
            self.error_iterations.append(float(error))
            if store_iterations:
                self.object_iterations.append(np.asarray(self._object).copy())
                self.probe_iterations.append(np.fft.fftshift(self._probe, axes=(-2, -1)))
        self.object = np.asarray(self._object)
        self.probe = np.fft.fftshift(self._probe, axes=(-2, -1))
        self.error = float(error)
        if not use_projection:  # singleslice_ptychography.py:990-992
            self._exit_waves = None
        return self


# ---------------------------------------------------------------------------
# Mixed-state ptychography (several incoherent probe modes)
# ---------------------------------------------------------------------------
def probe_orthogonalization(probe):
    """ptychographic_constraints.py:1180-1221: modes rotated onto the
    eigenvectors of their Gram matrix (upper triangle, complex64), then sorted
    by real-space intensity, largest first."""
    n_probes = probe.shape[0]
    pairwise_dot_product = np.empty((n_probes, n_probes), dtype=probe.dtype)
    for i in range(n_probes):
        for j in range(i, n_probes):
            pairwise_dot_product[i, j] = np.sum(probe[i].conj() * probe[j])
    _, evecs = np.linalg.eigh(pairwise_dot_product, UPLO="U")
    probe = np.tensordot(evecs.T, probe, axes=1)
    intensities = np.sum(np.abs(probe) ** 2, axis=(-2, -1))
    intensities_order = np.argsort(intensities, axis=None)[::-1]
    return probe[intensities_order]


def mixed_probe_modes(probe, num_probes):
    """ptychographic_methods.py:1416-1462: mode 0 is the single-mode probe,
    mode i the previous one times random sub-pixel phase ramps drawn from the
    global NumPy generator."""
    sx, sy = probe.shape
    probes = np.zeros((num_probes, sx, sy), dtype=np.complex64)
    probes[0] = probe
    for i_probe in range(1, num_probes):
        shift_x = np.exp(-2j * np.pi * (np.random.rand() - 0.5) * np.fft.fftfreq(sx))
        shift_y = np.exp(-2j * np.pi * (np.random.rand() - 0.5) * np.fft.fftfreq(sy))
        probes[i_probe] = probes[i_probe - 1] * shift_x[:, None] * shift_y[None]
    return probes


class OracleMixedstatePtychography(OracleSingleslicePtychography):
    """``MixedstatePtychography`` (mixedstate_ptychography.py:45-1027) for the
    gradient-descent path: probe [K][S][S], the mixed-state overlap, Fourier
    projection and adjoint (ptychographic_methods.py:2696-2805, 2906-3000) and
    the mixed probe constraints (ptychographic_constraints.py:1150-1286)."""

    def __init__(self, energy, datacube=None, num_probes=None, initial_probe_guess=None,
                 **kwargs):
        # mixedstate_ptychography.py:150-168
        if initial_probe_guess is None:
            if num_probes is None:
                raise ValueError("If initial_probe_guess is None, or a ComplexProbe object, "
                                 "num_probes must be specified.")
        else:
            if len(initial_probe_guess.shape) != 3:
                raise ValueError("Specified initial_probe_guess must have dimensions (N,Sx,Sy).")
            num_probes = initial_probe_guess.shape[0]
        self._num_probes = num_probes
        super().__init__(energy, datacube=datacube, initial_probe_guess=initial_probe_guess,
                         **kwargs)

    def preprocess(self, **kwargs):
        # the FOV mask comes from the first mode (mixedstate_ptychography.py:548-563)
        given = None if self._probe is None else np.asarray(self._probe, dtype=np.complex64)
        if given is not None:
            self._probe = given[0]
        super().preprocess(**kwargs)
        self._probe = (given.copy() if given is not None
                       else mixed_probe_modes(self._probe, self._num_probes))
        self._probe_initial = self._probe.copy()
        self._probe_initial_aperture = np.abs(_fft2(self._probe, self._mode))
        return self

    def forward(self, obj, probe, positions_px, amplitudes, fourier_mask=None):
        """Mixed-state _forward for GD: shifted probes [B][K], far field
        sqrt(sum_k |F_k|^2), chi_k = ifft2(A/far F_k - F_k)."""
        mode = self._mode
        frac = positions_px - np.round(positions_px)
        rows, cols = patch_indices(positions_px, self._region_of_interest_shape, obj.shape)
        shifted_probes = fft_shift(probe, frac, mode)
        object_patches = obj[rows, cols]
        overlap = shifted_probes * np.expand_dims(object_patches, axis=1)
        fourier_overlap = _fft2(overlap, mode)
        if fourier_mask is not None:
            fourier_overlap *= fourier_mask
        farfield_amplitudes = np.sqrt(np.sum(np.abs(fourier_overlap) ** 2, axis=1))
        error = np.sum(np.abs(amplitudes - farfield_amplitudes) ** 2)
        farfield_amplitudes[farfield_amplitudes == 0.0] = np.inf
        amplitude_modification = amplitudes / farfield_amplitudes
        fourier_modified_overlap = amplitude_modification[:, None] * fourier_overlap
        fourier_modified_overlap = fourier_modified_overlap - fourier_overlap
        if fourier_mask is not None:
            fourier_modified_overlap *= fourier_mask
        exit_waves = _ifft2(fourier_modified_overlap, mode)
        return shifted_probes, object_patches, exit_waves, error

    def adjoint(self, obj, probe, object_patches, shifted_probes, positions_px, exit_waves,
                step_size, normalization_min, fix_probe):
        """Mixed-state _gradient_descent_adjoint for a complex object: per-mode
        scatters added in float32 (complex64 accumulators, as the reference
        allocates them with zeros_like(object))."""
        roi = self._region_of_interest_shape
        shape = obj.shape
        probe_normalization = np.zeros_like(obj)
        object_update = np.zeros_like(obj)
        for i_probe in range(probe.shape[0]):
            probe_normalization += scatter_bincount(np.abs(shifted_probes[:, i_probe]) ** 2,
                                                    positions_px, roi, shape)
            object_update += step_size * scatter_bincount(
                np.conj(shifted_probes[:, i_probe]) * exit_waves[:, i_probe], positions_px, roi,
                shape)
        probe_normalization = 1 / np.sqrt(
            1e-16
            + ((1 - normalization_min) * probe_normalization) ** 2
            + (normalization_min * np.max(probe_normalization)) ** 2
        )
        obj += object_update * probe_normalization
        if not fix_probe:
            object_normalization = np.sum(np.abs(object_patches) ** 2, axis=0)
            object_normalization = 1 / np.sqrt(
                1e-16
                + ((1 - normalization_min) * object_normalization) ** 2
                + (normalization_min * np.max(object_normalization)) ** 2
            )
            probe += step_size * (
                np.sum(np.expand_dims(np.conj(object_patches), axis=1) * exit_waves, axis=0)
                * object_normalization[None]
            )
        return obj, probe

    def probe_com_constraint(self, probe):
        """ptychographic_constraints.py:1156-1178: CoM of mode 0, every mode
        shifted by it."""
        intensity = np.abs(probe[0]) ** 2
        x0, y0 = get_CoM(intensity, corner_centered=True)
        return fft_shift(probe, -np.array([x0, y0]), self._mode)

    def probe_extra_constraints(self, probe):
        if self._orthogonalize_probe:
            return probe_orthogonalization(probe)
        return probe

    def reconstruct(self, *args, orthogonalize_probe=True, **kwargs):
        """mixedstate_ptychography.py:631-1027 (GD)."""
        method = kwargs.get("reconstruction_method", "gradient-descent")
        if method not in ("GD", "gradient-descent"):
            raise NotImplementedError("the mixed-state oracle restates gradient descent only")
        if getattr(self, "_object_type", "complex") == "potential":
            raise NotImplementedError("the mixed-state oracle restates complex objects only")
        self._orthogonalize_probe = orthogonalize_probe
        return super().reconstruct(*args, **kwargs)
