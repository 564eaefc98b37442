"""Drop-in ``SingleslicePtychography`` running its GD loop on MI355X.

Keeps the constructor / ``preprocess()`` / ``reconstruct()`` signatures,
return-self chaining, error behaviour and result attributes of
py4DSTEM's class (py4DSTEM/process/phase/singleslice_ptychography.py:42-996),
so callers such as ``PtychographyOptimizer``
(phase/parameter_optimize.py:580-660) can switch by import.

``preprocess()`` uploads the patterns once; their centre of mass, the
bilinear shift + sqrt into kernel-layout amplitudes and the probe-overlap
field run in libpty_hip.so, the small per-scan fits (origin plane, rotation
search, FOV mask filter) on the host.  Every ``reconstruct()`` iteration runs
in libpty_hip.so (one native call per run).  Options of the reference that
this build does not implement raise ``NotImplementedError`` rather than being
silently ignored.
"""

from __future__ import annotations

import sys
import warnings

import numpy as np
import scipy.ndimage
import torch

from . import kernels as K
from ._lib import PtyGdParams, require_gpu
from .resample import axis_maps, resample_frames
from .preprocess import (
    ScaledQCalibration,
    bilinear_shift,
    center_of_mass,
    center_positions,
    extract_calibrations,
    fit_origin,
    initial_object_shape,
    real_space_sampling,
    scan_positions_in_pixels,
    solve_com_rotation,
)
from .probe import POLAR_ALIASES, POLAR_SYMBOLS, ComplexProbe, polar_parameters_from

warnings.simplefilter("always", UserWarning)

_GD_NAMES = ("GD", "gradient-descent")


def _method_parameters(method, parameter, a, b, c):
    """(projection_a, projection_b, projection_c) of a projection-set method,
    or None for gradient descent (phase_base_class.py:2005-2108, same
    ValueErrors)."""
    if method == "generalized-projections":
        if a is None or b is None or c is None:
            raise ValueError("reconstruction_parameter_a/b/c must all be specified "
                             "when using reconstruction_method='generalized-projections'.")
        return (a, b, c)
    if method in ("DM_AP", "difference-map_alternating-projections"):
        if parameter < 0.0 or parameter > 1.0:
            raise ValueError("reconstruction_parameter must be between 0-1.")
        return (-parameter, 1, 1 + parameter)
    if method in ("RAAR", "relaxed-averaged-alternating-reflections"):
        if parameter < 0.0 or parameter > 1.0:
            raise ValueError("reconstruction_parameter must be between 0-1.")
        return (1 - 2 * parameter, parameter, 2)
    if method in ("RRR", "relax-reflect-reflect"):
        if parameter < 0.0 or parameter > 2.0:
            raise ValueError("reconstruction_parameter must be between 0-2.")
        return (-parameter, parameter, 2)
    if method in ("SUPERFLIP", "charge-flipping"):
        return (0, 1, 2)
    if method in _GD_NAMES:
        return None
    raise ValueError(
        "reconstruction_method must be one of 'generalized-projections', "
        "'DM_AP' (or 'difference-map_alternating-projections'), "
        "'RAAR' (or 'relaxed-averaged-alternating-reflections'), "
        "'RRR' (or 'relax-reflect-reflect'), "
        "'SUPERFLIP' (or 'charge-flipping'), "
        f"or 'GD' (or 'gradient-descent'), not  {method}.")


def _warn(msg):
    print(msg, file=sys.stderr)


def subdivide_into_batches(num_items, max_batch):
    """Near-equal batch sizes, larger first (phase/utils.py:539-579)."""
    num_batches = (num_items + (-num_items % max_batch)) // max_batch
    if num_items < num_batches:
        raise RuntimeError("num_batches may not be larger than num_items")
    base, extra = divmod(num_items, num_batches)
    return [base + 1] * extra + [base] * (num_batches - extra)


class _DeviceMirror:
    """A host array not yet downloaded: the device tensor holding its values
    (never written afterwards)."""

    __slots__ = ("tensor",)

    def __init__(self, tensor):
        self.tensor = tensor

    def download(self):
        return self.tensor.cpu().numpy()


def _host_attr(name):
    """Property for a host array that may be held as a _DeviceMirror until
    it is first read."""
    key = name + "_host"

    def get(self):
        v = self.__dict__.get(key)
        if isinstance(v, _DeviceMirror):
            v = v.download()
            self.__dict__[key] = v
        return v

    def put(self, value):
        self.__dict__[key] = value

    return property(get, put)


class SingleslicePtychography:
    """Single-slice ptychographic reconstruction (GD) on MI355X.

    Parameters mirror py4DSTEM's ``SingleslicePtychography``
    (singleslice_ptychography.py:53-116).  ``device`` / ``storage`` accept
    'cpu' or 'gpu' for call compatibility; the iteration always runs in the
    HIP library on the current ROCm device.
    """

    _class_specific_metadata = ()

    # the starting object after the learned initialiser, and the object after
    # a reset from a device-resident start, stay on the device until the host
    # array is read
    _object = _host_attr("_object")
    _object_initial = _host_attr("_object_initial")

    def __init__(self, energy, datacube=None, semiangle_cutoff=None, semiangle_cutoff_pixels=None,
                 rolloff=2.0, vacuum_probe_intensity=None, polar_parameters=None,
                 initial_object_guess=None, initial_probe_guess=None, initial_scan_positions=None,
                 object_fov_ang=None, positions_offset_ang=None, object_padding_px=None,
                 object_type="complex", positions_mask=None, verbose=True, device="cpu",
                 storage=None, clear_fft_cache=True, name="ptychographic_reconstruction",
                 **kwargs):
        if storage is None:
            storage = device
        self.set_device(device, clear_fft_cache)
        self.set_storage(storage)
        for key in kwargs:
            if key not in POLAR_SYMBOLS and key not in POLAR_ALIASES:
                raise ValueError("{} not a recognized parameter".format(key))
        merged = dict(polar_parameters or {})
        merged.update(kwargs)
        self._polar_parameters = polar_parameters_from(merged)
        if object_type != "potential" and object_type != "complex":
            raise ValueError(
                f"object_type must be either 'potential' or 'complex', not {object_type}")
        self.name = name
        self.set_save_defaults()
        self._datacube = datacube
        self._object = initial_object_guess
        self._probe = initial_probe_guess
        self._vacuum_probe_intensity = vacuum_probe_intensity
        self._scan_positions = initial_scan_positions
        self._positions_offset_ang = positions_offset_ang
        self._energy = energy
        self._semiangle_cutoff = semiangle_cutoff
        self._semiangle_cutoff_pixels = semiangle_cutoff_pixels
        self._rolloff = rolloff
        self._object_type = object_type
        self._object_padding_px = object_padding_px
        self._object_fov_ang = object_fov_ang
        self._positions_mask = positions_mask
        self._verbose = verbose
        self._preprocessed = False

    # -- device / bookkeeping (phase_base_class.py:46-221) -------------------
    def set_device(self, device, clear_fft_cache=None):
        if clear_fft_cache is not None:
            self._clear_fft_cache = clear_fft_cache
        if device is None:
            return self
        if device not in ("cpu", "gpu"):
            raise ValueError(f"device must be either 'cpu' or 'gpu', not {device}")
        self._device = device
        return self

    def set_storage(self, storage):
        if storage not in ("cpu", "gpu"):
            raise ValueError(f"storage must be either 'cpu' or 'gpu', not {storage}")
        if storage == "gpu" and getattr(self, "_device", "cpu") == "cpu":
            raise ValueError("storage='gpu' and device='cpu' is not supported")
        self._storage = storage
        return self

    def set_save_defaults(self, save_datacube=False, save_exit_waves=False, save_iterations=True,
                          save_iterations_frequency=1):
        self._save_datacube = save_datacube
        self._save_exit_waves = save_exit_waves
        self._save_iterations = save_iterations
        self._save_iterations_frequency = save_iterations_frequency
        return self

    def attach_datacube(self, datacube):
        self._datacube = datacube
        return self

    # -- properties (phase_base_class.py:2197-2227; ptychographic_methods.py) --
    @property
    def angular_sampling(self):
        return getattr(self, "_angular_sampling", None)

    @property
    def sampling(self):
        if self.angular_sampling is None:
            return None
        return real_space_sampling(self._energy, self._angular_sampling, self._amplitudes_shape)

    @property
    def positions(self):
        if self.angular_sampling is None:
            return None
        pos = np.array(self._positions_px, dtype=np.float32, copy=True)
        pos[:, 0] *= self.sampling[0]
        pos[:, 1] *= self.sampling[1]
        return pos

    @property
    def probe_centered(self):
        if getattr(self, "_probe", None) is None:
            return None
        return np.fft.fftshift(np.asarray(self._probe, dtype=np.complex64), axes=(-2, -1))

    @property
    def probe_fourier(self):
        if getattr(self, "_probe", None) is None:
            return None
        return np.fft.fftshift(np.fft.fft2(np.asarray(self._probe)), axes=(-2, -1))

    @property
    def object_cropped(self):
        """Object rotated into the scan frame and cropped to the scanned FOV
        (ptychographic_methods.py:85-124)."""
        angle = self._rotation_best_rad if self._rotation_best_transpose else -self._rotation_best_rad
        pos = np.asarray(self._positions_px, dtype=np.float64)
        c, s = np.cos(angle), np.sin(angle)
        rot = np.array([[c, -s], [s, c]], dtype=np.float32)
        origin = pos.mean(0).astype(np.float32)
        pts = (pos - origin) @ rot + origin
        min_x, min_y = np.floor(np.amin(pts, axis=0)).astype(int)
        min_x, min_y = max(min_x, 0), max(min_y, 0)
        max_x, max_y = np.ceil(np.amax(pts, axis=0)).astype(int)
        obj = np.asarray(self.object)
        rotated = scipy.ndimage.rotate(obj, np.rad2deg(-angle), order=1, reshape=False,
                                       axes=(-2, -1))[..., min_x:max_x, min_y:max_y]
        if self._rotation_best_transpose:
            rotated = rotated.swapaxes(-2, -1)
        return rotated

    # -- preprocess -----------------------------------------------------------
    def preprocess(self, diffraction_intensities_shape=None, reshaping_method="bilinear",
                   padded_diffraction_intensities_shape=None, region_of_interest_shape=None,
                   dp_mask=None, in_place_datacube_modification=False, fit_function="plane",
                   plot_center_of_mass="default", plot_rotation=True, maximize_divergence=False,
                   rotation_angles_deg=None, plot_probe_overlaps=True, force_com_rotation=None,
                   force_com_transpose=None, force_com_shifts=None, force_com_measured=None,
                   vectorized_com_calculation=True, force_scan_sampling=None,
                   force_angular_sampling=None, force_reciprocal_sampling=None,
                   object_fov_mask=None, crop_patterns=False, center_positions_in_fov=True,
                   store_initial_arrays=True, device=None, clear_fft_cache=None,
                   max_batch_size=None, **kwargs):
        """singleslice_ptychography.py:191-604.  Returns self."""
        self.set_device(device, clear_fft_cache)
        if self._datacube is None:
            raise ValueError("The preprocess() method requires a DataCube. "
                             "Please run ptycho.attach_datacube(DataCube) first.")
        reshape = (diffraction_intensities_shape is not None
                   or padded_diffraction_intensities_shape is not None)
        if reshape and (self._vacuum_probe_intensity is not None or dp_mask is not None):
            raise NotImplementedError("resampling or padding together with vacuum_probe_intensity or "
                                      "dp_mask is not supported")
        if crop_patterns:
            raise NotImplementedError("crop_patterns is not supported")

        dc = self._datacube
        Rshape = tuple(dc.data.shape[:2])
        if force_com_shifts is not None and np.isscalar(force_com_shifts[0]):
            force_com_shifts = (np.ones(Rshape) * force_com_shifts[0],
                                np.ones(Rshape) * force_com_shifts[1])
        if force_com_measured is not None and np.isscalar(force_com_measured[0]):
            force_com_measured = (np.ones(Rshape) * force_com_measured[0],
                                  np.ones(Rshape) * force_com_measured[1])
        if self._positions_mask is not None:
            self._positions_mask = np.asarray(self._positions_mask, dtype=bool)

        intensities = np.asarray(dc.data, dtype=np.float32)
        self._grid_scan_shape = intensities.shape[:2]
        cal_source = dc
        if reshape:
            # phase_base_class.py:223-420: resample, then pad, every pattern;
            # forced centres scale with the resampling (:277-300)
            maps = axis_maps(intensities.shape[-2:], diffraction_intensities_shape,
                             reshaping_method, padded_diffraction_intensities_shape)
            cal_source = ScaledQCalibration(dc, maps[2])
            if diffraction_intensities_shape is not None:
                rf = diffraction_intensities_shape[0] / intensities.shape[-2]
                if force_com_shifts is not None:
                    force_com_shifts = (force_com_shifts[0] * rf, force_com_shifts[1] * rf)
                if force_com_measured is not None:
                    force_com_measured = (force_com_measured[0] * rf, force_com_measured[1] * rf)
        (self._scan_sampling, self._angular_sampling,
         self._reciprocal_sampling) = extract_calibrations(
            cal_source, self._energy, True, force_scan_sampling, force_angular_sampling,
            force_reciprocal_sampling, self._verbose)
        shape = (maps[0].shape[0], maps[1].shape[0]) if reshape else intensities.shape[-2:]
        if shape[0] != shape[1]:
            raise NotImplementedError("only square diffraction patterns are supported")
        S = int(shape[0])
        K.check_frame(S)
        if region_of_interest_shape is not None and tuple(
                int(v) for v in np.ravel(region_of_interest_shape)) != (S, S):
            # the reference resamples every far field between the two shapes
            # (ptychographic_methods.py:1603-1631); at the patterns' own shape
            # the order-1 zoom by 1 keeps the values, and its float64 divisor
            # only moves the projection to complex128, so this path runs the
            # plain complex64 projection (within 1e-4 of the reference's run)
            raise NotImplementedError("region_of_interest_shape different from the diffraction "
                                      "pattern shape (resampled exit waves) is not supported")
        if self._semiangle_cutoff_pixels:
            self._semiangle_cutoff = self._semiangle_cutoff_pixels * self._angular_sampling[0]

        # the patterns go to the GPU once; centre of mass and normalisation
        # run there (phase_base_class.py:613-767, 1347-1460)
        dev = require_gpu()
        Qx, Qy = intensities.shape[-2:]
        frames_dev = torch.from_numpy(np.ascontiguousarray(intensities).reshape(-1, Qx, Qy)).to(dev)
        if reshape:
            frames_dev = resample_frames(frames_dev, maps[0], maps[1])
            # only the shape of the host array is used from here on
            intensities = np.broadcast_to(np.float32(0), tuple(self._grid_scan_shape) + (S, S))

        # centre of mass and origin fit (phase_base_class.py:613-767)
        if force_com_measured:
            com_mx = np.asarray(force_com_measured[0], dtype=np.float32)
            com_my = np.asarray(force_com_measured[1], dtype=np.float32)
        else:
            if dp_mask is not None and dp_mask.shape != intensities.shape[-2:]:
                raise ValueError(f"Mask shape should be (Qx,Qy):{intensities.shape[-2:]}, "
                                 f"not {dp_mask.shape}")
            com_mx, com_my = K.pattern_com(frames_dev, dp_mask)
            com_mx = com_mx.reshape(intensities.shape[:2])
            com_my = com_my.reshape(intensities.shape[:2])
        if force_com_shifts is None:
            if fit_function is not None:
                finite = np.isfinite(com_mx)
                com_fx = np.asarray(fit_origin(com_mx, finite, fit_function), dtype=np.float32)
                com_fy = np.asarray(fit_origin(com_my, finite, fit_function), dtype=np.float32)
            else:
                com_fx = np.asarray(com_mx, dtype=np.float32)
                com_fy = np.asarray(com_my, dtype=np.float32)
        else:
            com_fx = np.asarray(force_com_shifts[0], dtype=np.float32)
            com_fy = np.asarray(force_com_shifts[1], dtype=np.float32)
        self._com_measured_x, self._com_measured_y = com_mx, com_my
        self._com_fitted_x, self._com_fitted_y = com_fx, com_fy
        self._com_normalized_x = np.nan_to_num(com_mx - com_fx) * self._reciprocal_sampling[0]
        self._com_normalized_y = np.nan_to_num(com_my - com_fy) * self._reciprocal_sampling[1]

        (self._rotation_best_rad, self._rotation_best_transpose, self._com_x,
         self._com_y) = solve_com_rotation(
            self._com_normalized_x, self._com_normalized_y, rotation_angles_deg,
            maximize_divergence, force_com_rotation, force_com_transpose, self._verbose)

        # corner-centred amplitudes (phase_base_class.py:1347-1460), kernel layout
        self._amplitudes_dev, self._mean_diffraction_intensity = self._normalize_intensities(
            frames_dev, com_fx, com_fy)
        self._amplitudes_host = None
        self._crop_mask, self._crop_mask_shape = None, intensities.shape[-2:]
        del intensities, frames_dev
        self._num_diffraction_patterns = self._amplitudes_dev.shape[0]
        self._amplitudes_shape = np.array([S, S])
        self._resample_exit_waves = False
        self._region_of_interest_shape = np.array([S, S])

        # positions, object, probe
        pos, self._object_padding_px = scan_positions_in_pixels(
            self._scan_positions, self._grid_scan_shape, self._scan_sampling,
            self._rotation_best_rad, self._rotation_best_transpose, self.sampling,
            self._region_of_interest_shape, self._object_padding_px, self._positions_mask,
            self._positions_offset_ang)
        if self._object is None:
            shape = initial_object_shape(pos, self._object_padding_px, self._region_of_interest_shape,
                                         self._object_fov_ang, self.sampling)
            # ptychographic_methods.py:40-79: zero potential or unit transmission
            self._object = (np.zeros(shape, dtype=np.float32) if self._object_type == "potential"
                            else np.ones(shape, dtype=np.complex64))
        else:
            self._object = np.array(self._object, dtype=self._object_dtype(), copy=True)
        if self._object.shape[0] < S or self._object.shape[1] < S:
            raise ValueError(f"object {self._object.shape} is smaller than the {S}x{S} probe window")
        if store_initial_arrays:
            self._object_initial_dev = None
            self._object_initial = self._object.copy()
            self._object_type_initial = self._object_type
        self._object_shape = self._object.shape
        self._positions_px = np.asarray(pos, dtype=np.float32)
        if center_positions_in_fov:
            center_positions(self._positions_px, self._object_shape)
        self._positions_px_initial_com = self._positions_px.mean(0)
        self._positions_px_initial = self._positions_px.copy()
        self._positions_initial = self._positions_px_initial.copy()
        self._positions_initial[:, 0] *= self.sampling[0]
        self._positions_initial[:, 1] *= self.sampling[1]

        self._probe = self._initialize_probe(self._probe)
        self._known_aberrations_array = ComplexProbe(
            energy=self._energy, gpts=self._region_of_interest_shape, sampling=self.sampling,
            parameters=self._polar_parameters)._evaluate_ctf()
        if store_initial_arrays:
            self._probe_initial = self._probe.copy()
            self._probe_initial_aperture = np.abs(np.fft.fft2(self._probe))
        else:
            self._probe_initial_aperture = None

        # device-resident positions for the loop (amplitudes are already there)
        self._positions_dev = torch.from_numpy(self._positions_px.copy()).to(dev)
        self._positions_initial_dev = torch.from_numpy(self._positions_px_initial.copy()).to(dev)

        # probe-overlap field and FOV mask (singleslice_ptychography.py:504-537)
        if object_fov_mask is None or plot_probe_overlaps:
            self._probe_overlap = self._probe_overlap_field(max_batch_size)
        if object_fov_mask is None:
            blurred = scipy.ndimage.gaussian_filter(self._probe_overlap, 1.0)
            self._object_fov_mask = blurred > 0.25 * blurred.max()
        elif object_fov_mask is True:
            self._object_fov_mask = np.full(self._object_shape, True)
        else:
            self._object_fov_mask = np.asarray(object_fov_mask)
        self._object_fov_mask_inverse = np.invert(self._object_fov_mask)
        self._preprocessed = True
        if hasattr(self, "_probe_dev"):
            del self._probe_dev
        # the starting object stays resident: resets and the learned
        # initialiser's judge use device copies
        self._object_dev = torch.from_numpy(np.ascontiguousarray(self._object,
                                                                 dtype=self._object_dtype())).to(dev)
        if store_initial_arrays:
            self._object_initial_dev = self._object_dev.clone()
        return self

    def _normalize_intensities(self, frames_dev, com_fx, com_fy):
        """Shift each pattern by -CoM (bilinear, periodic), take
        sqrt(max(I, 0)) into the kernel layout on the GPU; the float32 mean
        intensity is accumulated pattern by pattern in scan order, as the
        reference loop does."""
        n_in, S, _ = frames_dev.shape
        fx = (-np.asarray(com_fx, dtype=np.float32)).reshape(-1)
        fy = (-np.asarray(com_fy, dtype=np.float32)).reshape(-1)
        sel = None
        if self._positions_mask is not None:
            sel = np.nonzero(self._positions_mask.reshape(-1))[0].astype(np.int32)
        amps, sums = K.shift_amplitudes(frames_dev, fx, fy, sel)
        mean_intensity = np.float32(0)
        for v in sums.cpu().numpy().astype(np.float32):
            mean_intensity += v
        mean_intensity /= amps.shape[0]
        return amps, mean_intensity

    @property
    def _amplitudes(self):
        """Host copy of the corner-centred amplitudes [N][S][S] (natural
        order), downloaded from the device on first use."""
        if getattr(self, "_amplitudes_host", None) is None:
            if not hasattr(self, "_amplitudes_dev"):
                raise AttributeError("_amplitudes: preprocess() has not run")
            self._amplitudes_host = K.to_kernel_layout(self._amplitudes_dev, inverse=True).cpu().numpy()
        return self._amplitudes_host

    def _object_dtype(self):
        """float32 potential or complex64 transmission (ptychographic_methods.py:69-77)."""
        return np.float32 if self._object_type == "potential" else np.complex64

    def _initialize_probe(self, initial_probe):
        """ptychographic_methods.py:1011-1102."""
        if initial_probe is None:
            vacuum = None
            if self._vacuum_probe_intensity is not None:
                # the measured vacuum intensity, its centre of mass moved to
                # the corner by a periodic bilinear shift, is the aperture
                # (:1033-1064); the semiangle cut-off no longer applies
                vacuum = np.asarray(self._vacuum_probe_intensity, dtype=np.float32)
                x0, y0 = center_of_mass(vacuum)
                vacuum = bilinear_shift(vacuum, -x0, -y0)
                if vacuum.shape != tuple(self._region_of_interest_shape):
                    raise NotImplementedError(
                        f"vacuum_probe_intensity of shape {vacuum.shape} would need resampling to "
                        f"{tuple(self._region_of_interest_shape)}, which is not supported")
                self._semiangle_cutoff = np.inf
            probe = ComplexProbe(energy=self._energy, gpts=self._region_of_interest_shape,
                                 sampling=self.sampling,
                                 semiangle_cutoff=(np.inf if vacuum is not None
                                                   else self._semiangle_cutoff),
                                 rolloff=self._rolloff, vacuum_probe_intensity=vacuum,
                                 parameters=self._polar_parameters).build()._array
            intensity = np.sum(np.abs(np.fft.fft2(probe)) ** 2)
            probe *= np.sqrt(self._mean_diffraction_intensity / intensity)
            return probe.astype(np.complex64)
        if isinstance(initial_probe, ComplexProbe):
            if tuple(initial_probe._gpts) != tuple(self._region_of_interest_shape):
                raise ValueError()
            probe = (initial_probe._array if hasattr(initial_probe, "_array")
                     else initial_probe.build()._array)
            intensity = np.sum(np.abs(np.fft.fft2(probe)) ** 2)
            probe = probe * np.sqrt(self._mean_diffraction_intensity / intensity)
            return probe.astype(np.complex64)
        return np.array(initial_probe, dtype=np.complex64, copy=True)

    def _probe_overlap_field(self, max_batch_size=None):
        """sum over batches of scatter(|fft_shift(P, frac)|^2), float32."""
        dev = self._amplitudes_dev.device
        N = self._num_diffraction_patterns
        B = N if max_batch_size is None else int(max_batch_size)
        probe = torch.from_numpy(np.ascontiguousarray(self._probe)).to(dev)
        phat = K.probe_spectrum(probe)
        ones = torch.ones(self._object_shape, dtype=torch.complex64, device=dev)
        overlap = torch.zeros(self._object_shape, dtype=torch.float32, device=dev)
        start = 0
        for nb in subdivide_into_batches(N, B):
            order = torch.arange(start, start + nb, dtype=torch.int32, device=dev)
            ps, chi, _ = K.forward_adjoint(phat, ones, self._positions_dev, self._amplitudes_dev, order)
            _, spsum, _ = K.scatter(ps, chi, self._positions_dev, order, self._object_shape)
            overlap += spsum
            start += nb
            del ps, chi
        return overlap.cpu().numpy()

    # -- learned starting object ----------------------------------------------------
    def initialize_object_learned(self, model=None, stride=4, judge_positions=512):
        """Replace the starting object with the learned initialiser's stitched
        prediction (py4dstem_amd.learned_init) when its data error on
        `judge_positions` seeded positions is below the current starting
        object's; otherwise keep it (flat-object fallback).  Call after
        preprocess() and before reconstruct(); the choice and timings are in
        ``learned_init_info``.  Returns self."""
        from . import learned_init
        if self._object_type == "potential":
            raise NotImplementedError("the learned initialiser predicts complex transmissions")
        if not getattr(self, "_preprocessed", False):
            raise ValueError("preprocess() must run before initialize_object_learned()")
        if model is None or isinstance(model, str):
            model = learned_init.ConvNet.load(model or learned_init.DEFAULT_MODEL)
        obj, used, info = learned_init.initialize_object(self, model, stride=stride,
                                                         judge_positions=judge_positions)
        self.learned_init_info = info
        if used:
            # the next run (and any reset) starts from the device copy; the
            # host arrays are downloaded when first read
            self._object = _DeviceMirror(obj)
            self._object_initial = _DeviceMirror(obj)
            self._object_initial_dev = obj
            self._object_dev = obj.clone()
            if hasattr(self, "_probe_dev"):
                del self._probe_dev
        return self

    # -- execution tuning (results do not depend on it) --------------------------
    def set_stage(self, stage, pipeline=False):
        """Positions per fused launch inside a mini-batch (0 = whole batch),
        and whether gradient descent pipelines those launches: the scatter of
        one stage runs on a second stream beside the fused pass of the next.
        Performance settings only: the update is bitwise identical."""
        self._stage = int(stage)
        self._pipeline = bool(pipeline)
        return self

    def _stage_for(self, batch):
        """(positions per fused launch, pipelined) for mini-batches of `batch`."""
        stage = getattr(self, "_stage", None)
        if stage is None:
            from .autotune import default_execution
            return default_execution(int(self._region_of_interest_shape[0]), batch)
        return stage, getattr(self, "_pipeline", False)

    # -- reconstruct ------------------------------------------------------------
    def _reset_reconstruction(self, store_iterations, reset):
        """ptychographic_methods.py:252-294."""
        if store_iterations and (not hasattr(self, "object_iterations") or reset):
            self.object_iterations = []
            self.probe_iterations = []
        if reset is True:
            self.error_iterations = []
            init_dev = getattr(self, "_object_initial_dev", None)
            # (the device copy holds the values of _object_initial)
            self._object = (_DeviceMirror(init_dev) if init_dev is not None
                            else self._object_initial.copy())
            self._probe = self._probe_initial.copy()
            self._positions_px = self._positions_px_initial.copy()
            if hasattr(self, "_positions_initial_dev"):
                self._positions_dev.copy_(self._positions_initial_dev)
            self._object_type = self._object_type_initial
            self._exit_waves = None
            for attr in ("_object_dev", "_probe_dev"):
                if hasattr(self, attr):
                    delattr(self, attr)
            if init_dev is not None:
                self._object_dev = init_dev.clone()
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

    def reconstruct(self, num_iter=8, reconstruction_method="gradient-descent",
                    reconstruction_parameter=1.0, reconstruction_parameter_a=None,
                    reconstruction_parameter_b=None, reconstruction_parameter_c=None,
                    max_batch_size=None, seed_random=None, step_size=0.5, normalization_min=1,
                    positions_step_size=0.5, pure_phase_object=False, fix_probe_com=True,
                    fix_probe=False, fix_probe_aperture=False, constrain_probe_amplitude=False,
                    constrain_probe_amplitude_relative_radius=0.5,
                    constrain_probe_amplitude_relative_width=0.05,
                    constrain_probe_fourier_amplitude=False,
                    constrain_probe_fourier_amplitude_max_width_pixels=3.0,
                    constrain_probe_fourier_amplitude_constant_intensity=False, fix_positions=True,
                    fix_positions_com=True, max_position_update_distance=None,
                    max_position_total_distance=None, global_affine_transformation=False,
                    gaussian_filter_sigma=None, gaussian_filter=True, fit_probe_aberrations=False,
                    fit_probe_aberrations_max_angular_order=4,
                    fit_probe_aberrations_max_radial_order=4,
                    fit_probe_aberrations_remove_initial=False,
                    fit_probe_aberrations_using_scikit_image=True, butterworth_filter=True,
                    q_lowpass=None, q_highpass=None, butterworth_order=2, tv_denoise=True,
                    tv_denoise_weight=None, tv_denoise_inner_iter=40, object_positivity=True,
                    shrinkage_rad=0.0, fix_potential_baseline=True, detector_fourier_mask=None,
                    store_iterations=False, progress_bar=True, reset=None, device=None,
                    clear_fft_cache=None, object_type=None):
        """singleslice_ptychography.py:606-996 (GD).  Returns self."""
        kwargs = {k: v for k, v in locals().items() if k != "self"}
        run = self._prepare_run(**kwargs)
        self._execute_run(run)
        return self._finish_run(run)

    def _prepare_run(self, num_iter=8, reconstruction_method="gradient-descent",
                    reconstruction_parameter=1.0, reconstruction_parameter_a=None,
                    reconstruction_parameter_b=None, reconstruction_parameter_c=None,
                    max_batch_size=None, seed_random=None, step_size=0.5, normalization_min=1,
                    positions_step_size=0.5, pure_phase_object=False, fix_probe_com=True,
                    fix_probe=False, fix_probe_aperture=False, constrain_probe_amplitude=False,
                    constrain_probe_amplitude_relative_radius=0.5,
                    constrain_probe_amplitude_relative_width=0.05,
                    constrain_probe_fourier_amplitude=False,
                    constrain_probe_fourier_amplitude_max_width_pixels=3.0,
                    constrain_probe_fourier_amplitude_constant_intensity=False, fix_positions=True,
                    fix_positions_com=True, max_position_update_distance=None,
                    max_position_total_distance=None, global_affine_transformation=False,
                    gaussian_filter_sigma=None, gaussian_filter=True, fit_probe_aberrations=False,
                    fit_probe_aberrations_max_angular_order=4,
                    fit_probe_aberrations_max_radial_order=4,
                    fit_probe_aberrations_remove_initial=False,
                    fit_probe_aberrations_using_scikit_image=True, butterworth_filter=True,
                    q_lowpass=None, q_highpass=None, butterworth_order=2, tv_denoise=True,
                    tv_denoise_weight=None, tv_denoise_inner_iter=40, object_positivity=True,
                    shrinkage_rad=0.0, fix_potential_baseline=True, detector_fourier_mask=None,
                    store_iterations=False, progress_bar=True, reset=None, device=None,
                    clear_fft_cache=None, object_type=None, _allocate=True):
        """Validation, batch orders and device state for reconstruct()."""
        self.set_device(device, clear_fft_cache)
        if not getattr(self, "_preprocessed", False):
            raise ValueError("preprocess() must run before reconstruct()")
        self._reset_reconstruction(store_iterations, reset)
        if object_type is not None and object_type != self._object_type:
            raise NotImplementedError("switching object_type is not supported")

        # method parameters (phase_base_class.py:2005-2108)
        projection = _method_parameters(reconstruction_method, reconstruction_parameter,
                                        reconstruction_parameter_a, reconstruction_parameter_b,
                                        reconstruction_parameter_c)
        if projection is not None:
            # the reference raises this from its verbose summary
            # (phase_base_class.py:2130-2137); it is raised here whatever the
            # verbosity, since mini-batches would mix exit waves of different
            # positions
            if max_batch_size is not None:
                raise ValueError(
                    "Stochastic object/probe updating is inconsistent with 'DM_AP', 'RAAR', "
                    "'RRR', and 'SUPERFLIP'. Use reconstruction_method='GD' or set "
                    "max_batch_size=None.")
            if detector_fourier_mask is not None:  # ptychographic_methods.py:1686-1687
                raise NotImplementedError()
        tv = tv_denoise and tv_denoise_weight is not None
        unsupported = {
            # the reference runs pylops' split Bregman solver here
            # (ptychographic_constraints.py:233-256)
            "tv_denoise_weight with object_type='potential'": tv and self._object_type == "potential",
            "fit_probe_aberrations": fit_probe_aberrations and not fix_probe,
            # the aperture constraint takes precedence (ptychographic_constraints.py:1127-1137)
            "constrain_probe_fourier_amplitude": (constrain_probe_fourier_amplitude
                                                  and not fix_probe and not fix_probe_aperture),
        }
        bad = [k for k, v in unsupported.items() if v]
        if bad:
            raise NotImplementedError("not supported on this path: " + ", ".join(bad))
        if tv:
            # complex objects: TV denoising leaves the object as it is and only
            # warns (ptychographic_constraints.py:224-231, called every
            # iteration from :443-446)
            warnings.warn("TV denoising is currently only supported for object_type=='potential'.",
                          UserWarning)

        N = self._num_diffraction_patterns
        S = int(self._region_of_interest_shape[0])
        if S not in K.LDS_FRAMES:
            # 256 px frames: gradient descent through line passes (no LDS frame)
            for what, used in (("projection-set methods", projection is not None),
                               ("position correction", not fix_positions),
                               ("fix_probe_aperture", fix_probe_aperture and not fix_probe),
                               ("constrain_probe_amplitude",
                                constrain_probe_amplitude and not fix_probe)):
                if used:
                    raise NotImplementedError(f"{what} with {S} px frames is not supported on "
                                              "this path (gradient descent only)")
        if self._verbose:
            first = (f"Performing {num_iter} iterations using a {self._object_type} object type, "
                     f"with the {reconstruction_method} algorithm, with normalization_min: "
                     f"{normalization_min}")
            if projection is None:
                warnings.warn(first + f" and step _size: {step_size}" + (
                    f", in batches of max {max_batch_size} measurements."
                    if max_batch_size is not None else "."), UserWarning)
            elif reconstruction_method in ("generalized-projections", "SUPERFLIP",
                                           "charge-flipping"):
                warnings.warn(first + f" and (a,b,c): {projection}.", UserWarning)
            else:
                warnings.warn(first + f" and α: {reconstruction_parameter}.", UserWarning)

        # batch order: identical RNG use to the reference loop (:835-857).
        # The affine position constraint draws from the same generator after
        # every iteration, so its orders are drawn one iteration at a time.
        affine = (not fix_positions) and bool(global_affine_transformation)
        shuffled = np.arange(N)
        if max_batch_size is not None:
            np.random.seed(seed_random)
            B = int(max_batch_size)
        else:
            B = N
        B = max(1, min(B, N))
        if projection is not None:
            # projection sets: one batch of all positions in scan order, no
            # shuffle (singleslice_ptychography.py:853-855)
            orders = np.arange(N, dtype=np.int32)[None, :]
        else:
            orders = np.empty((max(num_iter, 1), N), dtype=np.int32)
        for it in range(0 if (affine or projection is not None) else num_iter):
            np.random.shuffle(shuffled)
            orders[it] = shuffled

        dev = require_gpu()
        if not hasattr(self, "_object_dev"):
            self._object_dev = torch.from_numpy(np.ascontiguousarray(
                self._object, dtype=self._object_dtype())).to(dev)
        if not hasattr(self, "_probe_dev"):
            self._probe_dev = torch.from_numpy(np.ascontiguousarray(self._probe,
                                                                    dtype=np.complex64)).to(dev)
        use_mask = fix_potential_baseline and self._object_fov_mask_inverse.sum() > 0
        mask_dev = (torch.from_numpy(self._object_fov_mask_inverse.astype(np.uint8)).to(dev)
                    if use_mask else None)
        fmask_dev = None
        if detector_fourier_mask is not None:
            fmask = np.asarray(detector_fourier_mask, dtype=np.float32)
            if fmask.shape != tuple(self._amplitudes_shape):
                raise ValueError(f"detector_fourier_mask must have shape "
                                 f"{tuple(self._amplitudes_shape)}, got {fmask.shape}")
            # corner-centred like the amplitudes, stored in their kernel layout
            fmask_dev = K.to_kernel_layout(
                torch.from_numpy(np.ascontiguousarray(fmask[None])).to(dev))[0].contiguous()
        potential = self._object_type == "potential"
        if potential:
            # the potential forward/adjoint and constraints (ptychographic_methods.py:
            # 1522-1523, 1868-1880; ptychographic_constraints.py:60-118) run in
            # gradient descent with fixed positions
            for what, used in (("projection-set methods", projection is not None),
                               ("position correction", not fix_positions),
                               ("detector_fourier_mask", detector_fourier_mask is not None),
                               ("q_lowpass/q_highpass", butterworth_filter and (
                                   q_lowpass is not None or q_highpass is not None))):
                if used:
                    raise NotImplementedError(f"{what} with object_type='potential' is not "
                                              "supported on this path")
        problem = K.make_problem(self._object_dev, self._probe_dev, self._positions_dev,
                                 self._amplitudes_dev, mask_dev, self._mean_diffraction_intensity,
                                 fmask_dev)
        per_call = 1 if (store_iterations or progress_bar or affine) else max(num_iter, 1)
        params = PtyGdParams(per_call, B, int(bool(fix_probe)), int(bool(fix_probe_com)),
                             # singleslice_ptychography.py:976
                             int(bool(pure_phase_object) and not potential), float(step_size),
                             float(normalization_min), float(shrinkage_rad))
        stage, pipeline = self._stage_for(B)
        params.stage = int(stage)
        params.pipeline = int(bool(pipeline))
        params.object_positivity = int(bool(object_positivity))
        aperture_dev = None
        if fix_probe_aperture and not fix_probe:
            # |fft2| of the initial probe (mode 0 with several modes), in the
            # kernel layout of the transforms
            ap = getattr(self, "_probe_initial_aperture", None)
            if ap is None:
                raise ValueError("fix_probe_aperture needs the initial probe aperture, which "
                                 "preprocess(store_initial_arrays=True) records")
            ap = np.asarray(ap, dtype=np.float32)
            ap = ap[0] if ap.ndim == 3 else ap
            aperture_dev = K.to_kernel_layout(
                torch.from_numpy(np.ascontiguousarray(ap[None])).to(dev))[0].contiguous()
            params.fix_probe_aperture = 1
            params.probe_aperture = aperture_dev.data_ptr()
        if constrain_probe_amplitude and not fix_probe:
            params.constrain_probe_amplitude = 1
            params.probe_amplitude_radius = float(constrain_probe_amplitude_relative_radius)
            params.probe_amplitude_width = float(constrain_probe_amplitude_relative_width)
        gauss_dev = None
        if gaussian_filter and gaussian_filter_sigma is not None:
            # ptychographic_constraints.py:142: sigma from Angstrom to pixels
            sigma_px = float(gaussian_filter_sigma) / float(self.sampling[0])
            if sigma_px > 1e-15:
                w, r = K.gaussian_kernel1d(sigma_px)
                gauss_dev = torch.from_numpy(w).to(dev)
                params.gaussian_weights = gauss_dev.data_ptr()
                params.gaussian_radius = r
        if butterworth_filter and (q_lowpass is not None or q_highpass is not None):
            # singleslice_ptychography.py:961-965; a cut-off of None or 0 drops
            # its factor (ptychographic_constraints.py:188-191)
            params.butterworth = 1
            params.q_lowpass = float(q_lowpass or 0.0)
            params.q_highpass = float(q_highpass or 0.0)
            params.butterworth_order = float(butterworth_order)
            params.sampling[0] = float(self.sampling[0])
            params.sampling[1] = float(self.sampling[1])
        pos_params = singular = proj_params = None
        if projection is not None:
            if getattr(self, "_exit_waves", None) is None:
                # the reference starts from exit waves = overlap (None here)
                self._exit_waves = torch.empty((N,) + tuple(self._amplitudes_shape),
                                               dtype=torch.complex64, device=dev)
                self._exit_valid = False
            proj_params = K.projection_params(self._exit_waves, *projection,
                                              exit_valid=self._exit_valid)
            nbytes = K.projection_workspace_bytes(problem, params)
        if not fix_positions:
            # position correction (singleslice_ptychography.py:913-926) and the
            # centre-of-mass constraint, which the reference applies when its
            # own flag arrives False (fix_positions_com and not fix_positions
            # at :956, tested as `if not fix_positions_com` in
            # ptychographic_constraints.py:1367)
            norm = np.sqrt(self.sampling[0] ** 2 + self.sampling[1] ** 2)
            singular = torch.zeros(1, dtype=torch.int32, device=dev)
            pos_params = K.position_params(
                self._positions_dev, self._positions_initial_dev, positions_step_size,
                None if max_position_update_distance is None
                else max_position_update_distance / norm,
                None if max_position_total_distance is None
                else max_position_total_distance / norm,
                center_of_mass=not fix_positions_com,
                initial_com=self._positions_px_initial.mean(0), singular=singular)
            nbytes = (K.gd_positions_workspace_bytes(problem, params) if projection is None
                      else K.projection_positions_workspace_bytes(problem, params))
        elif projection is None:
            nbytes = K.gd_workspace_bytes(problem, params)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev) if _allocate else None
        orders_dev = torch.from_numpy(orders).to(dev)
        errors_dev = torch.zeros(max(num_iter, 1), dtype=torch.float64, device=dev)
        self._last_run = dict(problem=problem, params=params, workspace=ws, orders=orders_dev,
                              errors=errors_dev, mask=mask_dev, fourier_mask=fmask_dev,
                              probe_aperture=aperture_dev, gaussian_weights=gauss_dev,
                              num_iter=num_iter,
                              per_call=per_call, store_iterations=store_iterations,
                              progress_bar=progress_bar, position_params=pos_params,
                              singular=singular, affine=affine, shuffled=shuffled,
                              projection_params=proj_params)
        return self._last_run

    def _execute_run(self, run):
        problem, params, ws = run["problem"], run["params"], run["workspace"]
        orders_dev, errors_dev = run["orders"], run["errors"]
        num_iter, per_call = run["num_iter"], run["per_call"]
        it = 0
        bar = None
        if run["progress_bar"] and num_iter > 0:
            try:
                from tqdm import tqdm
                bar = tqdm(total=num_iter, desc="Reconstructing object and probe", unit=" iter")
            except ImportError:
                bar = None
        pos_params = run["position_params"]
        while it < num_iter:
            n = min(per_call, num_iter - it)
            params.num_iter = n
            if run["affine"] and run["projection_params"] is None:
                # gradient descent shuffles before every iteration; the
                # projection sets keep the scan order
                np.random.shuffle(run["shuffled"])
                orders_dev[it].copy_(torch.from_numpy(run["shuffled"].astype(np.int32)))
            if run["projection_params"] is not None:
                if pos_params is None:
                    K.projection_run(problem, params, run["projection_params"], orders_dev[0],
                                     errors_dev[it:], ws)
                else:
                    K.projection_run_positions(problem, params, run["projection_params"],
                                               pos_params, orders_dev[0], errors_dev[it:], ws)
                run["projection_params"].exit_valid = 1
                self._exit_valid = True
            elif pos_params is None:
                K.gd_run(problem, params, orders_dev[it:], errors_dev[it:], ws)
            else:
                K.gd_run_positions(problem, params, pos_params, orders_dev[it:], errors_dev[it:],
                                   ws)
            if run["affine"]:
                self._affine_positions_constraint()
            it += n
            if run["store_iterations"]:
                self.object_iterations.append(self._object_dev.cpu().numpy().copy())
                self.probe_iterations.append(
                    np.fft.fftshift(self._probe_dev.cpu().numpy(), axes=(-2, -1)))
            if bar is not None:
                bar.update(n)
        if bar is not None:
            bar.close()

    def _affine_positions_constraint(self):
        """ptychographic_constraints.py:1313-1353: positions replaced by the
        RANSAC affine map of the initial scan (host; N x 2 floats)."""
        from .positions import estimate_global_transformation_ransac
        initial = self._positions_px_initial
        current = self._positions_dev.cpu().numpy()
        origin = initial.mean(0)
        tf, _ = estimate_global_transformation_ransac(initial, current, origin=origin,
                                                      translation_allowed=True,
                                                      min_sample=initial.shape[0] // 10)
        self._tf = tf
        new = np.asarray(tf(initial, origin=origin), dtype=np.float32)
        self._positions_dev.copy_(torch.from_numpy(np.ascontiguousarray(new)))

    def _finish_run(self, run):
        num_iter = run["num_iter"]
        if run.get("position_params") is not None:
            self._positions_px = self._positions_dev.cpu().numpy()
            if int(run["singular"].item()):
                raise np.linalg.LinAlgError(
                    "Singular matrix in position correction "
                    f"({int(run['singular'].item())} position(s) got no update)")
        errors = run["errors"][:num_iter].cpu().numpy()
        self.error_iterations.extend(float(e) for e in errors)
        self._object = self._object_dev.cpu().numpy()
        self._probe = self._probe_dev.cpu().numpy()
        self.object = self._object.copy()
        self.probe = self.probe_centered
        if num_iter > 0:
            self.error = float(errors[-1])
        elif not hasattr(self, "error"):
            self.error = 0.0
        if run.get("projection_params") is None:
            self._exit_waves = None  # singleslice_ptychography.py:990-992
        return self
