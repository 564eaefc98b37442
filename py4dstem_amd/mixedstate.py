"""Drop-in ``MixedstatePtychography``: single-slice ptychography with several
incoherent probe modes, gradient-descent path on MI355X.

Same constructor / ``preprocess()`` / ``reconstruct()`` keywords, chaining and
errors as py4DSTEM's class (py4DSTEM/process/phase/mixedstate_ptychography.py:
45-1027).  The probe is [K][S][S]; every iteration runs in libpty_hip.so
(``pty_mixed_run``: per position and mode the shifted probe, exit wave and far
field, the common amplitude factor sqrt of the summed mode intensities, and
the per-mode adjoint terms; object and probe steps with the reference's
float32 mode sums; recentring of every mode by mode 0's centre of mass).  The
mode orthogonalisation of ptychographic_constraints.py:1180-1221 (an eigen
decomposition of the K x K Gram matrix) runs on the host between the
per-iteration native calls, with the reference's NumPy arithmetic.
"""

from __future__ import annotations

import numpy as np
import torch

from . import kernels as K
from .ptychography import SingleslicePtychography


def mixed_probe_modes(probe, num_probes):
    """ptychographic_methods.py:1416-1462: mode 0 is the single-mode probe,
    mode i the previous mode times sub-pixel phase ramps whose offsets are
    drawn from the global NumPy generator (two draws per mode)."""
    sx, sy = probe.shape
    probes = np.zeros((num_probes, sx, sy), dtype=np.complex64)
    probes[0] = probe
    for i_probe in range(1, num_probes):
        shift_x = np.exp(-2j * np.pi * (np.random.rand() - 0.5) * np.fft.fftfreq(sx))
        shift_y = np.exp(-2j * np.pi * (np.random.rand() - 0.5) * np.fft.fftfreq(sy))
        probes[i_probe] = probes[i_probe - 1] * shift_x[:, None] * shift_y[None]
    return probes


def orthogonalize_probe_modes(probe):
    """ptychographic_constraints.py:1180-1221: the modes rotated onto the
    eigenvectors of their Gram matrix (upper triangle, the probe's complex64
    sums), then ordered by real-space intensity, largest first."""
    n = probe.shape[0]
    gram = np.empty((n, n), dtype=probe.dtype)
    for i in range(n):
        for j in range(i, n):
            gram[i, j] = np.sum(probe[i].conj() * probe[j])
    _, evecs = np.linalg.eigh(gram, UPLO="U")
    probe = np.tensordot(evecs.T, probe, axes=1)
    intensities = np.sum(np.abs(probe) ** 2, axis=(-2, -1))
    return probe[np.argsort(intensities, axis=None)[::-1]]


class MixedstatePtychography(SingleslicePtychography):
    def __init__(self, energy, datacube=None, num_probes=None, semiangle_cutoff=None,
                 semiangle_cutoff_pixels=None, rolloff=2.0, vacuum_probe_intensity=None,
                 polar_parameters=None, object_padding_px=None, initial_object_guess=None,
                 initial_probe_guess=None, initial_scan_positions=None, object_fov_ang=None,
                 positions_offset_ang=None, object_type="complex", positions_mask=None,
                 verbose=True, device="cpu", storage=None, clear_fft_cache=True,
                 name="mixed-state_ptychographic_reconstruction", **kwargs):
        """mixedstate_ptychography.py:119-207."""
        if initial_probe_guess is None or hasattr(initial_probe_guess, "build"):
            if num_probes is None:
                raise ValueError("If initial_probe_guess is None, or a ComplexProbe object, "
                                 "num_probes must be specified.")
        else:
            if len(initial_probe_guess.shape) != 3:
                raise ValueError("Specified initial_probe_guess must have dimensions (N,Sx,Sy).")
            num_probes = initial_probe_guess.shape[0]
        super().__init__(energy, datacube=datacube, semiangle_cutoff=semiangle_cutoff,
                         semiangle_cutoff_pixels=semiangle_cutoff_pixels, rolloff=rolloff,
                         vacuum_probe_intensity=vacuum_probe_intensity,
                         polar_parameters=polar_parameters,
                         initial_object_guess=initial_object_guess,
                         initial_probe_guess=initial_probe_guess,
                         initial_scan_positions=initial_scan_positions,
                         object_fov_ang=object_fov_ang, positions_offset_ang=positions_offset_ang,
                         object_padding_px=object_padding_px, object_type=object_type,
                         positions_mask=positions_mask, verbose=verbose, device=device,
                         storage=storage, clear_fft_cache=clear_fft_cache, name=name, **kwargs)
        self._num_probes = int(num_probes)

    # -- probe modes ----------------------------------------------------------------
    def _initialize_probe(self, initial_probe):
        """ptychographic_methods.py:1416-1462."""
        if initial_probe is None or hasattr(initial_probe, "build"):
            single = super()._initialize_probe(initial_probe)
            return mixed_probe_modes(single, self._num_probes)
        return np.array(initial_probe, dtype=np.complex64, copy=True)

    def _probe_overlap_field(self, max_batch_size=None):
        """The FOV overlap of the first mode (mixedstate_ptychography.py:548-563)."""
        modes = self._probe
        self._probe = np.ascontiguousarray(modes[0])
        try:
            return super()._probe_overlap_field(max_batch_size)
        finally:
            self._probe = modes

    def initialize_object_learned(self, *args, **kwargs):
        raise NotImplementedError("the learned initialiser is trained for a single probe mode")

    # -- reconstruct --------------------------------------------------------------------
    def reconstruct(self, num_iter=8, reconstruction_method="gradient-descent",
                    reconstruction_parameter=1.0, reconstruction_parameter_a=None,
                    reconstruction_parameter_b=None, reconstruction_parameter_c=None,
                    max_batch_size=None, seed_random=None, step_size=0.5, normalization_min=1,
                    positions_step_size=0.9, pure_phase_object=False, fix_probe_com=True,
                    orthogonalize_probe=True, fix_probe=False, fix_probe_aperture=False,
                    constrain_probe_amplitude=False,
                    constrain_probe_amplitude_relative_radius=0.5,
                    constrain_probe_amplitude_relative_width=0.05,
                    constrain_probe_fourier_amplitude=False,
                    constrain_probe_fourier_amplitude_max_width_pixels=3.0,
                    constrain_probe_fourier_amplitude_constant_intensity=False, fix_positions=True,
                    fix_positions_com=True, global_affine_transformation=False,
                    max_position_update_distance=None, max_position_total_distance=None,
                    gaussian_filter_sigma=None, gaussian_filter=True, fit_probe_aberrations=False,
                    fit_probe_aberrations_max_angular_order=4,
                    fit_probe_aberrations_max_radial_order=4,
                    fit_probe_aberrations_remove_initial=False,
                    fit_probe_aberrations_using_scikit_image=True,
                    num_probes_fit_aberrations=np.inf, butterworth_filter=True, q_lowpass=None,
                    q_highpass=None, butterworth_order=2, tv_denoise=True, tv_denoise_weight=None,
                    tv_denoise_inner_iter=40, object_positivity=True, shrinkage_rad=0.0,
                    fix_potential_baseline=True, detector_fourier_mask=None,
                    store_iterations=False, progress_bar=True, reset=None, device=None,
                    clear_fft_cache=None, object_type=None):
        """mixedstate_ptychography.py:631-1027 (gradient descent).  Returns self."""
        if reconstruction_method not in ("GD", "gradient-descent"):
            from .ptychography import _method_parameters
            _method_parameters(reconstruction_method, reconstruction_parameter,
                               reconstruction_parameter_a, reconstruction_parameter_b,
                               reconstruction_parameter_c)  # the reference's ValueErrors first
            raise NotImplementedError("mixed-state projection-set methods are not supported on "
                                      "this path (gradient descent only)")
        if self._object_type == "potential":
            raise NotImplementedError("potential objects with several probe modes are not "
                                      "supported on this path")
        if int(self._region_of_interest_shape[0]) not in K.LDS_FRAMES:
            raise NotImplementedError("several probe modes take frames of 32, 64 or 128 px on "
                                      "this path")
        run = self._prepare_run(
            num_iter=num_iter, reconstruction_method=reconstruction_method,
            max_batch_size=max_batch_size, seed_random=seed_random, step_size=step_size,
            normalization_min=normalization_min, pure_phase_object=pure_phase_object,
            fix_probe_com=fix_probe_com, fix_probe=fix_probe,
            fix_probe_aperture=fix_probe_aperture, positions_step_size=positions_step_size,
            fix_positions=fix_positions, fix_positions_com=fix_positions_com,
            global_affine_transformation=global_affine_transformation,
            max_position_update_distance=max_position_update_distance,
            max_position_total_distance=max_position_total_distance,
            constrain_probe_amplitude=constrain_probe_amplitude,
            constrain_probe_amplitude_relative_radius=constrain_probe_amplitude_relative_radius,
            constrain_probe_amplitude_relative_width=constrain_probe_amplitude_relative_width,
            constrain_probe_fourier_amplitude=constrain_probe_fourier_amplitude,
            gaussian_filter_sigma=gaussian_filter_sigma, gaussian_filter=gaussian_filter,
            fit_probe_aberrations=fit_probe_aberrations, butterworth_filter=butterworth_filter,
            q_lowpass=q_lowpass, q_highpass=q_highpass, tv_denoise=tv_denoise,
            tv_denoise_weight=tv_denoise_weight, shrinkage_rad=shrinkage_rad,
            fix_potential_baseline=fix_potential_baseline,
            detector_fourier_mask=detector_fourier_mask, store_iterations=store_iterations,
            progress_bar=progress_bar, reset=reset, device=device,
            clear_fft_cache=clear_fft_cache, object_type=object_type, _allocate=False)
        dev = run["orders"].device
        nbytes = (K.mixed_workspace_bytes if run["position_params"] is None
                  else K.mixed_positions_workspace_bytes)(run["problem"], run["params"],
                                                          self._num_probes)
        run["workspace"] = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        run["orthogonalize"] = bool(orthogonalize_probe)
        if orthogonalize_probe:
            run["per_call"] = 1  # the host step follows every iteration
        self._execute_run(run)
        return self._finish_run(run)

    def _execute_run(self, run):
        problem, params, ws = run["problem"], run["params"], run["workspace"]
        orders_dev, errors_dev = run["orders"], run["errors"]
        num_iter, per_call = run["num_iter"], run["per_call"]
        bar = None
        if run["progress_bar"] and num_iter > 0:
            try:
                from tqdm import tqdm
                bar = tqdm(total=num_iter, desc="Reconstructing object and probe", unit=" iter")
            except ImportError:
                bar = None
        it = 0
        while it < num_iter:
            n = min(per_call, num_iter - it)
            params.num_iter = n
            if run["affine"]:  # one iteration per call: orders drawn between the RANSAC fits
                np.random.shuffle(run["shuffled"])
                orders_dev[it].copy_(torch.from_numpy(run["shuffled"].astype(np.int32)))
            K.mixed_run(problem, params, self._num_probes, orders_dev[it:], errors_dev[it:], ws,
                        run["position_params"])
            if run["orthogonalize"]:
                modes = orthogonalize_probe_modes(self._probe_dev.cpu().numpy())
                self._probe_dev.copy_(torch.from_numpy(np.ascontiguousarray(modes)))
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
