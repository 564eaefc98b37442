"""Seeded synthetic 4D-STEM ptychography workloads.

The forward model is the single-slice one the reconstruction inverts: the
probe, sub-pixel shifted in Fourier space, multiplies a corner-centred
periodic patch of the complex object, and the detector records
``|fft2(.)|**2`` with Poisson noise, stored with the pattern centre at pixel
(Q/2, Q/2) like a raw datacube.  Simulation is done in complex128 NumPy on
the host (it is data generation, not the measured path).

Workloads (sizes in the docstrings of :func:`small_workload` and
:func:`production_workload`) follow the two plain GD configurations: a
32x32-scan / 64x64-pixel CPU-sized case and the 128x128-scan /
128x128-pixel production case.
"""

from __future__ import annotations

import numpy as np
import scipy.ndimage

from .datacube import Calibration, DataCube
from .preprocess import (
    center_positions,
    real_space_sampling,
    scan_positions_in_pixels,
)
from .probe import ComplexProbe, electron_wavelength_angstrom


def random_field_phase_object(shape, sigma_px, max_phase, rng):
    """|O| = 1, phase = Gaussian-smoothed white noise rescaled to [0, max_phase]."""
    field = scipy.ndimage.gaussian_filter(rng.standard_normal(shape), sigma_px)
    field = (field - field.min()) / (field.max() - field.min()) * max_phase
    return np.exp(1j * field).astype(np.complex64)


def gaussian_atoms_phase_object(shape, sampling, lattice_ang, sigma_ang, peak_rad,
                                jitter_frac, rng, vacancy_frac=0.0):
    """|O| = 1, phase = sum of Gaussian 'atoms' on a jittered square lattice."""
    ny, nx = shape
    phase = np.zeros(shape, dtype=np.float64)
    step = lattice_ang / sampling
    sig = sigma_ang / sampling
    half = int(np.ceil(4 * sig))
    xs = np.arange(step / 2, shape[0], step)
    ys = np.arange(step / 2, shape[1], step)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    centres = np.stack([gx.ravel(), gy.ravel()], axis=1)
    centres += rng.uniform(-jitter_frac, jitter_frac, centres.shape) * step
    if vacancy_frac > 0:
        centres = centres[rng.uniform(size=len(centres)) >= vacancy_frac]
    for cx, cy in centres:
        i0, j0 = int(np.floor(cx)), int(np.floor(cy))
        ii = np.arange(i0 - half, i0 + half + 2)
        jj = np.arange(j0 - half, j0 + half + 2)
        g = np.exp(-((ii[:, None] - cx) ** 2 + (jj[None, :] - cy) ** 2) / (2 * sig * sig))
        phase[np.ix_(ii % shape[0], jj % shape[1])] += peak_rad * g
    return np.exp(1j * phase).astype(np.complex64), phase


def scaled_probe(energy, S, sampling, semiangle_cutoff, rolloff, dose_per_pattern, **aberrations):
    """Unit-norm probe scaled so that sum |fft2 P|^2 = dose_per_pattern."""
    probe = ComplexProbe(energy=energy, gpts=(S, S), sampling=sampling,
                         semiangle_cutoff=semiangle_cutoff, rolloff=rolloff,
                         parameters=aberrations).build()._array.astype(np.complex128)
    probe *= np.sqrt(dose_per_pattern / np.sum(np.abs(np.fft.fft2(probe)) ** 2))
    return probe


def simulate_intensities(obj, probe, positions_px, chunk=512):
    """Noise-free corner-centred intensities (N, S, S) float64."""
    S = probe.shape[0]
    N = positions_px.shape[0]
    P_hat = np.fft.fft2(probe)
    k = np.fft.fftfreq(S, 1.0)
    idx = np.fft.fftfreq(S, 1.0 / S).astype(int)
    out = np.empty((N, S, S), dtype=np.float64)
    for s in range(0, N, chunk):
        pos = np.asarray(positions_px[s:s + chunk], dtype=np.float64)
        r = np.round(pos)
        frac = pos - r
        ramp = (np.exp(-2j * np.pi * k[None, :, None] * frac[:, 0, None, None])
                * np.exp(-2j * np.pi * k[None, None, :] * frac[:, 1, None, None]))
        shifted = np.fft.ifft2(P_hat[None] * ramp)
        x0 = r[:, 0].astype(int)
        y0 = r[:, 1].astype(int)
        rows = (x0[:, None, None] + idx[None, :, None]) % obj.shape[0]
        cols = (y0[:, None, None] + idx[None, None, :]) % obj.shape[1]
        out[s:s + chunk] = np.abs(np.fft.fft2(shifted * obj[rows, cols])) ** 2
    return out


def simulate_intensities_gpu(obj, probe, positions_px, chunk=8192):
    """Noise-free intensities (N, S, S) float32 from the HIP forward-model
    kernel (float32 arithmetic); returns a torch tensor on the GPU."""
    import torch

    from . import _lib, kernels

    dev = _lib.require_gpu()
    S = probe.shape[0]
    o = torch.from_numpy(np.ascontiguousarray(obj, dtype=np.complex64)).to(dev)
    phat = kernels.probe_spectrum(torch.from_numpy(np.ascontiguousarray(probe, dtype=np.complex64)).to(dev))
    pos = torch.from_numpy(np.ascontiguousarray(positions_px, dtype=np.float32)).to(dev)
    N = pos.shape[0]
    out = torch.empty((N, S, S), dtype=torch.float32, device=dev)
    for s in range(0, N, chunk):
        n = min(chunk, N - s)
        _lib.call("pty_simulate_intensities", _lib.ptr(phat), _lib.ptr(o), o.shape[0], o.shape[1],
                  _lib.ptr(pos[s:s + n]), n, S, _lib.ptr(out[s:s + n]), _lib.stream_handle())
    return out


def poisson_datacube_gpu(intensities, grid_shape, scan_step_ang, dk_inv_ang, seed):
    """Poisson-sample GPU intensities, centre the patterns, return a DataCube."""
    import torch

    g = torch.Generator(device=intensities.device)
    g.manual_seed(int(seed))
    counts = torch.poisson(intensities.clamp_min(0), generator=g)
    data = torch.fft.fftshift(counts, dim=(-2, -1)).cpu().numpy()
    S = data.shape[-1]
    cal = Calibration(R_pixel_size=scan_step_ang, R_pixel_units="A",
                      Q_pixel_size=dk_inv_ang, Q_pixel_units="A^-1")
    return DataCube(data.reshape(tuple(grid_shape) + (S, S)), cal)


def to_datacube(intensities, grid_shape, scan_step_ang, dk_inv_ang, rng=None, poisson=True):
    """Poisson-sample, centre the patterns and wrap in a calibrated DataCube."""
    data = rng.poisson(intensities).astype(np.float32) if poisson else intensities.astype(np.float32)
    data = np.fft.fftshift(data, axes=(-2, -1))
    S = data.shape[-1]
    data = data.reshape(grid_shape + (S, S))
    cal = Calibration(R_pixel_size=scan_step_ang, R_pixel_units="A",
                      Q_pixel_size=dk_inv_ang, Q_pixel_units="A^-1")
    return DataCube(data, cal)


def small_workload(seed=0, scan=(32, 32), S=64, object_px=256, poisson=True):
    """32x32 scan, 1.0 A step, +/-0.05 A jitter; 64x64 px at 0.2 A; 256^2 object
    (phase 0-0.3 rad random field, sigma 2 px); 300 keV, 20 mrad, defocus=200;
    1e5 counts per pattern; 8 iterations, full batch, fixed probe."""
    rng = np.random.default_rng(seed)
    energy, semiangle, rolloff, defocus = 300e3, 20.0, 2.0, 200.0
    dk = 1.0 / (S * 0.2)
    lam = electron_wavelength_angstrom(energy)
    sampling = real_space_sampling(energy, (dk * lam * 1e3,) * 2, (S, S))
    obj = random_field_phase_object((object_px, object_px), 2.0, 0.3, rng)
    step_px = 1.0 / sampling[0]
    ext = (np.array(scan) - 1) * step_px
    gx = object_px / 2 - ext[0] / 2 + np.arange(scan[0]) * step_px
    gy = object_px / 2 - ext[1] / 2 + np.arange(scan[1]) * step_px
    cx, cy = np.meshgrid(gx, gy, indexing="ij")
    centres = np.stack([cx.ravel(), cy.ravel()], axis=1)
    centres += rng.uniform(-0.05, 0.05, centres.shape) / sampling[0]
    positions_ang = (centres - S / 2) * sampling[0]
    pos_px, _ = scan_positions_in_pixels(positions_ang, scan, (1.0, 1.0), 0.0, False,
                                         sampling, (S, S))
    pos_px = pos_px.astype(np.float32)
    dose = 1e5
    probe = scaled_probe(energy, S, sampling, semiangle, rolloff, dose, defocus=defocus)
    I = simulate_intensities(obj.astype(np.complex128), probe, pos_px)
    dc = to_datacube(I, scan, 1.0, dk, rng, poisson)
    return {
        "datacube": dc, "object_true": obj, "positions_ang": positions_ang,
        "positions_px": pos_px, "probe_true": probe.astype(np.complex64),
        "init": dict(energy=energy, semiangle_cutoff=semiangle, rolloff=rolloff,
                     defocus=defocus, initial_object_guess=np.ones((object_px,) * 2, np.complex64),
                     initial_scan_positions=positions_ang, verbose=False),
        "preprocess": dict(force_com_rotation=0, force_com_transpose=False,
                           force_com_shifts=(S / 2, S / 2), center_positions_in_fov=False,
                           plot_center_of_mass=False, plot_rotation=False,
                           plot_probe_overlaps=False),
        "reconstruct": dict(num_iter=8, max_batch_size=None, fix_probe=True, step_size=0.5,
                            normalization_min=1, seed_random=seed, progress_bar=False),
    }


def production_workload(seed=0, scan=(128, 128), S=128, object_px=720, poisson=True,
                        lattice_ang=2.5, num_iter=50, batch=1024):
    """128x128 scan at 0.4 A; 128x128 px at 0.1 A; 720^2 object of Gaussian atoms
    (sigma 0.3 A, peak 0.15 rad) on a jittered 2.5 A lattice; 300 keV,
    20 mrad, defocus=200; 1e6 e/A^2 (1.6e5 counts per pattern); 50 iterations,
    mini-batch 1024 in seeded order, probe update enabled."""
    rng = np.random.default_rng(seed)
    energy, semiangle, rolloff, defocus = 300e3, 20.0, 2.0, 200.0
    dk = 1.0 / (S * 0.1)
    lam = electron_wavelength_angstrom(energy)
    sampling = real_space_sampling(energy, (dk * lam * 1e3,) * 2, (S, S))
    obj, _ = gaussian_atoms_phase_object((object_px, object_px), sampling[0], lattice_ang,
                                         0.3, 0.15, 0.1, rng)
    step = 0.4
    pos_px, _ = scan_positions_in_pixels(None, scan, (step, step), 0.0, False, sampling, (S, S))
    pos_px = center_positions(pos_px.astype(np.float32), (object_px, object_px))
    dose = 1e6 * step * step
    probe = scaled_probe(energy, S, sampling, semiangle, rolloff, dose, defocus=defocus)
    I = simulate_intensities(obj.astype(np.complex128), probe, pos_px)
    dc = to_datacube(I, scan, step, dk, rng, poisson)
    return {
        "datacube": dc, "object_true": obj, "positions_px": pos_px,
        "probe_true": probe.astype(np.complex64),
        "init": dict(energy=energy, semiangle_cutoff=semiangle, rolloff=rolloff,
                     defocus=defocus, initial_object_guess=np.ones((object_px,) * 2, np.complex64),
                     verbose=False),
        "preprocess": dict(force_com_rotation=0, force_com_transpose=False,
                           force_com_shifts=(S / 2, S / 2), plot_center_of_mass=False,
                           plot_rotation=False, plot_probe_overlaps=False),
        "reconstruct": dict(num_iter=num_iter, max_batch_size=batch, fix_probe=False,
                            step_size=0.5, normalization_min=1, seed_random=seed,
                            progress_bar=False),
    }


def many_workloads(count=64, seed=0, scan_range=(32, 96), frames=(64, 128),
                   defocus_range=(100.0, 400.0), num_iter=30, batch=None, on_gpu=True):
    """Independent seeded problems for batched reconstruction: square scans
    with edge uniform in scan_range, frames alternating between the given
    sizes (half each), defocus uniform in defocus_range, Gaussian-atom
    objects sized by the reconstruction's own default rule.  64 px frames use
    0.2 A sampling and 1.0 A steps, 128 px frames 0.1 A and 0.4 A."""
    from .preprocess import initial_object_shape

    rng = np.random.default_rng(seed)
    energy, semiangle, rolloff = 300e3, 20.0, 2.0
    lam = electron_wavelength_angstrom(energy)
    out = []
    for i in range(count):
        S = frames[i % len(frames)]
        n = int(rng.integers(scan_range[0], scan_range[1] + 1))
        defocus = float(rng.uniform(*defocus_range))
        px = 0.2 if S == 64 else 0.1
        step = 1.0 if S == 64 else 0.4
        dk = 1.0 / (S * px)
        sampling = real_space_sampling(energy, (dk * lam * 1e3,) * 2, (S, S))
        pos, pad = scan_positions_in_pixels(None, (n, n), (step, step), 0.0, False, sampling, (S, S))
        shape = initial_object_shape(pos, pad, np.array([S, S]))
        pos = center_positions(pos.astype(np.float32), shape)
        obj, _ = gaussian_atoms_phase_object(shape, sampling[0], 2.5, 0.3, 0.15, 0.1, rng)
        dose = 1e5 * step * step
        probe = scaled_probe(energy, S, sampling, semiangle, rolloff, dose, defocus=defocus)
        if on_gpu:
            I = simulate_intensities_gpu(obj, probe, pos)
            dc = poisson_datacube_gpu(I, (n, n), step, dk, seed * 1000 + i)
            del I
        else:
            I = simulate_intensities(obj.astype(np.complex128), probe, pos)
            dc = to_datacube(I, (n, n), step, dk, rng, True)
        out.append({
            "datacube": dc, "object_true": obj,
            "init": dict(energy=energy, semiangle_cutoff=semiangle, rolloff=rolloff,
                         defocus=defocus, verbose=False),
            "preprocess": dict(force_com_rotation=0, force_com_transpose=False,
                               force_com_shifts=(S / 2, S / 2), plot_center_of_mass=False,
                               plot_rotation=False, plot_probe_overlaps=False),
            "reconstruct": dict(num_iter=num_iter, max_batch_size=batch, fix_probe=False,
                                step_size=0.5, normalization_min=1, seed_random=seed * 1000 + i,
                                progress_bar=False),
        })
    return out


def object_nmse(recon, truth, mask=None):
    """min_c ||c*recon - truth||^2 / ||truth||^2 over ``mask`` (complex scalar c
    absorbs the global phase/scale ambiguity)."""
    r = np.asarray(recon, dtype=np.complex128)
    t = np.asarray(truth, dtype=np.complex128)
    if mask is not None:
        r, t = r[mask], t[mask]
    c = np.vdot(r, t) / np.vdot(r, r)
    return float(np.sum(np.abs(c * r - t) ** 2) / np.sum(np.abs(t) ** 2))
