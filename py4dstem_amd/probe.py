"""Probe optics: electron wavelength, aberration function and aperture.

Host-side (NumPy) set-up code for the reconstruction; it runs once per
``preprocess()`` and is not on the GPU hot path.

References (py4DSTEM 0.14.18):
  * wavelength: py4DSTEM/process/utils/utils.py:90-97
  * polar symbols / aliases and the defocus sign flip:
    process/phase/utils.py:29-43, 141-160; phase_base_class.py:1788-1809
  * aperture with cosine rolloff: process/phase/utils.py:162-191
  * aberration function chi: process/phase/utils.py:290-368
  * corner-centred probe build: process/phase/utils.py:401-426
"""

from __future__ import annotations

import math

import numpy as np

POLAR_SYMBOLS = (
    "C10", "C12", "phi12",
    "C21", "phi21", "C23", "phi23",
    "C30", "C32", "phi32", "C34", "phi34",
    "C41", "phi41", "C43", "phi43", "C45", "phi45",
    "C50", "C52", "phi52", "C54", "phi54", "C56", "phi56",
)
POLAR_ALIASES = {
    "defocus": "C10", "astigmatism": "C12", "astigmatism_angle": "phi12",
    "coma": "C21", "coma_angle": "phi21", "Cs": "C30", "C5": "C50",
}

# Radial order n of alpha^(n+1)/(n+1), with the (magnitude, angle, m) terms of
# cos(m (phi - angle)) that it collects, and the symbols whose non-zero value
# switches the order on.  The fourth-order guard lists phi41 twice and never
# phi45, exactly as the reference does (process/phase/utils.py:335-341).
_CHI_ORDERS = (
    (1, (("C10", None, 0), ("C12", "phi12", 2)), ("C10", "C12", "phi12")),
    (2, (("C21", "phi21", 1), ("C23", "phi23", 3)), ("C21", "phi21", "C23", "phi23")),
    (3, (("C30", None, 0), ("C32", "phi32", 2), ("C34", "phi34", 4)),
     ("C30", "C32", "phi32", "C34", "phi34")),
    (4, (("C41", "phi41", 1), ("C43", "phi43", 3), ("C45", "phi45", 5)),
     ("C41", "phi41", "C43", "phi43", "C45", "phi41")),
    (5, (("C50", None, 0), ("C52", "phi52", 2), ("C54", "phi54", 4), ("C56", "phi56", 6)),
     ("C50", "C52", "phi52", "C54", "phi54", "C56", "phi56")),
)


def electron_wavelength_angstrom(E_eV: float) -> float:
    """Relativistically corrected de Broglie wavelength in Angstrom."""
    m = 9.109383 * 10**-31
    e = 1.602177 * 10**-19
    c = 299792458
    h = 6.62607 * 10**-34
    return h / math.sqrt(2 * m * e * E_eV) / math.sqrt(1 + e * E_eV / 2 / m / c**2) * 10**10


def polar_parameters_from(parameters=None, **kwargs) -> dict:
    """All 25 polar coefficients, with aliases resolved (defocus -> -C10)."""
    out = dict.fromkeys(POLAR_SYMBOLS, 0.0)
    merged = dict(parameters or {})
    merged.update(kwargs)
    for key, value in merged.items():
        if key in out:
            out[key] = value
        elif key == "defocus":
            out["C10"] = -value
        elif key in POLAR_ALIASES:
            out[POLAR_ALIASES[key]] = value
        else:
            raise ValueError("{} not a recognized parameter".format(key))
    return out


def _fftfreq32(n, d):
    return np.fft.fftfreq(int(n), d).astype(np.float32)


def scattering_angles(gpts, sampling, wavelength):
    """(alpha, phi) on the corner-centred float32 frequency grid."""
    kx = _fftfreq32(gpts[0], sampling[0]) * wavelength
    ky = _fftfreq32(gpts[1], sampling[1]) * wavelength
    alpha = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    phi = np.arctan2(ky[None, :], kx[:, None])
    return alpha, phi


def aberration_chi(alpha, phi, params, wavelength):
    """chi(alpha, phi) = 2 pi / lambda * sum_n alpha^(n+1)/(n+1) * (...)."""
    alpha2 = alpha**2
    chi = np.zeros(alpha.shape, dtype=np.float32)
    # radial factors, multiplied left to right after the 1/(n+1) prefactor
    radial = {1: (alpha2,), 2: (alpha2, alpha), 3: (alpha2**2,), 4: (alpha2**2, alpha),
              5: (alpha2**3,)}
    for order, terms, guard in _CHI_ORDERS:
        if not any(params[s] != 0.0 for s in guard):
            continue
        angular = 0.0
        for mag, ang, m in terms:
            if ang is None:
                angular = angular + params[mag]
            elif m == 1:
                angular = angular + params[mag] * np.cos(phi - params[ang])
            else:
                angular = angular + params[mag] * np.cos(m * (phi - params[ang]))
        term = 1 / (order + 1)
        for factor in radial[order]:
            term = term * factor
        chi += term * angular
    return 2 * np.pi / wavelength * chi


def aperture(alpha, semiangle_cutoff_mrad, rolloff_mrad, vacuum_probe_intensity=None):
    """Soft-edged objective aperture (cosine rolloff over ``rolloff`` mrad)."""
    if vacuum_probe_intensity is not None:
        v = np.asarray(vacuum_probe_intensity, dtype=np.float32)
        return np.sqrt(np.maximum(v, 0))
    if semiangle_cutoff_mrad == np.inf:
        return np.ones_like(alpha)
    cutoff = semiangle_cutoff_mrad / 1000
    if rolloff_mrad > 0.0:
        r = rolloff_mrad / 1000.0
        edge = 0.5 * (1 + np.cos(np.pi * (alpha - cutoff + r) / r))
        edge[alpha > cutoff] = 0.0
        return np.where(alpha > cutoff - r, edge, np.ones_like(alpha, dtype=np.float32))
    return np.array(alpha < cutoff).astype(np.float32)


class ComplexProbe:
    """Corner-centred complex probe on a (Sx, Sy) grid of ``sampling`` Angstrom.

    Mirrors the constructor and ``build()`` / ``_array`` surface of
    py4DSTEM's ``ComplexProbe`` (process/phase/utils.py:50-139, 420-426).
    """

    def __init__(self, energy, gpts, sampling, semiangle_cutoff=np.inf, rolloff=2.0,
                 vacuum_probe_intensity=None, parameters=None, **kwargs):
        self._energy = energy
        self._wavelength = electron_wavelength_angstrom(energy)
        self._gpts = tuple(int(g) for g in gpts)
        self._sampling = tuple(float(s) for s in sampling)
        self._semiangle_cutoff = semiangle_cutoff
        self._rolloff = rolloff
        self._vacuum_probe_intensity = vacuum_probe_intensity
        self._parameters = polar_parameters_from(parameters, **kwargs)

    def _evaluate_ctf(self):
        alpha, phi = scattering_angles(self._gpts, self._sampling, self._wavelength)
        ctf = np.exp(-1.0j * aberration_chi(alpha, phi, self._parameters, self._wavelength))
        if self._semiangle_cutoff < np.inf or self._vacuum_probe_intensity is not None:
            ctf *= aperture(alpha, self._semiangle_cutoff, self._rolloff,
                            self._vacuum_probe_intensity)
        return ctf

    def build(self):
        array = np.fft.ifft2(self._evaluate_ctf())
        self._array = array / np.sqrt((np.abs(array) ** 2).sum())
        return self
