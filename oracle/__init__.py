"""CPU oracle for single-slice gradient-descent ptychography.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference algorithm (py4DSTEM 0.14.18, ``SingleslicePtychography``) used as
the checker for the HIP product path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``py4dstem_amd`` never imports, links or runs
anything from here, and it has no CPU fallback that could route through it.

Every function cites the reference file:line it restates (paths relative to
the reference repository root).  The pinning status of each function (which
closed form, invariant or brute-force check in ``tests/test_oracle_*.py``
fixes it) is listed in ``ptycho_oracle.PINNING``.
"""

from .ptycho_oracle import (  # noqa: F401
    PINNING,
    ComplexProbeOracle,
    OracleMixedstatePtychography,
    OracleSingleslicePtychography,
    electron_wavelength_angstrom,
    fft_shift,
    fourier_translation_operator,
    generate_batches,
    get_CoM,
    get_shifted_ar,
    patch_indices,
    scatter_bincount,
    subdivide_into_batches,
)
