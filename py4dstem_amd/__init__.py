"""MI355X-native single-slice gradient-descent ptychography.

A drop-in for py4DSTEM's ``SingleslicePtychography`` GD path whose iteration
runs in hand-written HIP kernels for gfx950 (``libpty_hip.so``, C-ABI in
``include/pty_hip.h``).
"""

from .datacube import Calibration, DataCube  # noqa: F401
from .probe import ComplexProbe, electron_wavelength_angstrom  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # the reconstruction class pulls in torch and the HIP library lazily
    if name == "SingleslicePtychography":
        from .ptychography import SingleslicePtychography
        return SingleslicePtychography
    if name == "MixedstatePtychography":
        from .mixedstate import MixedstatePtychography
        return MixedstatePtychography
    raise AttributeError(name)
