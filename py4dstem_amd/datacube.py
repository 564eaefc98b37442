"""Minimal 4D-STEM data container for the drop-in reconstruction class.

py4DSTEM's ``DataCube`` (py4DSTEM/datacube/datacube.py:21-71) is an emdfile
node; the reconstruction only touches ``.data`` (Rx, Ry, Qx, Qy), ``.Rshape``
(datacube.py:186) and the four calibration getters of ``Calibration``
(py4DSTEM/data/calibration.py:266-352, defaults 1 / "pixels" at :204-208).
Any object with that surface, including a real py4DSTEM ``DataCube``, is
accepted by :class:`py4dstem_amd.SingleslicePtychography`.
"""

from __future__ import annotations

import numpy as np


class Calibration:
    """Pixel sizes and units for real (R) and reciprocal (Q) space."""

    def __init__(self, R_pixel_size=1.0, R_pixel_units="pixels",
                 Q_pixel_size=1.0, Q_pixel_units="pixels"):
        self._R = [float(R_pixel_size), str(R_pixel_units)]
        self._Q = [float(Q_pixel_size), str(Q_pixel_units)]

    def get_R_pixel_size(self):
        return self._R[0]

    def get_R_pixel_units(self):
        return self._R[1]

    def get_Q_pixel_size(self):
        return self._Q[0]

    def get_Q_pixel_units(self):
        return self._Q[1]

    def set_R_pixel_size(self, x):
        self._R[0] = float(x)

    def set_R_pixel_units(self, x):
        self._R[1] = str(x)

    def set_Q_pixel_size(self, x):
        self._Q[0] = float(x)

    def set_Q_pixel_units(self, x):
        self._Q[1] = str(x)


class DataCube:
    """A (Rx, Ry, Qx, Qy) intensity array with a :class:`Calibration`."""

    def __init__(self, data, calibration: Calibration | None = None, name="datacube"):
        data = np.asarray(data)
        if data.ndim != 4:
            raise ValueError(f"DataCube data must be 4D (Rx,Ry,Qx,Qy), got shape {data.shape}")
        self.data = data
        self.calibration = calibration if calibration is not None else Calibration()
        self.name = name

    @property
    def shape(self):
        return self.data.shape

    @property
    def Rshape(self):
        return self.data.shape[:2]

    @property
    def Qshape(self):
        return self.data.shape[2:]
