"""Writes tests/golden/small_gd.npz: inputs and oracle outputs of a small
seeded GD reconstruction (6x6 scan, 64 px frames, 96^2 object, mini-batches of
12 with probe update, 3 iterations, seeded non-flat start).

The expected outputs come from this repository's NumPy oracle (the reference
itself cannot be imported in this environment); tests/test_golden.py checks
that the oracle still reproduces them and that the GPU path matches them.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.ndimage

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import OracleSingleslicePtychography  # noqa: E402
from py4dstem_amd.synthetic import small_workload  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_gd.npz")
RECON = dict(num_iter=3, max_batch_size=12, seed_random=7, fix_probe=False, step_size=0.5,
             normalization_min=1)


def case():
    w = small_workload(seed=21, scan=(6, 6), S=64, object_px=96)
    rng = np.random.default_rng(22)
    init = np.exp(0.02j * scipy.ndimage.gaussian_filter(rng.standard_normal((96, 96)), 3)
                  / 0.05).astype(np.complex64)
    return w, init


def main():
    w, init = case()
    o = OracleSingleslicePtychography(datacube=w["datacube"], **dict(
        w["init"], initial_object_guess=init.copy()))
    o.preprocess(**w["preprocess"])
    o.reconstruct(**RECON)
    np.savez_compressed(
        OUT, data=w["datacube"].data, positions_ang=w["positions_ang"], init_object=init,
        object=o.object, probe=o.probe, error_iterations=np.array(o.error_iterations),
        fov_mask=o._object_fov_mask, mean_intensity=np.float32(o._mean_diffraction_intensity))
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()
