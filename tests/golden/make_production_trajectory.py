"""Writes tests/golden/production_gd50.npz: the oracle's 50-iteration
gradient-descent trajectory on the production workload, from the flat start.

Workload: ``py4dstem_amd.synthetic.production_workload(seed=0)`` -- 128x128
scan at 0.4 A, 128 px patterns at 0.1 A, 720x720 object initialised to ones,
mini-batches of 1024 in the seeded order, probe update, step 0.5,
normalization_min 1 (phase/singleslice_ptychography.py:847-983).

The oracle runs twice, in complex64 (NumPy >= 2 semantics, the parity target)
and complex128 (NumPy < 2 semantics, the reference's pinned NumPy), in two
processes.  Stored: both runs' per-iteration errors, the complex64 run's
final object and probe, the complex64/complex128 gaps of the final object and
probe (relative L2), and a checksum of the generated datacube so the GPU test
can confirm it regenerated the same patterns.  Each run takes ~20-40 min on
one core.

Run: python tests/golden/make_production_trajectory.py
"""
import hashlib
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "production_gd50.npz")
NUM_ITER = 50


def datacube_digest(data):
    """sha256 of the float32 patterns' bytes (C order)."""
    return hashlib.sha256(np.ascontiguousarray(data, dtype=np.float32).tobytes()).hexdigest()


def run(mode):
    from oracle import OracleSingleslicePtychography
    from py4dstem_amd.synthetic import production_workload

    w = production_workload(seed=0, num_iter=NUM_ITER)
    o = OracleSingleslicePtychography(datacube=w["datacube"], mode=mode, **w["init"])
    o.preprocess(**w["preprocess"])
    t0 = time.time()
    recon = {k: v for k, v in w["reconstruct"].items() if k != "progress_bar"}
    o.reconstruct(**recon)
    print(f"[{mode}] {NUM_ITER} iterations in {time.time() - t0:.0f} s", flush=True)
    return dict(object=np.asarray(o.object, np.complex64), probe=np.asarray(o.probe, np.complex64),
                errors=np.asarray(o.error_iterations, np.float64),
                digest=datacube_digest(w["datacube"].data))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def main():
    with mp.get_context("spawn").Pool(2) as pool:
        r64, r128 = pool.map(run, ["c64", "c128"])
    assert r64["digest"] == r128["digest"]
    np.savez_compressed(
        OUT, object=r64["object"], probe=r64["probe"], errors_c64=r64["errors"],
        errors_c128=r128["errors"], object_gap=np.float64(rel(r64["object"], r128["object"])),
        probe_gap=np.float64(rel(r64["probe"], r128["probe"])),
        datacube_sha256=np.array(r64["digest"]))
    print(OUT, os.path.getsize(OUT), "object gap", rel(r64["object"], r128["object"]),
          "probe gap", rel(r64["probe"], r128["probe"]))


if __name__ == "__main__":
    main()
