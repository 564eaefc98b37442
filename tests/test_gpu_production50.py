"""The production workload at its stated iteration count, against the
oracle's committed trajectory.

tests/golden/production_gd50.npz holds the oracle's 50 gradient-descent
iterations (phase/singleslice_ptychography.py:847-983) on
production_workload(seed=0): 128x128 scan at 0.4 A, 128 px patterns, 720x720
object from the flat start, mini-batches of 1024 in the seeded order, probe
update.  It was written by tests/golden/make_production_trajectory.py, which
also ran the oracle in complex128 (NumPy < 2 semantics) to measure how far
two correct float implementations drift apart on this run.

Tolerance: every per-iteration error, the final object and the final probe
within max(1e-4, 3 x the oracle's own complex64/complex128 gap) relative;
the flat start takes exp(i angle F) of far-field pixels that are exactly zero
in the model, where rounding sets the angle.
"""

import hashlib
import os

import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "production_gd50.npz")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.mark.skipif(not os.path.exists(GOLDEN),
                    reason="tests/golden/production_gd50.npz not written yet "
                           "(tests/golden/make_production_trajectory.py)")
def test_production_50_iterations_match_oracle_trajectory():
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import production_workload
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    g = dict(np.load(GOLDEN, allow_pickle=False))
    say("production workload (seed 0), 50 iterations from the flat start")
    w = production_workload(seed=0, num_iter=50)
    digest = hashlib.sha256(np.ascontiguousarray(w["datacube"].data, dtype=np.float32)
                            .tobytes()).hexdigest()
    assert digest == str(g["datacube_sha256"]), "the generated patterns differ from the fixture's"
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    p.reconstruct(**w["reconstruct"])
    e64, e128 = g["errors_c64"], g["errors_c128"]
    err_gap = float(np.max(np.abs(e64 / e128 - 1)))
    errs = np.array(p.error_iterations)
    assert errs.shape == (50,)
    d_err = float(np.max(np.abs(errs / e64 - 1)))
    d_obj, d_probe = rel(p.object, g["object"]), rel(p.probe, g["probe"])
    say(f"object {d_obj:.2e} (oracle gap {float(g['object_gap']):.2e}), probe {d_probe:.2e} "
        f"(gap {float(g['probe_gap']):.2e}), errors {d_err:.2e} (gap {err_gap:.2e}); "
        f"error {errs[0]:.4e} -> {errs[-1]:.4e}")
    assert d_obj <= max(1e-4, 3 * float(g["object_gap"]))
    assert d_probe <= max(1e-4, 3 * float(g["probe_gap"]))
    assert d_err <= max(1e-4, 3 * err_gap)
