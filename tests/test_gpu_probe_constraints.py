"""GPU probe constraints fix_probe_aperture and constrain_probe_amplitude
(ptychographic_constraints.py:920-957, 1004-1033, applied after the probe
recentring at :1126-1147 and, with several modes, :1258-1282) against the
oracle, through whole gradient-descent reconstructions with probe update.

Tolerances: object, probe and every per-iteration error <= 1e-4 relative
(float32 transforms and sums on the GPU; the oracle follows the reference's
NumPy dtypes, including the float64 top-hat mask).
"""

import numpy as np
import pytest
import scipy.ndimage

from oracle import OracleMixedstatePtychography, OracleSingleslicePtychography
from tests.conftest import rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    from py4dstem_amd import _lib
    _lib.load()
    return torch.device("cuda", 0)


def _init_object(shape, seed):
    rng = np.random.default_rng(seed)
    return np.exp(0.02j * scipy.ndimage.gaussian_filter(rng.standard_normal(shape), 3)
                  / 0.05).astype(np.complex64)


def _err_rel(g, o):
    return float(np.max(np.abs(np.array(g.error_iterations) / np.array(o.error_iterations) - 1)))


CASES = {
    "aperture": dict(fix_probe_aperture=True),
    "amplitude": dict(constrain_probe_amplitude=True,
                      constrain_probe_amplitude_relative_radius=0.3,
                      constrain_probe_amplitude_relative_width=0.04),
    "both": dict(fix_probe_aperture=True, constrain_probe_amplitude=True,
                 constrain_probe_amplitude_relative_radius=0.35),
}


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("case", list(CASES))
def test_reconstruct_with_probe_constraints_matches_oracle(dev, case, mixed):
    from py4dstem_amd import MixedstatePtychography
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import small_workload
    say(f"reconstruct with probe constraints {case}, mixed={mixed}")
    w = small_workload(seed=1)
    recon = dict(w["reconstruct"], max_batch_size=256, fix_probe=False, num_iter=4,
                 **CASES[case])
    if mixed:
        recon["orthogonalize_probe"] = False
    init_obj = _init_object((256, 256), 17)
    out = {}
    pairs = ((("gpu", MixedstatePtychography), ("oracle", OracleMixedstatePtychography)) if mixed
             else (("gpu", SingleslicePtychography), ("oracle", OracleSingleslicePtychography)))
    for kind, cls in pairs:
        kw = dict(w["init"], initial_object_guess=init_obj.copy())
        if mixed:
            kw["num_probes"] = 2
        p = cls(datacube=w["datacube"], **kw)
        np.random.seed(5)
        p.preprocess(**w["preprocess"])
        np.random.seed(1234)
        out[kind] = p.reconstruct(**recon)
    g, o = out["gpu"], out["oracle"]
    gp = np.asarray(g.probe)
    say(f"object {rel(g.object, o.object):.2e} probe {rel(gp, o.probe):.2e} "
        f"error {_err_rel(g, o):.2e}")
    assert rel(g.object, o.object) <= 1e-4
    assert rel(gp, o.probe) <= 1e-4
    assert _err_rel(g, o) <= 1e-4
    if CASES[case].get("fix_probe_aperture") and not CASES[case].get("constrain_probe_amplitude"):
        # the Fourier amplitude of (mode 0 of) the probe is the initial one, up
        # to the intensity normalisation
        p0 = gp[0] if mixed else gp
        a = np.abs(np.fft.fft2(np.fft.ifftshift(p0)))
        ap = np.asarray(g._probe_initial_aperture, dtype=np.float64)
        ap = ap[0] if ap.ndim == 3 else ap
        scale = a.sum() / ap.sum()
        assert np.max(np.abs(a - scale * ap)) <= 1e-4 * scale * ap.max()


def test_probe_constraints_off_when_probe_fixed(dev):
    """With fix_probe the constraints do not run (the reference applies every
    probe constraint only when the probe is updated): the probe is unchanged."""
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import small_workload
    w = small_workload(seed=1, scan=(8, 8), object_px=128)
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    before = p._probe.copy()
    p.reconstruct(num_iter=2, max_batch_size=16, fix_probe=True, fix_probe_aperture=True,
                  constrain_probe_amplitude=True, progress_bar=False)
    assert np.array_equal(p._probe_dev.cpu().numpy(), before)
