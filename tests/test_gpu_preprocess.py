"""Preprocessing kernels against the oracle.

* pty_shift_amplitudes: the bilinear periodic shift by -CoM and sqrt are
  bit-identical to the oracle's get_shifted_ar + sqrt(max(I, 0))
  (reference float32 arithmetic), including selected subsets; the float64
  per-pattern sums agree with NumPy float64 sums to 1e-12.
* pty_pattern_com: centre of mass against the oracle's get_CoM (float64),
  relative 1e-6, with and without a detector mask.
* preprocess(): kernel-layout amplitudes and the mean intensity against the
  oracle's preprocess on the same datacube (bitwise / 1e-6).
"""

import numpy as np
import pytest

import oracle.ptycho_oracle as po

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


@pytest.mark.parametrize("S", [64, 128])
def test_shift_amplitudes_bitwise(dev, S):
    from py4dstem_amd import kernels as K
    say(f"shift amplitudes S={S}")
    rng = np.random.default_rng(S)
    n = 11
    frames = rng.uniform(0, 50, (n, S, S)).astype(np.float32)
    frames[0, :3, :3] = -2.0  # negative counts clip to 0 before the sqrt
    sx = (-rng.uniform(S / 2 - 3, S / 2 + 3, n)).astype(np.float32)
    sy = (-rng.uniform(S / 2 - 3, S / 2 + 3, n)).astype(np.float32)
    sx[1], sy[1] = -S / 2, -S / 2  # integer shift: an exact roll
    sel = np.array([3, 0, 7, 1, 10], np.int32)
    for select in (None, sel):
        amps, sums = K.shift_amplitudes(torch.from_numpy(frames).to(dev), sx, sy, select)
        nat = K.to_kernel_layout(amps, inverse=True).cpu().numpy()
        idx = np.arange(n) if select is None else select
        for k, src in enumerate(idx):
            shifted = po.get_shifted_ar(frames[src], sx[src], sy[src])
            np.testing.assert_array_equal(nat[k], np.sqrt(np.maximum(shifted, 0)))
            assert abs(sums[k].item() / np.sum(shifted, dtype=np.float64) - 1) < 1e-12


def test_pattern_com(dev):
    from py4dstem_amd import kernels as K
    say("pattern centre of mass")
    rng = np.random.default_rng(5)
    frames = rng.uniform(0, 10, (6, 64, 64)).astype(np.float32)
    mask = (rng.uniform(size=(64, 64)) > 0.2).astype(np.float32)
    for m in (None, mask):
        cx, cy = K.pattern_com(torch.from_numpy(frames).to(dev), m)
        for k in range(6):
            ar = frames[k].astype(np.float64) * (1 if m is None else m)
            rx, ry = po.get_CoM(ar)
            assert abs(cx[k] / rx - 1) < 1e-6 and abs(cy[k] / ry - 1) < 1e-6


def test_preprocess_matches_oracle(dev, small_case):
    from py4dstem_amd import SingleslicePtychography
    from py4dstem_amd import kernels as K
    say("preprocess vs oracle")
    w, o = small_case
    init = dict(w["init"], initial_object_guess=w["init"]["initial_object_guess"].copy())
    p = SingleslicePtychography(datacube=w["datacube"], **init)
    p.preprocess(**w["preprocess"])
    nat = K.to_kernel_layout(p._amplitudes_dev, inverse=True).cpu().numpy()
    np.testing.assert_array_equal(nat, o._amplitudes)
    np.testing.assert_array_equal(p._amplitudes, o._amplitudes)
    assert abs(p._mean_diffraction_intensity / o._mean_diffraction_intensity - 1) < 1e-6
    np.testing.assert_array_equal(p._object_fov_mask, o._object_fov_mask)
