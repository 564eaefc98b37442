"""Degenerate inputs on the GPU against the oracle, and C-ABI errors raised
through the bound entry points.

* All-zero patterns: A = 0 gives chi = -psi and err = sum |F|^2
  (phase/ptychographic_methods.py:1576-1635).
* Zero far field: an object that is zero over a whole window gives F = 0
  there.  The GPU takes angle(0) = 0 there (chi = ifft2(A)).  NumPy's
  exp(1j angle(F)) at an exact zero follows the sign bits pocketfft leaves on
  it (angle pi at -0 + 0j), which depend on the signs of the zeros it was
  given, so the reference's value there is an artefact of its FFT's
  operation order that no other FFT reproduces; the oracle comparison below
  therefore keeps every far field nonzero.
* NaN counts: NumPy carries NaN through sqrt(max(I, 0)), the mean intensity,
  the error and every object pixel the NaN pattern's window touches; the GPU
  path must give the same NaN pattern and the same finite values.
* Errors: bad sizes and unsupported frames reach the caller as RuntimeError
  (library codes) or ValueError / NotImplementedError (argument checks), the
  reference's error types for bad arguments.
"""

import numpy as np
import pytest
import scipy.ndimage

from oracle import OracleSingleslicePtychography
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


def _start(shape, seed=17):
    rng = np.random.default_rng(seed)
    return np.exp(0.02j * scipy.ndimage.gaussian_filter(rng.standard_normal(shape), 3)
                  / 0.05).astype(np.complex64)


def _pair(w, init_obj, recon, probe=None, pre=None):
    from py4dstem_amd.ptychography import SingleslicePtychography
    out = {}
    for kind, cls in (("gpu", SingleslicePtychography), ("oracle", OracleSingleslicePtychography)):
        kw = dict(w["init"], initial_object_guess=init_obj.copy())
        if probe is not None:
            kw["initial_probe_guess"] = probe.copy()
        p = cls(datacube=w["datacube"], **kw)
        p.preprocess(**(pre or w["preprocess"]))
        rc = dict(recon)
        if kind == "gpu":
            rc["progress_bar"] = False
        out[kind] = p.reconstruct(**rc)
    return out["gpu"], out["oracle"]


def test_zero_patterns_match_oracle(dev):
    """Two all-zero patterns (A = 0: chi = -psi, err = sum |F|^2) and an
    object that is zero over part of the field: 3 iterations with
    mini-batches and probe update, object, probe and every error <= 1e-4
    against the oracle."""
    from py4dstem_amd.synthetic import small_workload
    w = small_workload(seed=2, scan=(10, 10), object_px=160)
    w["datacube"].data[3, 4] = 0.0
    w["datacube"].data[7, 1] = 0.0
    init = _start((160, 160))
    init[20:60, 20:60] = 0.0          # zero over part of some windows, never a whole one
    recon = dict(w["reconstruct"], num_iter=3, max_batch_size=25, fix_probe=False)
    g, o = _pair(w, init, recon)
    err = float(np.max(np.abs(np.array(g.error_iterations) / np.array(o.error_iterations) - 1)))
    say(f"zero frames: object {rel(g.object, o.object):.2e} probe {rel(g.probe, o.probe):.2e} "
        f"error {err:.2e}")
    assert rel(g.object, o.object) <= 1e-4
    assert rel(g.probe, o.probe) <= 1e-4
    assert err <= 1e-4


def test_zero_pattern_and_zero_patch_single_batch_terms(dev):
    """The fused kernel's terms for one batch against the oracle's forward: a
    zero pattern gives chi = -psi.  A position whose whole window lies in a
    zero region of the object has F = 0 exactly; the GPU takes angle 0 there
    (chi = ifft2(A)), and the oracle's NumPy value is ifft2(A u) with u = +1
    or -1 by the sign bit pocketfft leaves on each zero."""
    from py4dstem_amd import kernels as K
    from py4dstem_amd.synthetic import small_workload
    w = small_workload(seed=5, scan=(4, 4), object_px=128)
    o = OracleSingleslicePtychography(datacube=w["datacube"], **w["init"])
    o.preprocess(**w["preprocess"])
    amps = o._amplitudes.copy()
    amps[2] = 0.0
    pos = o._positions_px
    phat = K.probe_spectrum(torch.from_numpy(o._probe).to(dev))
    ampk = K.to_kernel_layout(torch.from_numpy(amps).to(dev))
    posd = torch.from_numpy(pos.copy()).to(dev)
    # (a) zero pattern, ordinary object: every slot against the oracle
    obj = _start((128, 128))
    sp, op, chi, _ = o.forward(obj, o._probe, pos, amps)
    np.testing.assert_allclose(chi[2], -(sp[2] * op[2]), atol=1e-5 * np.abs(sp[2]).max())
    gchi = K.forward_adjoint(phat, torch.from_numpy(obj).to(dev), posd, ampk,
                             torch.arange(16, dtype=torch.int32, device=dev))[1].cpu().numpy()
    say(f"zero pattern: chi {rel(gchi, chi):.2e}, slot 2 {rel(gchi[2], chi[2]):.2e}")
    assert rel(gchi[2], chi[2]) <= 1e-5
    assert rel(gchi, chi) <= 2e-5
    # (b) slot 7 alone over an object that is zero around it: F = 0 exactly
    x0, y0 = np.round(pos[7]).astype(int)
    obj[x0 - 40:x0 + 40, y0 - 40:y0 + 40] = 0.0
    sp, op, chi, _ = o.forward(obj, o._probe, pos[7:8], amps[7:8])
    assert np.all(op[0] == 0)
    u = np.exp(1j * np.angle(np.fft.fft2((sp[0] * op[0]).astype(np.complex64))))
    np.testing.assert_allclose(chi[0], np.fft.ifft2(amps[7] * u), atol=1e-6 * np.abs(amps[7]).max())
    g7 = K.forward_adjoint(phat, torch.from_numpy(obj).to(dev), posd, ampk,
                           torch.tensor([7], dtype=torch.int32, device=dev))[1].cpu().numpy()
    assert rel(g7[0], np.fft.ifft2(amps[7].astype(np.float64))) <= 1e-6


def test_nan_pattern_propagates_as_numpy(dev):
    """A NaN count in one pattern: the bilinear shift spreads it to four
    amplitudes of that pattern and the mean intensity becomes NaN.  With the
    probe given (so the NaN mean does not reach it), one full-batch
    iteration with a fixed probe and no shrinkage (whose masked mean phase
    would spread NaN over the whole object in both implementations) gives a
    NaN error, NaN object pixels exactly where the oracle has them (the
    pattern's 64 x 64 window), and the finite pixels within 1e-4."""
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import small_workload
    w = small_workload(seed=3, scan=(8, 8), object_px=128)
    clean = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    clean.preprocess(**w["preprocess"])
    probe = np.asarray(clean._probe).copy()
    w["datacube"].data[4, 5, 30, 31] = np.nan
    recon = dict(w["reconstruct"], num_iter=1, max_batch_size=None, fix_probe=True,
                 fix_potential_baseline=False)
    g, o = _pair(w, _start((128, 128)), recon, probe=probe)
    assert np.isnan(g._amplitudes).sum() == np.isnan(o._amplitudes).sum() == 4
    assert np.isnan(o._mean_diffraction_intensity) and np.isnan(g._mean_diffraction_intensity)
    assert np.isnan(o.error_iterations[0]) and np.isnan(g.error_iterations[0])
    nan_o, nan_g = np.isnan(o.object), np.isnan(g.object)
    say(f"NaN pixels oracle {nan_o.sum()} gpu {nan_g.sum()}")
    assert nan_o.sum() == 64 * 64
    np.testing.assert_array_equal(nan_g, nan_o)
    fin = ~nan_o
    assert rel(g.object[fin], o.object[fin]) <= 1e-4


def test_abi_errors_reach_the_caller(dev):
    from py4dstem_amd import _lib
    from py4dstem_amd import kernels as K
    x = torch.zeros((2, 48, 48), dtype=torch.complex64, device=dev)
    with pytest.raises(NotImplementedError):
        K.fft2(x)
    y = torch.zeros((2, 64, 64), dtype=torch.complex64, device=dev)
    with pytest.raises(RuntimeError, match="not supported"):
        _lib.call("pty_fft2_c64", _lib.ptr(x), _lib.ptr(y), 2, 48, 0, _lib.stream_handle())
    with pytest.raises(RuntimeError, match="in place"):
        _lib.call("pty_fft2_c64", _lib.ptr(y), _lib.ptr(y), 2, 64, 0, _lib.stream_handle())
    with pytest.raises(ValueError):
        K.fft2(torch.zeros((2, 64, 64), dtype=torch.float32, device=dev))
    # a valid call afterwards still works: the failed ones launched nothing
    z = K.fft2(torch.ones((1, 64, 64), dtype=torch.complex64, device=dev))
    torch.cuda.synchronize()
    assert abs(complex(z[0, 0, 0].item()) - 4096) < 1e-3
