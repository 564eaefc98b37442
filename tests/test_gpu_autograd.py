"""Differentiation check of the hand-written adjoint behind
torch.autograd.Function.

* Small case: loss and gradients against the oracle's loss_gradients
  (complex128): relative L2 <= 1e-4; with integer positions the probe
  gradient equals the reference's probe term -2 S^2 sum_b conj(O_b) chi_b.
* Production shapes (128x128 scan, 128 px frames, 720^2 object): 1,000
  seeded object pixels and 200 probe pixels against float64 central finite
  differences (h = 1e-6) evaluated on the GPU with torch complex128 as the
  checker; tolerance 1e-3 relative (L2 over the sampled set, and per pixel
  where |g| exceeds 1 % of the sampled RMS).  The full object/probe gradient
  of a 2,048-position subset against the oracle adjoint: <= 1e-4.
"""

import numpy as np
import pytest
import scipy.ndimage

import oracle.ptycho_oracle as po
from tests.conftest import rel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


def _tensors(o, obj, probe, dev):
    from py4dstem_amd import kernels as K
    amps = K.to_kernel_layout(torch.from_numpy(np.ascontiguousarray(o._amplitudes)).to(dev))
    pos = torch.from_numpy(o._positions_px.copy()).to(dev)
    O = torch.from_numpy(obj).to(dev).requires_grad_(True)
    P = torch.from_numpy(probe).to(dev).requires_grad_(True)
    return O, P, pos, amps


def test_autograd_matches_oracle_adjoint_small(dev, frame_case):
    from py4dstem_amd.autograd import ptychographic_loss
    _, o = frame_case
    say(f"autograd small vs oracle, {o._amplitudes.shape[-1]} px frames")
    rng = np.random.default_rng(31)
    obj = (o._object * np.exp(1j * 0.1 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    probe = o._probe.astype(np.complex64)
    O, P, pos, amps = _tensors(o, obj, probe, dev)
    loss = ptychographic_loss(O, P, pos, amps, chunk=300)
    loss.backward()
    o128 = po.OracleSingleslicePtychography(energy=300e3, mode="c128")
    o128._region_of_interest_shape = o._region_of_interest_shape
    L, g_obj, g_probe = o128.loss_gradients(obj.astype(np.complex128), probe.astype(np.complex128),
                                            o._positions_px, o._amplitudes)
    say(f"loss {loss.item() / L - 1:.2e} obj {rel(O.grad.cpu().numpy(), g_obj):.2e} "
        f"probe {rel(P.grad.cpu().numpy(), g_probe):.2e}")
    assert abs(loss.item() / L - 1) < 1e-6
    assert rel(O.grad.cpu().numpy(), g_obj) <= 1e-4
    assert rel(P.grad.cpu().numpy(), g_probe) <= 1e-4


def test_autograd_probe_gradient_reduces_to_reference_term(dev, small_case):
    from py4dstem_amd.autograd import ptychographic_loss
    say("autograd integer positions")
    _, o = small_case
    rng = np.random.default_rng(32)
    obj = (o._object * np.exp(1j * 0.1 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    probe = o._probe.astype(np.complex64)
    ipos = np.round(o._positions_px).astype(np.float32)
    O, P, _, amps = _tensors(o, obj, probe, dev)
    pos = torch.from_numpy(ipos).to(dev)
    ptychographic_loss(O, P, pos, amps).backward()
    o128 = po.OracleSingleslicePtychography(energy=300e3, mode="c128")
    o128._region_of_interest_shape = o._region_of_interest_shape
    sp, op, chi, _ = o128.forward(obj.astype(np.complex128), probe.astype(np.complex128), ipos,
                                  o._amplitudes)
    ref_probe = -2 * 64 * 64 * np.sum(np.conj(op) * chi, axis=0)
    ref_obj = -2 * 64 * 64 * po.scatter_bincount(np.conj(sp) * chi, ipos, (64, 64), obj.shape)
    assert rel(P.grad.cpu().numpy(), ref_probe) <= 1e-4
    assert rel(O.grad.cpu().numpy(), ref_obj) <= 1e-4


@pytest.fixture(scope="module")
def production(dev):
    from py4dstem_amd import kernels as K
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import production_workload
    say("production workload for the finite-difference check")
    w = production_workload(seed=5)
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    rng = np.random.default_rng(33)
    phase = scipy.ndimage.gaussian_filter(rng.standard_normal(p._object_shape), 4)
    obj = np.exp(1j * 0.3 * phase / phase.std()).astype(np.complex64)
    return p, obj, p._probe.astype(np.complex64)


def _fp64_state(p, obj, probe, dev, idx):
    """float64 torch copies of the model quantities for positions idx."""
    S = probe.shape[0]
    pos = torch.from_numpy(p._positions_px[idx].astype(np.float64)).to(dev)
    r = torch.round(pos)
    frac = (pos - r).to(torch.float32).to(torch.float64)
    k = torch.fft.fftfreq(S, 1.0, dtype=torch.float64, device=dev)
    T = (torch.exp(-2j * np.pi * k[None, :, None] * frac[:, 0, None, None])
         * torch.exp(-2j * np.pi * k[None, None, :] * frac[:, 1, None, None]))
    P = torch.from_numpy(probe.astype(np.complex128)).to(dev)
    Pshift = torch.fft.ifft2(torch.fft.fft2(P)[None] * T)
    fi = torch.from_numpy(np.fft.fftfreq(S, 1 / S).astype(np.int64)).to(dev)
    x0 = r[:, 0].long()
    y0 = r[:, 1].long()
    Px, Py = obj.shape
    rows = (x0[:, None] + fi[None, :]) % Px
    cols = (y0[:, None] + fi[None, :]) % Py
    O = torch.from_numpy(obj.astype(np.complex128)).to(dev)
    Ob = O[rows[:, :, None], cols[:, None, :]]
    F = torch.fft.fft2(Pshift * Ob)
    A = torch.from_numpy(p._amplitudes[idx].astype(np.float64)).to(dev)
    return dict(T=T, Pshift=Pshift, Ob=Ob, F=F, A=A, rows=rows, cols=cols)


def _central_diff(F, A, D, h):
    lp = (torch.abs(F + h * D) - A) ** 2
    lm = (torch.abs(F - h * D) - A) ** 2
    return float(torch.sum(lp - lm).item()) / (2 * h)


def test_gradients_against_fp64_finite_differences_production(dev, production):
    from py4dstem_amd import kernels as K
    from py4dstem_amd.autograd import ptychographic_loss
    p, obj, probe = production
    S = probe.shape[0]
    N = p._num_diffraction_patterns
    say("autograd production gradients")
    O = torch.from_numpy(obj).to(dev).requires_grad_(True)
    P = torch.from_numpy(probe).to(dev).requires_grad_(True)
    ptychographic_loss(O, P, p._positions_dev, p._amplitudes_dev, chunk=4096).backward()
    g_obj = O.grad.cpu().numpy().astype(np.complex128)
    g_probe = P.grad.cpu().numpy().astype(np.complex128)
    h = 1e-6
    rng = np.random.default_rng(34)
    st = _fp64_state(p, obj, probe, dev, np.arange(N))
    kk = torch.arange(S, dtype=torch.float64, device=dev)

    # object pixels: only windows covering the pixel change; D_b = P_shift_b(i,j) e^{-2pi i (k i + l j)/S}
    fov = np.argwhere(p._object_fov_mask)
    pix = fov[rng.choice(len(fov), 1000, replace=False)]
    fd_o = np.zeros(len(pix), np.complex128)
    for n, (r0, c0) in enumerate(pix):
        hit_r = (st["rows"] == int(r0))
        hit_c = (st["cols"] == int(c0))
        cover = torch.nonzero(hit_r.any(1) & hit_c.any(1)).flatten()
        i = torch.argmax(hit_r[cover].int(), 1).double()
        j = torch.argmax(hit_c[cover].int(), 1).double()
        ps = st["Pshift"][cover, i.long(), j.long()]
        D = ps[:, None, None] * torch.exp(-2j * np.pi * (kk[None, :, None] * i[:, None, None]
                                                          + kk[None, None, :] * j[:, None, None]) / S)
        F, A = st["F"][cover], st["A"][cover]
        fd_o[n] = _central_diff(F, A, D, h) + 1j * _central_diff(F, A, 1j * D, h)
    got_o = g_obj[pix[:, 0], pix[:, 1]]

    # probe pixels: D_b = fft2(fft_shift(delta_uv, frac_b) * O_b), all positions
    ppix = np.stack([rng.integers(0, S, 200), rng.integers(0, S, 200)], axis=1)
    fd_p = np.zeros(len(ppix), np.complex128)
    for n, (u, v) in enumerate(ppix):
        e = torch.exp(-2j * np.pi * (kk[:, None] * int(u) + kk[None, :] * int(v)) / S)
        tot_r = tot_i = 0.0
        for s0 in range(0, N, 4096):
            sl = slice(s0, s0 + 4096)
            D = torch.fft.fft2(torch.fft.ifft2(e[None] * st["T"][sl]) * st["Ob"][sl])
            tot_r += _central_diff(st["F"][sl], st["A"][sl], D, h)
            tot_i += _central_diff(st["F"][sl], st["A"][sl], 1j * D, h)
        fd_p[n] = tot_r + 1j * tot_i
    got_p = g_probe[ppix[:, 0], ppix[:, 1]]

    for name, got, fd in (("object", got_o, fd_o), ("probe", got_p, fd_p)):
        l2 = np.linalg.norm(got - fd) / np.linalg.norm(fd)
        big = np.abs(fd) > 1e-2 * np.sqrt(np.mean(np.abs(fd) ** 2))
        worst = np.max(np.abs(got[big] - fd[big]) / np.abs(fd[big]))
        say(f"{name}: rel L2 {l2:.2e}, worst per-pixel {worst:.2e} over {big.sum()} pixels")
        assert l2 <= 1e-3
        assert worst <= 1e-3


def test_full_gradient_against_oracle_adjoint_production_subset(dev, production):
    from py4dstem_amd.autograd import ptychographic_loss
    p, obj, probe = production
    say("autograd production subset vs oracle")
    idx = np.random.default_rng(35).permutation(p._num_diffraction_patterns)[:2048]
    order = torch.from_numpy(idx.astype(np.int32)).to(dev)
    O = torch.from_numpy(obj).to(dev).requires_grad_(True)
    P = torch.from_numpy(probe).to(dev).requires_grad_(True)
    ptychographic_loss(O, P, p._positions_dev, p._amplitudes_dev, order=order).backward()
    o = po.OracleSingleslicePtychography(energy=300e3, mode="c64")
    o._region_of_interest_shape = np.array([128, 128])
    _, g_obj, g_probe = o.loss_gradients(obj, probe, p._positions_px[idx], p._amplitudes[idx])
    eo, ep = rel(O.grad.cpu().numpy(), g_obj), rel(P.grad.cpu().numpy(), g_probe)
    say(f"object {eo:.2e} probe {ep:.2e}")
    assert eo <= 1e-4 and ep <= 1e-4
