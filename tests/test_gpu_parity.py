"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Tolerances (relative L2 unless stated):
  * transforms and per-position terms: 1e-6 .. 2e-5 (float32 rounding);
  * scatter with an exact product (chi = 1): bit-exact against np.bincount;
  * reconstructions, well-conditioned start (non-flat seeded object):
    object and probe <= 1e-4 and every per-iteration error <= 1e-4;
  * reconstructions from the flat object of the small configuration:
    object and per-iteration error are each held to max(1e-4, 3 x floor),
    floor = the oracle's own complex64-vs-complex128 gap on the same input
    (1.9e-4 for the object, 3.9e-5 for the error at iteration 2), because
    exp(i angle F) of far-field pixels that are exactly zero in the
    flat-object model amplifies rounding noise in any reimplementation.
"""

import numpy as np
import pytest
import scipy.ndimage

import oracle.ptycho_oracle as po
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


def K():
    from py4dstem_amd import kernels
    return kernels


@pytest.mark.parametrize("S", [64, 128])
def test_fft2_matches_numpy(dev, S):
    say(f"fft2 S={S}")
    rng = np.random.default_rng(S)
    x = (rng.standard_normal((5, S, S)) + 1j * rng.standard_normal((5, S, S))).astype(np.complex64)
    xt = torch.from_numpy(x).to(dev)
    f = K().fft2(xt).cpu().numpy()
    assert rel(f, np.fft.fft2(x.astype(np.complex128))) < 5e-7
    b = K().fft2(torch.from_numpy(f).to(dev), inverse=True).cpu().numpy()
    assert rel(b, x) < 5e-7


@pytest.mark.parametrize("S", [64, 128])
def test_kernel_layout_roundtrip(dev, S):
    say(f"permute S={S}")
    rng = np.random.default_rng(1)
    a = rng.standard_normal((3, S, S)).astype(np.float32)
    t = torch.from_numpy(a).to(dev)
    k = K().to_kernel_layout(t)
    idx = K().kernel_index(S)
    np.testing.assert_array_equal(k.cpu().numpy(), a[:, idx][:, :, idx])
    np.testing.assert_array_equal(K().to_kernel_layout(k, inverse=True).cpu().numpy(), a)


def _prepared(o, dev):
    S = int(o._region_of_interest_shape[0])
    amps = K().to_kernel_layout(torch.from_numpy(np.ascontiguousarray(o._amplitudes)).to(dev))
    pos = torch.from_numpy(o._positions_px.copy()).to(dev)
    return S, amps, pos


def test_fused_forward_matches_oracle(dev, small_case):
    say("fused forward S=64")
    _, o = small_case
    S, amps, pos = _prepared(o, dev)
    rng = np.random.default_rng(4)
    obj = (o._object * np.exp(1j * 0.1 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    order = rng.permutation(o._num_diffraction_patterns)[:200].astype(np.int32)
    phat = K().probe_spectrum(torch.from_numpy(o._probe.astype(np.complex64)).to(dev))
    ps, chi, err, far = K().forward_adjoint(phat, torch.from_numpy(obj).to(dev), pos, amps,
                                            torch.from_numpy(order).to(dev), with_far=True)
    sp, op, ew, e = o.forward(obj, o._probe, o._positions_px[order], o._amplitudes[order])
    assert rel(ps.cpu().numpy(), sp) < 1e-6
    assert rel(chi.cpu().numpy(), ew) < 2e-5
    assert abs(err.sum().item() / e - 1) < 1e-6
    # far field from the staging layout, in kernel layout, against fft2
    k = K().kernel_index(S)
    F = np.fft.fft2((sp * op).astype(np.complex128))[:, k][:, :, k]
    assert rel(K().far_to_kernel_layout(far).cpu().numpy(), F) < 1e-6


def test_fused_forward_matches_oracle_128(dev):
    say("fused forward S=128")
    from py4dstem_amd.synthetic import production_workload
    w = production_workload(seed=2, scan=(6, 7), object_px=200, num_iter=1)
    o = OracleSingleslicePtychography(datacube=w["datacube"], **dict(
        w["init"], initial_object_guess=np.ones((200, 200), np.complex64)))
    o.preprocess(**w["preprocess"])
    S, amps, pos = _prepared(o, dev)
    rng = np.random.default_rng(5)
    obj = np.exp(1j * 0.2 * rng.standard_normal((200, 200))).astype(np.complex64)
    order = rng.permutation(42).astype(np.int32)
    phat = K().probe_spectrum(torch.from_numpy(o._probe.astype(np.complex64)).to(dev))
    ps, chi, err = K().forward_adjoint(phat, torch.from_numpy(obj).to(dev), pos, amps,
                                       torch.from_numpy(order).to(dev))
    sp, op, ew, e = o.forward(obj, o._probe, o._positions_px[order], o._amplitudes[order])
    assert rel(ps.cpu().numpy(), sp) < 1e-6
    assert rel(chi.cpu().numpy(), ew) < 2e-5
    assert abs(err.sum().item() / e - 1) < 1e-6


def test_scatter_bitexact_and_updates_match_oracle(dev, small_case):
    say("scatter + object/probe update")
    _, o = small_case
    S, amps, pos = _prepared(o, dev)
    rng = np.random.default_rng(6)
    B = 300
    order = rng.permutation(o._num_diffraction_patterns)[:B].astype(np.int32)
    order_t = torch.from_numpy(order).to(dev)
    psn = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    ps = torch.from_numpy(psn).to(dev)
    grad, spsum, tmax = K().scatter(ps, torch.ones_like(ps), pos, order_t, o._object.shape)
    ref = po.scatter_bincount(np.conj(psn), o._positions_px[order], (S, S), o._object.shape)
    np.testing.assert_array_equal(grad.cpu().numpy(), ref)
    assert tmax.max().item() == spsum.max().item()

    # several hit-list segments, repeated positions
    B2 = 2500
    order2 = rng.integers(0, o._num_diffraction_patterns, B2).astype(np.int32)
    ps2n = (rng.standard_normal((B2, S, S)) + 1j * rng.standard_normal((B2, S, S))).astype(
        np.complex64)
    ps2 = torch.from_numpy(ps2n).to(dev)
    g2, _, _ = K().scatter(ps2, torch.ones_like(ps2), pos, torch.from_numpy(order2).to(dev),
                           o._object.shape)
    np.testing.assert_array_equal(g2.cpu().numpy(), po.scatter_bincount(
        np.conj(ps2n), o._positions_px[order2], (S, S), o._object.shape))
    del ps2, g2

    obj = (o._object * np.exp(1j * 0.1 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    probe = o._probe.astype(np.complex64)
    obj_t = torch.from_numpy(obj).to(dev)
    probe_t = torch.from_numpy(probe).to(dev)
    phat = K().probe_spectrum(probe_t)
    ps, chi, _ = K().forward_adjoint(phat, obj_t, pos, amps, order_t)
    grad, spsum, tmax = K().scatter(ps, chi, pos, order_t, o._object.shape)
    K().probe_update(probe_t, chi, obj_t, pos, order_t, 0.5, 1)   # reads pre-update object
    K().object_update(obj_t, grad, spsum, tmax, 0.5, 1)
    sp, op, ew, _ = o.forward(obj, probe, o._positions_px[order], o._amplitudes[order])
    obj_ref, probe_ref = o.adjoint(obj.copy(), probe.copy(), op, sp, o._positions_px[order], ew,
                                   0.5, 1, False)
    assert rel(obj_t.cpu().numpy(), obj_ref) < 1e-6
    assert rel(probe_t.cpu().numpy(), probe_ref) < 1e-5


def test_constraints_match_oracle(dev, small_case):
    say("object constraints + probe CoM")
    _, o = small_case
    rng = np.random.default_rng(7)
    obj = (1.2 * np.exp(1j * (0.3 + 0.2 * rng.standard_normal(o._object.shape)))).astype(
        np.complex64)
    t = torch.from_numpy(obj).to(dev)
    mask = torch.from_numpy(o._object_fov_mask_inverse.astype(np.uint8)).to(dev)
    K().object_constraints(t, mask)
    ref = o.object_constraints(obj.copy(), o._object_fov_mask_inverse)
    assert rel(t.cpu().numpy(), ref) < 2e-6
    t2 = torch.from_numpy(obj).to(dev)
    K().object_constraints(t2, None, pure_phase_object=True)
    assert rel(t2.cpu().numpy(), o.object_constraints(obj.copy(), None, 0.0, True)) < 2e-6
    for S in (64, 128):
        p = po.ComplexProbeOracle(300e3, (S, S), (0.2, 0.2), 20.0, 2.0).build()._array
        p = (po.fft_shift(p.astype(np.complex64), np.array([1.3, -0.7])) * 30).astype(np.complex64)
        pt = torch.from_numpy(p).to(dev)
        K().probe_com_constraint(pt)
        assert rel(pt.cpu().numpy(), o.probe_com_constraint(p)) < 2e-6


def _init_object(shape, seed):
    rng = np.random.default_rng(seed)
    return np.exp(0.02j * scipy.ndimage.gaussian_filter(rng.standard_normal(shape), 3)
                  / 0.05).astype(np.complex64)


def _run_pair(w, init_obj, recon):
    from py4dstem_amd.ptychography import SingleslicePtychography
    out = {}
    for kind, cls in (("gpu", SingleslicePtychography), ("oracle", OracleSingleslicePtychography)):
        p = cls(datacube=w["datacube"], **dict(w["init"], initial_object_guess=init_obj.copy()))
        p.preprocess(**w["preprocess"])
        np.random.seed(1234)
        out[kind] = p.reconstruct(**recon)
    return out["gpu"], out["oracle"]


def _err_rel(g, o):
    return float(np.max(np.abs(np.array(g.error_iterations) / np.array(o.error_iterations) - 1)))


@pytest.mark.parametrize("variant", ["full_batch_fixed_probe", "minibatch_probe_update"])
def test_reconstruct_matches_oracle_well_conditioned(dev, variant):
    from py4dstem_amd.synthetic import small_workload
    say(f"reconstruct small {variant}")
    w = small_workload(seed=1)
    recon = dict(w["reconstruct"])
    if variant == "minibatch_probe_update":
        recon.update(max_batch_size=256, fix_probe=False, num_iter=8)
    g, o = _run_pair(w, _init_object((256, 256), 17), recon)
    np.testing.assert_array_equal(g._object_fov_mask, o._object_fov_mask)
    say(f"object {rel(g.object, o.object):.2e} probe {rel(g.probe, o.probe):.2e} "
        f"error {_err_rel(g, o):.2e}")
    assert rel(g.object, o.object) <= 1e-4
    assert rel(g.probe, o.probe) <= 1e-4
    assert _err_rel(g, o) <= 1e-4


def test_reconstruct_small_config_flat_init(dev):
    """The small configuration exactly as specified (flat object init)."""
    from py4dstem_amd.synthetic import small_workload
    say("reconstruct small config, flat init")
    w = small_workload(seed=1)
    g, o = _run_pair(w, np.ones((256, 256), np.complex64), dict(w["reconstruct"]))
    # the oracle's own complex64 vs complex128 gap on this input
    o128 = OracleSingleslicePtychography(datacube=w["datacube"], mode="c128", **dict(
        w["init"], initial_object_guess=np.ones((256, 256), np.complex64)))
    o128.preprocess(**w["preprocess"])
    np.random.seed(1234)
    o128.reconstruct(**w["reconstruct"])
    floor = rel(o.object, o128.object)
    err_floor = _err_rel(o, o128)
    say(f"object {rel(g.object, o.object):.2e} (c64-vs-c128 floor {floor:.2e}) "
        f"error {_err_rel(g, o):.2e} (floor {err_floor:.2e}, vs c128 {_err_rel(g, o128):.2e})")
    assert _err_rel(g, o) <= max(1e-4, 3 * err_floor)
    assert rel(g.object, o.object) <= max(1e-4, 3 * floor)


def test_reconstruct_is_deterministic(dev):
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import small_workload
    say("determinism")
    w = small_workload(seed=2, scan=(16, 16), object_px=160)
    res = []
    for _ in range(2):
        p = SingleslicePtychography(datacube=w["datacube"], **dict(
            w["init"], initial_object_guess=np.ones((160, 160), np.complex64)))
        p.preprocess(**w["preprocess"])
        p.reconstruct(**dict(w["reconstruct"], num_iter=3, max_batch_size=64, fix_probe=False))
        res.append((p.object.copy(), p.probe.copy(), list(p.error_iterations)))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]


@pytest.mark.parametrize("mixed", [False, True])
def test_reconstruct_with_detector_mask_matches_oracle(dev, mixed):
    """detector_fourier_mask (corner-centred like the amplitudes; 0/1 with a
    soft edge) through gradient descent with probe update, against the
    oracle: object, probe and per-iteration error <= 1e-4."""
    from oracle import OracleMixedstatePtychography
    from py4dstem_amd import MixedstatePtychography
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import small_workload
    say(f"reconstruct with detector mask, mixed={mixed}")
    w = small_workload(seed=1)
    q = np.fft.fftfreq(64)
    qq = np.sqrt(q[:, None] ** 2 + q[None, :] ** 2)
    mask = np.clip((0.4 - qq) / 0.05, 0, 1).astype(np.float32)  # soft high-q cut
    mask[5, 7] = 0.0  # a dead pixel
    recon = dict(w["reconstruct"], max_batch_size=256, fix_probe=False, num_iter=4,
                 detector_fourier_mask=mask)
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
    say(f"object {rel(g.object, o.object):.2e} probe {rel(np.asarray(g.probe), o.probe):.2e} "
        f"error {_err_rel(g, o):.2e}")
    assert rel(g.object, o.object) <= 1e-4
    assert rel(np.asarray(g.probe), o.probe) <= 1e-4
    assert _err_rel(g, o) <= 1e-4


def test_production_iteration_matches_oracle(dev):
    """One gradient-descent iteration at the production shapes (128x128 scan,
    128 px patterns, 720x720 object, 16 mini-batches of 1024 in seeded order,
    probe update) against the oracle: object, probe and error <= 1e-4, from
    the well-conditioned (seeded, non-flat) start of the reconstruction
    parity tests; from the flat start the far field has exactly-zero pixels
    whose exp(i angle F) amplifies rounding (see the flat-init test: 2e-4 at
    this size, as the oracle's own complex64-vs-complex128 gap is there).
    The oracle starts from the GPU class's preprocessed state (amplitudes,
    mean intensity, positions, probe, object, FOV mask), which the preprocess
    tests show to be bit-identical to the oracle's own preprocessing; its
    iteration takes ~20 s on one core."""
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import production_workload
    say("production-size iteration vs oracle")
    w = production_workload(seed=0, num_iter=1)
    g = SingleslicePtychography(datacube=w["datacube"], **dict(
        w["init"], initial_object_guess=_init_object((720, 720), 17)))
    g.preprocess(**w["preprocess"])
    o = OracleSingleslicePtychography(energy=w["init"]["energy"])
    for k in ("_amplitudes", "_mean_diffraction_intensity", "_num_diffraction_patterns", "_object",
              "_object_fov_mask_inverse", "_positions_px", "_region_of_interest_shape", "_probe"):
        v = getattr(g, k)
        setattr(o, k, np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
    o._object_initial = o._object.copy()
    o._probe_initial = o._probe.copy()
    o._positions_px_initial = o._positions_px.copy()
    o._object_fov_mask = np.invert(o._object_fov_mask_inverse)
    recon = dict(w["reconstruct"], num_iter=1, progress_bar=False)
    g.reconstruct(**recon)
    o.reconstruct(**{k: v for k, v in recon.items() if k != "progress_bar"})
    err = abs(g.error_iterations[0] / o.error_iterations[0] - 1)
    say(f"object {rel(g.object, o.object):.2e} probe {rel(g.probe, o.probe):.2e} error {err:.2e}")
    assert g.object.shape == (720, 720) and len(g.error_iterations) == 1
    assert rel(g.object, o.object) <= 1e-4
    assert rel(g.probe, o.probe) <= 1e-4
    assert err <= 1e-4


@pytest.mark.parametrize("scan,batch", [((6, 7), 10), ((6, 7), 1), ((1, 1), None), ((5, 5), 25)])
def test_ragged_and_degenerate_batches_match_oracle(dev, scan, batch):
    """Batching edge cases against the oracle, probe update on: near-equal
    ragged mini-batches (42 positions in batches of at most 10:
    9, 9, 8, 8, 8 as utils.py:539-589 splits them), one-position batches, a
    single scan position, and a batch equal to the scan.  Object, probe and
    every per-iteration error <= 1e-4."""
    from py4dstem_amd.synthetic import small_workload
    say(f"batches: scan {scan} max_batch_size {batch}")
    w = small_workload(seed=4, scan=scan, object_px=128)
    recon = dict(w["reconstruct"], max_batch_size=batch, fix_probe=False, num_iter=3)
    g, o = _run_pair(w, _init_object((128, 128), 21), recon)
    say(f"object {rel(g.object, o.object):.2e} probe {rel(g.probe, o.probe):.2e} "
        f"error {_err_rel(g, o):.2e}")
    assert len(g.error_iterations) == 3
    assert rel(g.object, o.object) <= 1e-4
    assert rel(g.probe, o.probe) <= 1e-4
    assert _err_rel(g, o) <= 1e-4
