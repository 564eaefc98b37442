"""Kernel phase logic on the CPU: the exact phase functions the HIP kernels
run, executed by the emulator (tests/emu/pty_emu.cpp) with LDS race and
bounds checking, compared with the oracle."""

import ctypes

import numpy as np
import pytest

import oracle.ptycho_oracle as po
from tests.conftest import cptr, far_staging_to_kernel, rel


def kidx(emu, S):
    return np.array([emu.emu_kidx(S, p) for p in range(S)])


@pytest.mark.parametrize("S", [32, 64, 128])
def test_fft_passes_match_numpy_and_are_race_free(emu, S):
    emu.emu_reset()
    rng = np.random.default_rng(S)
    x = (rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S))).astype(np.complex64)
    k = kidx(emu, S)
    assert sorted(k) == list(range(S))
    out = np.zeros_like(x)
    emu.emu_fft2_perm(S, 0, cptr(x), cptr(out))
    nat = np.empty_like(out)
    nat[np.ix_(k, k)] = out
    assert rel(nat, np.fft.fft2(x.astype(np.complex128))) < 5e-7
    back = np.zeros_like(x)
    emu.emu_fft2_perm(S, 1, cptr(out.copy()), cptr(back))
    assert rel(back, x) < 5e-7
    assert emu.emu_violations() == 0, emu.emu_first_violation()


def test_fused_position_matches_oracle_forward(emu, frame_case):
    w, o = frame_case
    S = o._amplitudes.shape[-1]
    rng = np.random.default_rng(4)
    obj = (o._object * np.exp(1j * 0.1 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    probe = o._probe.astype(np.complex64)
    order = rng.permutation(o._num_diffraction_patterns)[:12].astype(np.int32)
    pos = o._positions_px
    k = kidx(emu, S)
    phat = np.ascontiguousarray(np.fft.fft2(probe)[np.ix_(k, k)].astype(np.complex64))
    amps = np.ascontiguousarray(o._amplitudes[:, k][:, :, k])
    B = len(order)
    ps = np.zeros((B, S, S), np.complex64)
    chi = np.zeros_like(ps)
    far = np.zeros_like(ps)
    err = np.zeros(B)
    emu.emu_reset()
    emu.emu_fused(S, cptr(phat), cptr(obj), obj.shape[0], obj.shape[1], cptr(pos), len(pos),
                  cptr(amps), cptr(order), B, cptr(ps), cptr(chi), cptr(err), cptr(far))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    sp, op, ew, e = o.forward(obj, probe, pos[order], o._amplitudes[order])
    assert rel(ps, sp) < 1e-6
    assert rel(chi, ew) < 2e-5
    assert abs(err.sum() / e - 1) < 1e-6
    F = np.fft.fft2(sp * op)
    farn = np.empty_like(far)
    farn[:, k[:, None], k[None, :]] = far_staging_to_kernel(far)
    assert rel(farn, F) < 1e-6


@pytest.mark.parametrize("S", [32, 64])
def test_fused_position_edge_wrap_and_half_pixels(emu, S):
    """Positions on the object edges (periodic wrap) and exact .5 fractions."""
    rng = np.random.default_rng(8)
    Px, Py = 64, 70  # object as small as the frame
    obj = np.exp(1j * rng.uniform(0, 0.5, (Px, Py))).astype(np.complex64)
    probe = po.ComplexProbeOracle(300e3, (S, S), (0.2, 0.2), 20.0, 2.0).build()._array.astype(
        np.complex64) * 20
    pos = np.array([[0.5, 69.5], [63.49, -0.5], [31.5, 35.5], [-7.2, 80.1]], np.float32)
    order = np.arange(4, dtype=np.int32)
    amps_nat = rng.uniform(0, 3, (4, S, S)).astype(np.float32)
    k = kidx(emu, S)
    amps = np.ascontiguousarray(amps_nat[:, k][:, :, k])
    phat = np.ascontiguousarray(np.fft.fft2(probe)[np.ix_(k, k)].astype(np.complex64))
    ps = np.zeros((4, S, S), np.complex64)
    chi = np.zeros_like(ps)
    err = np.zeros(4)
    emu.emu_reset()
    emu.emu_fused(S, cptr(phat), cptr(obj), Px, Py, cptr(pos), 4, cptr(amps), cptr(order), 4,
                  cptr(ps), cptr(chi), cptr(err), None)
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    o = po.OracleSingleslicePtychography(energy=300e3)
    o._region_of_interest_shape = np.array([S, S])
    sp, _, ew, e = o.forward(obj, probe, pos, amps_nat)
    assert rel(ps, sp) < 1e-6 and rel(chi, ew) < 2e-5


def test_scatter_is_bincount_in_batch_order(emu, small_case):
    _, o = small_case
    S = 64
    rng = np.random.default_rng(6)
    B = 40
    order = rng.permutation(o._num_diffraction_patterns)[:B].astype(np.int32)
    pos = o._positions_px
    ps = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    ones = np.ones_like(ps)
    shape = o._object.shape
    ntiles = emu.emu_gather_num_tiles(shape[0], shape[1])
    grad = np.zeros(shape, np.complex64)
    spsum = np.zeros(shape, np.float32)
    tmax = np.zeros(ntiles, np.float32)
    emu.emu_reset()
    emu.emu_gather(S, cptr(ps), cptr(ones), cptr(pos), len(pos), cptr(order), B, shape[0],
                   shape[1], cptr(grad), cptr(spsum), cptr(tmax))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    # chi = 1 makes the float32 product exact: the float64 sums must then be
    # bit-identical to np.bincount in batch order
    ref = po.scatter_bincount(np.conj(ps), pos[order], (S, S), shape)
    np.testing.assert_array_equal(grad, ref)
    sq = (ps.real * ps.real + ps.imag * ps.imag).astype(np.float32)
    np.testing.assert_array_equal(spsum, po.scatter_bincount(sq, pos[order], (S, S), shape))
    assert tmax.max() == spsum.max()


def test_scatter_multi_segment_order(emu, small_case):
    """More batch slots than one hit-list segment holds (and repeated
    positions): per-pixel float64 order must still be the batch order."""
    _, o = small_case
    S = 64
    rng = np.random.default_rng(12)
    B = 1500
    order = rng.integers(0, o._num_diffraction_patterns, B).astype(np.int32)
    pos = o._positions_px
    ps = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    shape = o._object.shape
    grad = np.zeros(shape, np.complex64)
    spsum = np.zeros(shape, np.float32)
    tmax = np.zeros(emu.emu_gather_num_tiles(*shape), np.float32)
    emu.emu_reset()
    emu.emu_gather(S, cptr(ps), cptr(np.ones_like(ps)), cptr(pos), len(pos), cptr(order), B,
                   shape[0], shape[1], cptr(grad), cptr(spsum), cptr(tmax))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    np.testing.assert_array_equal(grad, po.scatter_bincount(np.conj(ps), pos[order], (S, S), shape))


def test_probe_gradient_pass_matches_oracle(emu, frame_case):
    """Backward probe pass: fft2(conj(O_b) chi_b) conj(T_b) per position;
    summed and inverse transformed it is the oracle's true probe adjoint."""
    _, o = frame_case
    S = o._amplitudes.shape[-1]
    rng = np.random.default_rng(13)
    obj = (o._object * np.exp(1j * 0.1 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    B = 10
    order = rng.permutation(o._num_diffraction_patterns)[:B].astype(np.int32)
    pos = o._positions_px
    chi = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    out = np.zeros_like(chi)
    emu.emu_reset()
    emu.emu_probe_grad(S, cptr(chi), cptr(obj), obj.shape[0], obj.shape[1], cptr(pos), len(pos),
                       cptr(order), B, cptr(out))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    k = kidx(emu, S)
    nat = np.empty_like(out)
    nat[:, k[:, None], k[None, :]] = out
    rows, cols = po.patch_indices(pos[order], (S, S), obj.shape)
    frac = pos[order] - np.round(pos[order])
    T = po.fourier_translation_operator(frac, (S, S))
    ref = np.fft.fft2(np.conj(obj[rows, cols]) * chi) * np.conj(T)
    assert rel(nat, ref) < 1e-6


def test_staged_scatter_is_bitwise_identical(emu, small_case):
    """A mini-batch scattered in stages with carried float64 sums equals the
    single-pass scatter bit for bit (same per-pixel summation order)."""
    _, o = small_case
    S = 64
    rng = np.random.default_rng(14)
    B = 300
    order = rng.permutation(o._num_diffraction_patterns)[:B].astype(np.int32)
    pos = o._positions_px
    ps = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    chi = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    shape = o._object.shape
    nt = emu.emu_gather_num_tiles(*shape)
    outs = []
    for stage in (B, 64, 128):
        grad = np.zeros(shape, np.complex64)
        spsum = np.zeros(shape, np.float32)
        tmax = np.zeros(nt, np.float32)
        run = np.zeros(shape + (3,), np.float64)
        emu.emu_reset()
        emu.emu_gather_staged(S, cptr(ps), cptr(chi), cptr(pos), len(pos), cptr(order), B, stage,
                              shape[0], shape[1], cptr(grad), cptr(spsum), cptr(tmax), cptr(run))
        assert emu.emu_violations() == 0, emu.emu_first_violation()
        outs.append((grad, spsum, tmax))
    for g, s_, t in outs[1:]:
        np.testing.assert_array_equal(g, outs[0][0])
        np.testing.assert_array_equal(s_, outs[0][1])
        np.testing.assert_array_equal(t, outs[0][2])


def test_simulation_pass_matches_forward_model(emu, frame_case):
    """Forward-model pass: |fft2(P_shift_b O_b)|^2 in natural order, equal to
    the oracle's forward model (and race-free)."""
    _, o = frame_case
    S = o._amplitudes.shape[-1]
    rng = np.random.default_rng(15)
    obj = (o._object * np.exp(1j * 0.2 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    pos = o._positions_px[:6].copy()
    k = kidx(emu, S)
    probe = o._probe.astype(np.complex64)
    phat = np.ascontiguousarray(np.fft.fft2(probe)[np.ix_(k, k)].astype(np.complex64))
    out = np.zeros((6, S, S), np.float32)
    emu.emu_reset()
    emu.emu_simulate(S, cptr(phat), cptr(obj), obj.shape[0], obj.shape[1], cptr(pos), 6, cptr(out))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    rows, cols = po.patch_indices(pos, (S, S), obj.shape)
    ref = np.abs(np.fft.fft2(po.fft_shift(probe, pos - np.round(pos)) * obj[rows, cols])) ** 2
    assert rel(out, ref) < 1e-6


def test_scatter_of_32px_stitching_windows(emu):
    """The scatter with S = 32 (learned-initialiser stitching): weighted
    patches of one object average back to that object where covered."""
    rng = np.random.default_rng(16)
    S, P = 32, 96
    obj = np.exp(1j * rng.uniform(0, 0.5, (P, P))).astype(np.complex64)
    g = np.arange(16, 81, 8)
    gx, gy = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32)
    B = len(pos)
    u = np.arange(S)
    w1 = np.sin(np.pi * (u + 0.5) / S) ** 2
    w = np.fft.ifftshift(np.outer(w1, w1))            # corner-centred window
    rows, cols = po.patch_indices(pos, (S, S), (P, P))
    patches = obj[rows, cols]
    ps = np.broadcast_to(np.sqrt(w), (B, S, S)).astype(np.complex64).copy()
    chi = (np.sqrt(w) * patches).astype(np.complex64)
    order = np.arange(B, dtype=np.int32)
    grad = np.zeros((P, P), np.complex64)
    wsum = np.zeros((P, P), np.float32)
    tmax = np.zeros(emu.emu_gather_num_tiles(P, P), np.float32)
    emu.emu_reset()
    emu.emu_gather(S, cptr(ps), cptr(chi), cptr(pos), B, cptr(order), B, P, P, cptr(grad),
                   cptr(wsum), cptr(tmax))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    ref = po.scatter_bincount(np.conj(ps) * chi, pos, (S, S), (P, P))
    assert rel(grad, ref) < 1e-6
    inner = (slice(16, 80), slice(16, 80))
    assert rel(grad[inner] / wsum[inner], obj[inner]) < 1e-6


def _position_batch(emu, o, seed=4, B=12):
    """Fused pass of a batch, the oracle's update of the object, and the
    inputs of the position-correction phases."""
    S = o._amplitudes.shape[-1]
    rng = np.random.default_rng(seed)
    obj = (o._object * np.exp(1j * 0.3 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    probe = o._probe.astype(np.complex64)
    order = rng.permutation(o._num_diffraction_patterns)[:B].astype(np.int32)
    pos = o._positions_px.copy()
    k = kidx(emu, S)
    phat = np.ascontiguousarray(np.fft.fft2(probe)[np.ix_(k, k)].astype(np.complex64))
    amps = np.ascontiguousarray(o._amplitudes[:, k][:, :, k])
    ps = np.zeros((B, S, S), np.complex64)
    chi = np.zeros_like(ps)
    far = np.zeros_like(ps)
    err = np.zeros(B)
    emu.emu_fused(S, cptr(phat), cptr(obj), obj.shape[0], obj.shape[1], cptr(pos), len(pos),
                  cptr(amps), cptr(order), B, cptr(ps), cptr(chi), cptr(err), cptr(far))
    sp, op, ew, _ = o.forward(obj, probe, pos[order], o._amplitudes[order])
    new_obj, _ = o.adjoint(obj.copy(), probe.copy(), op, sp, pos[order], ew, 0.5, 1, True)
    return dict(S=S, order=order, pos=pos, amps=amps, ps=ps, far=far, sp=sp, op=op,
                obj=np.ascontiguousarray(new_obj.astype(np.complex64)))


def _emu_positions(emu, d, clip_update=0, max_update=0.0, clip_total=0, max_total=0.0,
                   initial=None, step=0.5):
    pos = d["pos"].copy()
    init = d["pos"].copy() if initial is None else initial
    B = len(d["order"])
    sums = np.zeros((B, 5))
    singular = np.zeros(1, np.int32)
    emu.emu_reset()
    rc = emu.emu_position_correction(
        d["S"], cptr(d["ps"]), cptr(d["far"]), cptr(d["obj"]), d["obj"].shape[0], d["obj"].shape[1],
        cptr(d["amps"]), cptr(pos), cptr(init), len(pos), cptr(d["order"]), B,
        ctypes.c_float(step), clip_update, ctypes.c_double(max_update), clip_total,
        ctypes.c_double(max_total), cptr(sums), cptr(singular))
    assert rc == 0
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    return pos, sums, int(singular[0])


def test_position_correction_phases_match_oracle(emu, frame_case):
    """Position-correction phases (two shifted-object FFTs, float64 sums,
    2x2 solve) against the oracle's _position_correction on one batch."""
    _, o = frame_case
    d = _position_batch(emu, o)
    order = d["order"]
    pos, sums, singular = _emu_positions(emu, d)
    assert singular == 0
    diff, pdx, pdy = o.position_partials(d["obj"], d["pos"][order], d["sp"], d["sp"] * d["op"],
                                         o._amplitudes[order])
    pdx, pdy, diff = (x.astype(np.float64) for x in (pdx, pdy, diff))
    ref_sums = np.stack([np.sum(pdx * pdx, 1), np.sum(pdx * pdy, 1), np.sum(pdy * pdy, 1),
                         np.sum(pdx * diff, 1), np.sum(pdy * diff, 1)], 1)
    assert rel(sums, ref_sums) < 1e-5
    ref = o.position_correction(d["obj"], d["pos"][order].copy(), d["sp"], d["sp"] * d["op"],
                                o._amplitudes[order], d["pos"][order], 0.5)
    upd, ref_upd = d["pos"][order] - pos[order], d["pos"][order] - ref
    assert np.max(np.abs(upd)) > 1e-3  # the step is not trivially zero
    assert rel(upd, ref_upd) < 1e-4
    untouched = np.setdiff1d(np.arange(len(pos)), order)
    np.testing.assert_array_equal(pos[untouched], d["pos"][untouched])


def test_position_correction_clips_match_oracle(emu, small_case):
    _, o = small_case
    d = _position_batch(emu, o, seed=7)
    order = d["order"]
    norm = float(np.hypot(*o.sampling))
    free, _, _ = _emu_positions(emu, d)
    steps = np.linalg.norm(d["pos"][order] - free[order], axis=1)
    lim = float(np.median(steps))  # about half the positions get clipped
    pos, _, _ = _emu_positions(emu, d, clip_update=1, max_update=lim)
    ref = o.position_correction(d["obj"], d["pos"][order].copy(), d["sp"], d["sp"] * d["op"],
                                o._amplitudes[order], d["pos"][order], 0.5,
                                max_position_update_distance=lim * norm)
    np.testing.assert_allclose(pos[order], ref, rtol=0, atol=1e-4 * max(lim, 1e-3))
    # step lengths read back from float32 positions near 100 px: +-2e-5 px
    assert np.max(np.linalg.norm(d["pos"][order] - pos[order], axis=1)) <= lim + 2e-5
    # total-distance limit: half of the batch starts 1 px from its initial position
    init = d["pos"].copy()
    init[order[::2], 0] -= 1.0
    pos, _, _ = _emu_positions(emu, d, clip_total=1, max_total=0.5, initial=init)
    ref = o.position_correction(d["obj"], d["pos"][order].copy(), d["sp"], d["sp"] * d["op"],
                                o._amplitudes[order], init[order], 0.5,
                                max_position_total_distance=0.5 * norm)
    np.testing.assert_array_equal(pos[order[::2]], d["pos"][order[::2]])
    np.testing.assert_allclose(pos[order], ref, rtol=0, atol=1e-4)


@pytest.mark.parametrize("abc,valid", [((-1.0, 1, 2.0), 1), ((1 - 2 * 0.6, 0.6, 2), 1),
                                       ((0, 1, 2), 0)])
def test_projection_phases_match_oracle(emu, frame_case, abc, valid):
    """Projection-set exit-wave step (DM, RAAR, SUPERFLIP's first iteration)
    against the oracle's _projection_sets_fourier_projection, race-free."""
    _, o = frame_case
    S = o._amplitudes.shape[-1]
    rng = np.random.default_rng(21)
    obj = (o._object * np.exp(1j * 0.2 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    probe = o._probe.astype(np.complex64)
    B = 10
    order = np.arange(B, dtype=np.int32) * 7
    pos = o._positions_px
    k = kidx(emu, S)
    phat = np.ascontiguousarray(np.fft.fft2(probe)[np.ix_(k, k)].astype(np.complex64))
    amps = np.ascontiguousarray(o._amplitudes[:, k][:, :, k])
    sp, op, _, _ = o.forward(obj, probe, pos[order], o._amplitudes[order])
    e0 = (sp * op + 0.1 * np.abs(sp * op).mean()
          * (rng.standard_normal(sp.shape) + 1j * rng.standard_normal(sp.shape))).astype(np.complex64)
    exits = e0.copy()
    ps = np.zeros_like(exits)
    err = np.zeros(B)
    emu.emu_reset()
    emu.emu_projection(S, cptr(phat), cptr(obj), obj.shape[0], obj.shape[1], cptr(pos), len(pos),
                       cptr(amps), cptr(order), B, cptr(exits), valid, ctypes.c_double(abc[0]),
                       ctypes.c_double(abc[1]), ctypes.c_double(abc[2]), cptr(ps), cptr(err))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    rsp, _, ref, e = o.projection_forward(obj, probe, pos[order], o._amplitudes[order],
                                          e0.copy() if valid else None, *abc)
    assert rel(ps, rsp) < 1e-6
    assert rel(exits, ref) < 2e-5
    assert abs(err.sum() / e - 1) < 1e-5


@pytest.mark.parametrize("K", [1, 3])
def test_mixed_state_phases_match_oracle(emu, frame_case, K):
    """Mixed-state phases (K probe modes, staged far fields, common amplitude
    factor) against the oracle's mixed forward, race-free."""
    _, o = frame_case
    S = o._amplitudes.shape[-1]
    rng = np.random.default_rng(31 + K)
    obj = (o._object * np.exp(1j * 0.2 * rng.standard_normal(o._object.shape))).astype(np.complex64)
    np.random.seed(5)
    probes = po.mixed_probe_modes(o._probe.astype(np.complex64), K)
    probes[1:] *= 0.6
    B = 8
    order = (np.arange(B, dtype=np.int32) * 11) % o._num_diffraction_patterns
    pos = o._positions_px
    k = kidx(emu, S)
    phat = np.ascontiguousarray(np.fft.fft2(probes)[:, k][:, :, k].astype(np.complex64))
    amps = np.ascontiguousarray(o._amplitudes[:, k][:, :, k])
    ps = np.zeros((K, B, S, S), np.complex64)
    chi = np.zeros_like(ps)
    far = np.zeros((B, K, S, S), np.complex64)
    err = np.zeros(B)
    emu.emu_reset()
    emu.emu_mixed(S, cptr(phat), K, cptr(obj), obj.shape[0], obj.shape[1], cptr(pos), len(pos),
                  cptr(amps), cptr(order), B, cptr(ps), cptr(chi), cptr(err), cptr(far))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    m = po.OracleMixedstatePtychography(energy=300e3, num_probes=K)
    m._region_of_interest_shape = np.array([S, S])
    sp, _, ref, e = m.forward(obj, probes, pos[order], o._amplitudes[order])
    assert rel(ps.transpose(1, 0, 2, 3), sp) < 1e-6
    assert rel(chi.transpose(1, 0, 2, 3), ref) < 2e-5
    assert abs(err.sum() / e - 1) < 1e-5


@pytest.mark.parametrize("S,shape", [(32, (101, 83)), (64, (71, 130))])
def test_scatter_ragged_object_and_wrapped_windows(emu, S, shape):
    """Object edges that cut the scatter tiles (sizes no multiple of the tile
    rows or columns) and windows that wrap around both object edges: the
    float64 sums must still equal np.bincount in batch order bit for bit."""
    rng = np.random.default_rng(S + shape[0])
    B = 300
    pos = np.stack([rng.uniform(-S, shape[0] + S, B),
                    rng.uniform(-S, shape[1] + S, B)], 1).astype(np.float32)
    order = rng.permutation(B).astype(np.int32)
    ps = (rng.standard_normal((B, S, S)) + 1j * rng.standard_normal((B, S, S))).astype(np.complex64)
    grad = np.zeros(shape, np.complex64)
    spsum = np.zeros(shape, np.float32)
    tmax = np.zeros(emu.emu_gather_num_tiles(*shape), np.float32)
    emu.emu_reset()
    emu.emu_gather(S, cptr(ps), cptr(np.ones_like(ps)), cptr(pos), B, cptr(order), B,
                   shape[0], shape[1], cptr(grad), cptr(spsum), cptr(tmax))
    assert emu.emu_violations() == 0, emu.emu_first_violation()
    np.testing.assert_array_equal(grad, po.scatter_bincount(np.conj(ps), pos[order], (S, S), shape))
    sq = (ps.real * ps.real + ps.imag * ps.imag).astype(np.float32)
    np.testing.assert_array_equal(spsum, po.scatter_bincount(sq, pos[order], (S, S), shape))
    assert tmax.max() == spsum.max()
