"""Pins the CPU oracle to what mathematics and hand-derived values fix.

The reference ships no tests, fixtures or golden vectors for this path, and
it cannot be imported here (emdfile, pylops and others are absent), so each
oracle function is checked against closed forms, invariants, hand-derived
vectors or brute-force loops, chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails.
"""

import numpy as np
import pytest
import scipy.ndimage

import oracle.ptycho_oracle as po
from tests.conftest import rel


def test_wavelength_textbook_values():
    # relativistic electron wavelengths (textbook values, 4-5 significant digits)
    assert po.electron_wavelength_angstrom(300e3) == pytest.approx(0.019687, rel=1e-4)
    assert po.electron_wavelength_angstrom(200e3) == pytest.approx(0.025079, rel=1e-4)
    assert po.electron_wavelength_angstrom(80e3) == pytest.approx(0.041757, rel=1e-4)
    # the non-relativistic value at 300 kV would be 0.02239 A: must differ
    assert abs(po.electron_wavelength_angstrom(300e3) - 0.02239) > 1e-3


def _probe(**kw):
    args = dict(energy=300e3, gpts=(64, 64), sampling=(0.2, 0.2), semiangle_cutoff=20.0,
                rolloff=2.0)
    args.update(kw)
    return po.ComplexProbeOracle(**args)


def test_aperture_closed_form():
    p = _probe()
    alpha = np.array([0.0, 0.010, 0.0179, 0.019, 0.0201, 0.030], dtype=np.float32)
    a = p.evaluate_aperture(alpha.copy())
    assert a[0] == 1 and a[1] == 1 and a[2] == 1      # below cutoff - rolloff
    assert a[4] == 0 and a[5] == 0                    # above cutoff
    assert a[3] == pytest.approx(0.5, abs=1e-5)       # middle of the cosine edge
    hard = _probe(rolloff=0.0).evaluate_aperture(alpha.copy())
    np.testing.assert_array_equal(hard, [1, 1, 1, 1, 0, 0])


def test_chi_defocus_astigmatism_cs_closed_forms():
    lam = po.electron_wavelength_angstrom(300e3)
    alpha = np.linspace(0, 0.025, 7).astype(np.float32)[:, None] * np.ones((1, 5), np.float32)
    phi = np.linspace(-3, 3, 5).astype(np.float32)[None, :] * np.ones((7, 1), np.float32)
    chi = _probe(parameters={"defocus": 200.0}).evaluate_chi(alpha, phi)
    # defocus alias flips sign: C10 = -200
    np.testing.assert_allclose(chi, np.pi * (-200.0) * alpha.astype(np.float64) ** 2 / lam,
                               rtol=1e-5, atol=1e-5)
    chi = _probe(parameters={"C12": 50.0, "phi12": 0.3}).evaluate_chi(alpha, phi)
    np.testing.assert_allclose(chi, np.pi / lam * alpha.astype(np.float64) ** 2 * 50.0
                               * np.cos(2 * (phi - 0.3)), rtol=1e-5, atol=1e-5)
    chi = _probe(parameters={"Cs": 1e4}).evaluate_chi(alpha, phi)
    np.testing.assert_allclose(chi, 2 * np.pi / lam * 0.25 * alpha.astype(np.float64) ** 4 * 1e4,
                               rtol=1e-5, atol=1e-5)


def test_probe_build_unit_norm_aperture_and_centre():
    p = _probe().build()._array
    assert np.sum(np.abs(p) ** 2) == pytest.approx(1.0, rel=1e-5)
    # no aberrations: the probe peaks at the corner pixel (corner-centred)
    assert np.unravel_index(np.argmax(np.abs(p)), p.shape) == (0, 0)
    alpha, _ = _probe().get_scattering_angles()
    ap = _probe().evaluate_aperture(alpha.copy())
    spectrum = np.abs(np.fft.fft2(p))
    np.testing.assert_allclose(spectrum / spectrum.max(), ap / ap.max(), atol=1e-5)


def test_fft_shift_integer_is_roll_and_composes():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))).astype(np.complex64)
    np.testing.assert_allclose(po.fft_shift(a, np.array([3.0, -5.0])),
                               np.roll(a, (3, -5), axis=(0, 1)), atol=2e-5)
    s1 = po.fft_shift(po.fft_shift(a, np.array([0.3, 0.7])), np.array([0.4, -0.2]))
    s2 = po.fft_shift(a, np.array([0.7, 0.5]))
    np.testing.assert_allclose(s1, s2, atol=5e-5)
    # +1 moves a delta from index 0 to index 1 (np.roll direction)
    d = np.zeros((8, 8), np.complex64)
    d[0, 0] = 1
    out = po.fft_shift(d, np.array([1.0, 0.0]))
    assert np.argmax(np.abs(out)) == 8
    # batched operator: one frame per position
    op = po.fourier_translation_operator(np.array([[1.0, 0.0], [0.0, 2.0]]), (8, 8))
    assert op.shape == (2, 8, 8)


def test_subdivide_into_batches_hand_derived():
    assert po.subdivide_into_batches(10, max_batch=3) == [3, 3, 2, 2]
    assert po.subdivide_into_batches(16384, max_batch=1024) == [1024] * 16
    assert po.subdivide_into_batches(1000, max_batch=1024) == [1000]
    assert po.subdivide_into_batches(1025, max_batch=1024) == [513, 512]
    assert list(po.generate_batches(7, max_batch=3)) == [(0, 3), (3, 5), (5, 7)]
    with pytest.raises(RuntimeError):
        po.subdivide_into_batches(5, num_batches=2, max_batch=2)


def _brute_indices(positions, S, shape):
    B = len(positions)
    rows = np.zeros((B, S), int)
    cols = np.zeros((B, S), int)
    for b, (x, y) in enumerate(positions):
        x0, y0 = int(np.round(x)), int(np.round(y))  # numpy rounds half to even
        for i in range(S):
            f = i if i < S // 2 else i - S
            rows[b, i] = (x0 + f) % shape[0]
            cols[b, i] = (y0 + f) % shape[1]
    return rows, cols


def test_patch_indices_brute_force_round_half_even():
    pos = np.array([[2.5, 3.5], [0.2, -1.6], [9.5, 10.49], [-3.5, 12.5]], dtype=np.float32)
    S, shape = 8, (11, 13)
    r, c = po.patch_indices(pos, (S, S), shape)
    br, bc = _brute_indices(pos, S, shape)
    np.testing.assert_array_equal(r[:, :, 0], br)
    np.testing.assert_array_equal(c[:, 0, :], bc)
    assert r[0, 0, 0] == 2 and c[0, 0, 0] == 4  # 2.5 -> 2, 3.5 -> 4


def test_scatter_bincount_brute_force_float64_order():
    rng = np.random.default_rng(1)
    S, shape = 8, (12, 10)
    pos = rng.uniform(-2, 14, (6, 2)).astype(np.float32)
    patches = (rng.standard_normal((6, S, S)) + 1j * rng.standard_normal((6, S, S))).astype(
        np.complex64)
    out = po.scatter_bincount(patches, pos, (S, S), shape)
    br, bc = _brute_indices(pos, S, shape)
    re = np.zeros(shape)
    im = np.zeros(shape)
    for b in range(6):
        for i in range(S):
            for j in range(S):
                re[br[b, i], bc[b, j]] += float(patches[b, i, j].real)
                im[br[b, i], bc[b, j]] += float(patches[b, i, j].imag)
    np.testing.assert_array_equal(out.real, re.astype(np.float32))
    np.testing.assert_array_equal(out.imag, im.astype(np.float32))


def test_get_shifted_ar_integer_and_linear():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((16, 16)).astype(np.float32)
    np.testing.assert_array_equal(po.get_shifted_ar(a, 3.0, -2.0), np.roll(a, (3, -2), (0, 1)))
    ramp = np.tile(np.arange(16, dtype=np.float32)[:, None], (1, 16))
    out = po.get_shifted_ar(ramp, 0.5, 0.0)
    # interior rows: (1-w) f(i) + w f(i-1) = i - 0.5
    np.testing.assert_allclose(out[1:, 0], np.arange(1, 16) - 0.5, atol=1e-6)


def test_get_com_closed_form():
    a = np.zeros((9, 7))
    a[3, 5] = 2.0
    assert po.get_CoM(a) == (3.0, 5.0)
    b = np.zeros((8, 8))
    b[7, 1] = 1.0  # corner-centred coordinate of index 7 is -1
    assert po.get_CoM(b, corner_centered=True) == (-1.0, 1.0)


def test_fit_plane_recovers_noiseless_plane():
    x, y = np.linspace(0, 1, 12), np.linspace(0, 1, 9)
    ry, rx = np.meshgrid(y, x)
    plane = 3.0 * rx - 2.0 * ry + 1.5
    np.testing.assert_allclose(po.fit_plane(plane), plane, atol=1e-10)
    mask = np.ones_like(plane, bool)
    mask[::3] = False
    corrupted = plane.copy()
    corrupted[~mask] = 99.0
    np.testing.assert_allclose(po.fit_plane(corrupted, mask), plane, atol=1e-10)


@pytest.fixture(scope="module")
def tiny():
    """A small self-consistent problem: 16 px frames on a 40^2 object."""
    rng = np.random.default_rng(5)
    o = po.OracleSingleslicePtychography(energy=300e3)
    S = 16
    o._region_of_interest_shape = np.array([S, S])
    probe = po.ComplexProbeOracle(300e3, (S, S), (0.4, 0.4), 20.0, 2.0,
                                  parameters={"defocus": 50.0}).build()._array * 30
    obj = np.exp(1j * 0.3 * scipy.ndimage.gaussian_filter(rng.standard_normal((40, 40)), 2)
                 / 0.1).astype(np.complex64)
    pos = rng.uniform(8, 32, (12, 2)).astype(np.float32)
    return o, probe.astype(np.complex64), obj, pos


def test_forward_parseval_and_fixed_point(tiny):
    o, probe, obj, pos = tiny
    frac = pos - np.round(pos)
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    psi = po.fft_shift(probe, frac) * obj[rows, cols]
    F = np.fft.fft2(psi.astype(np.complex128))
    np.testing.assert_allclose(np.sum(np.abs(F) ** 2), 256 * np.sum(np.abs(psi) ** 2), rtol=1e-5)
    # measured amplitudes equal to the model: chi = 0 and err = 0
    amps = np.abs(F).astype(np.float32)
    _, _, chi, err = o.forward(obj, probe, pos, amps)
    assert np.max(np.abs(chi)) < 1e-4 * np.max(np.abs(psi))
    assert err < 1e-6 * np.sum(amps ** 2)
    # zero amplitudes: chi = ifft2(-F) = -psi
    _, _, chi0, err0 = o.forward(obj, probe, pos, np.zeros_like(amps))
    np.testing.assert_allclose(chi0, -psi, atol=1e-5)
    assert err0 == pytest.approx(np.sum(np.abs(F) ** 2), rel=1e-4)
    # zero far field: angle(0) = 0 so chi = ifft2(A)
    _, _, chiz, _ = o.forward(np.zeros_like(obj), probe, pos, amps)
    np.testing.assert_allclose(chiz, np.fft.ifft2(amps), atol=1e-5)


def test_gradient_is_scaled_adjoint_term_fp64_finite_differences(tiny):
    """dL/dRe O + i dL/dIm O = -2 S^2 scatter(conj(P_shift) chi) for
    L = sum (|F| - A)^2, checked by float64 central differences."""
    o, probe, obj, pos = tiny
    o = po.OracleSingleslicePtychography(energy=300e3, mode="c128")
    o._region_of_interest_shape = np.array([16, 16])
    rng = np.random.default_rng(9)
    obj64 = obj.astype(np.complex128)
    probe64 = probe.astype(np.complex128)
    amps = np.abs(np.fft.fft2(po.fft_shift(probe64, pos - np.round(pos), "c128")
                              * 1.05)).astype(np.float64) * rng.uniform(0.8, 1.2, (12, 16, 16))

    def loss(objv):
        frac = pos - np.round(pos)
        rows, cols = po.patch_indices(pos, (16, 16), objv.shape)
        F = np.fft.fft2(po.fft_shift(probe64, frac, "c128") * objv[rows, cols])
        return np.sum((np.abs(F) - amps) ** 2)

    sp, _, chi, _ = o.forward(obj64, probe64, pos, amps)
    x0 = np.round(pos[:, 0]).astype(int)
    y0 = np.round(pos[:, 1]).astype(int)
    idx = np.fft.fftfreq(16, 1 / 16).astype(int)
    grad = np.zeros(obj.shape, np.complex128)
    for b in range(len(pos)):
        grad[np.ix_((x0[b] + idx) % 40, (y0[b] + idx) % 40)] += np.conj(sp[b]) * chi[b]
    grad *= -2 * 256
    h = 1e-6
    for (i, j) in [(20, 20), (15, 27), (25, 12)]:
        for unit, comp in ((1.0, grad[i, j].real), (1j, grad[i, j].imag)):
            e = np.zeros_like(obj64)
            e[i, j] = unit * h
            fd = (loss(obj64 + e) - loss(obj64 - e)) / (2 * h)
            assert fd == pytest.approx(comp, rel=1e-5, abs=1e-6 * np.abs(grad).max())


def test_noise_free_truth_is_fixed_point_of_update(tiny):
    o, probe, obj, pos = tiny
    pos = np.round(pos).astype(np.float32)
    o._region_of_interest_shape = np.array([16, 16])
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    amps = np.abs(np.fft.fft2((probe[None] * obj[rows, cols]).astype(np.complex128))).astype(
        np.float32)
    obj2, probe2 = obj.copy(), probe.copy()
    sp, op, chi, err = o.forward(obj2, probe2, pos, amps)
    o.adjoint(obj2, probe2, op, sp, pos, chi, 0.5, 1, False)
    assert rel(obj2, obj) < 1e-5 and rel(probe2, probe) < 1e-5


def test_object_constraints_invariants():
    rng = np.random.default_rng(3)
    obj = (1.3 * np.exp(1j * (0.4 + 0.1 * rng.standard_normal((20, 20))))).astype(np.complex64)
    obj[:5] *= 0.5
    mask = np.zeros((20, 20), bool)
    mask[:, :4] = True
    out = po.OracleSingleslicePtychography.object_constraints(obj.copy(), mask)
    assert abs(np.angle(out)[mask].mean()) < 1e-6
    assert np.abs(out).max() <= 1.0 + 1e-6
    np.testing.assert_allclose(np.abs(out[:5]), np.minimum(np.abs(obj[:5]), 1), rtol=1e-6)
    pure = po.OracleSingleslicePtychography.object_constraints(obj.copy(), None,
                                                               pure_phase_object=True)
    np.testing.assert_allclose(np.abs(pure), 1.0, rtol=1e-6)


def test_probe_com_constraint_recentres():
    o = po.OracleSingleslicePtychography(energy=300e3)
    p = po.ComplexProbeOracle(300e3, (32, 32), (0.3, 0.3), 25.0, 2.0).build()._array
    p = po.fft_shift(p.astype(np.complex64), np.array([1.3, -0.7]))
    x0, y0 = po.get_CoM(np.abs(p) ** 2, corner_centered=True)
    assert abs(x0 - 1.3) < 0.05 and abs(y0 + 0.7) < 0.05
    q = o.probe_com_constraint(p)
    x1, y1 = po.get_CoM(np.abs(q) ** 2, corner_centered=True)
    # the corner-centred CoM of a periodic frame is not exactly shift
    # equivariant (tails wrap), so one recentring leaves a small residual
    assert abs(x1) < 0.05 and abs(y1) < 0.05
    assert np.hypot(x1, y1) < np.hypot(x0, y0) / 20


def test_probe_gradient_true_adjoint_fp64_finite_differences(tiny):
    """The probe gradient including the adjoint of the sub-pixel shift,
    checked by float64 central differences; and it reduces to the
    reference's sum_b conj(O_b) chi_b term for integer positions."""
    _, probe, obj, pos = tiny
    o = po.OracleSingleslicePtychography(energy=300e3, mode="c128")
    o._region_of_interest_shape = np.array([16, 16])
    rng = np.random.default_rng(10)
    probe64 = probe.astype(np.complex128)
    obj64 = obj.astype(np.complex128)
    amps = rng.uniform(0.5, 2.0, (12, 16, 16)) * np.abs(np.fft.fft2(probe64)).max() / 4
    _, g_obj, g_probe = o.loss_gradients(obj64, probe64, pos, amps)

    def loss(pv):
        frac = pos - np.round(pos)
        rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
        F = np.fft.fft2(po.fft_shift(pv, frac, "c128") * obj64[rows, cols])
        return np.sum((np.abs(F) - amps) ** 2)

    h = 1e-6
    for (i, j) in [(0, 0), (1, 15), (5, 3)]:
        for unit, comp in ((1.0, g_probe[i, j].real), (1j, g_probe[i, j].imag)):
            e = np.zeros_like(probe64)
            e[i, j] = unit * h
            fd = (loss(probe64 + e) - loss(probe64 - e)) / (2 * h)
            assert fd == pytest.approx(comp, rel=1e-5, abs=1e-6 * np.abs(g_probe).max())
    ipos = np.round(pos).astype(np.float32)
    _, _, g_int = o.loss_gradients(obj64, probe64, ipos, amps)
    sp, op, chi, _ = o.forward(obj64, probe64, ipos, amps)
    np.testing.assert_allclose(g_int, -2 * 256 * np.sum(np.conj(op) * chi, axis=0), rtol=1e-10,
                               atol=1e-10 * np.abs(g_int).max())


# ---------------------------------------------------------------------------
# position correction and position constraints
# ---------------------------------------------------------------------------
def _pos_problem(mode="c64", S=16, seed=21):
    """Tiny problem with calibrations set: S px frames at 0.4 A on a 48^2
    object, smooth phase object with a little amplitude contrast."""
    rng = np.random.default_rng(seed)
    o = po.OracleSingleslicePtychography(energy=300e3, mode=mode)
    o._region_of_interest_shape = np.array([S, S])
    o._amplitudes_shape = np.array([S, S])
    lam = po.electron_wavelength_angstrom(300e3)
    dk = lam * 1e3 / (0.4 * S)  # mrad per pixel giving 0.4 A real-space sampling
    o._angular_sampling = (dk, dk)
    probe = po.ComplexProbeOracle(300e3, (S, S), (0.4, 0.4), 20.0, 2.0,
                                  parameters={"defocus": 40.0}).build()._array * 30
    phase = scipy.ndimage.gaussian_filter(rng.standard_normal((48, 48)), 1.5)
    amp = 1 - 0.1 * scipy.ndimage.gaussian_filter(rng.random((48, 48)), 2)
    obj = (amp * np.exp(1j * phase / phase.std())).astype(np.complex64)
    pos = rng.uniform(12, 36, (10, 2)).astype(np.float32)
    return o, probe.astype(np.complex64), obj, pos


def _dft2_loops(x):
    """Plain DFT matrices in float64 (no FFT library)."""
    S = x.shape[-1]
    k = np.arange(S)
    W = np.exp(-2j * np.pi * np.outer(k, k) / S)
    return W @ x.astype(np.complex128) @ W.T


def test_position_partials_brute_force_dft_and_rolled_object():
    o, probe, obj, pos = _pos_problem(mode="c128", S=8)
    S = 8
    sp = po.fft_shift(probe[:S, :S] if probe.shape[0] != S else probe, pos - np.round(pos), "c128")
    amps = np.random.default_rng(3).random((len(pos), S, S)).astype(np.float32)
    rows, cols = _brute_indices(pos, S, obj.shape)
    overlap = np.stack([sp[b] * obj[np.ix_(rows[b], cols[b])] for b in range(len(pos))])
    d, pdx, pdy = o.position_partials(obj, pos, sp, overlap, amps)
    down = np.roll(obj, -1, axis=0)   # down[r, c] = obj[r + 1, c]
    right = np.roll(obj, -1, axis=1)  # right[r, c] = obj[r, c + 1]
    for b in range(len(pos)):
        F = _dft2_loops(overlap[b])
        Fx = _dft2_loops(sp[b] * down[np.ix_(rows[b], cols[b])])
        Fy = _dft2_loops(sp[b] * right[np.ix_(rows[b], cols[b])])
        ref_d = amps[b].astype(np.float64) ** 2 - np.abs(F) ** 2
        ref_x = 2 * np.real((F - Fx) * np.conj(F))
        ref_y = 2 * np.real((F - Fy) * np.conj(F))
        assert rel(d[b], ref_d.ravel()) < 1e-6
        assert rel(pdx[b], ref_x.ravel()) < 1e-6
        assert rel(pdy[b], ref_y.ravel()) < 1e-6
        # the two axes are not interchangeable on a random object
        assert rel(pdx[b], ref_y.ravel()) > 1e-2


def test_position_correction_recovers_planted_least_squares_step():
    """Measured intensities |F|^2 + alpha p_x + beta p_y: the least-squares
    step is exactly (alpha, beta) per position, scaled by the step size and
    subtracted from the positions."""
    o, probe, obj, pos = _pos_problem(mode="c128")
    S = 16
    sp = po.fft_shift(probe, pos - np.round(pos), "c128")
    rows, cols = po.patch_indices(pos, (S, S), obj.shape)
    overlap = sp * obj[rows, cols]
    _, pdx, pdy = o.position_partials(obj, pos, sp, overlap, np.zeros((len(pos), S, S)))
    rng = np.random.default_rng(4)
    ab = rng.uniform(-1, 1, (len(pos), 2)) * 1e-3
    F = np.fft.fft2(overlap.astype(np.complex128))
    I = np.abs(F).reshape(len(pos), -1) ** 2 + ab[:, :1] * pdx + ab[:, 1:] * pdy
    assert I.min() > 0
    amps = np.sqrt(I).reshape(len(pos), S, S)  # float64 amplitudes keep the plant exact
    out = o.position_correction(obj, pos.astype(np.float64), sp, overlap, amps, pos, 0.5)
    np.testing.assert_allclose(out, pos - 0.5 * ab, rtol=0, atol=1e-9)


def test_position_correction_clipping_closed_forms():
    o, probe, obj, pos = _pos_problem(mode="c128")
    S = 16
    sp = po.fft_shift(probe, pos - np.round(pos), "c128")
    rows, cols = po.patch_indices(pos, (S, S), obj.shape)
    overlap = sp * obj[rows, cols]
    _, pdx, pdy = o.position_partials(obj, pos, sp, overlap, np.zeros((len(pos), S, S)))
    ab = np.tile([[3e-3, -4e-3]], (len(pos), 1))  # |step| = 0.5 * 5e-3
    F = np.fft.fft2(overlap.astype(np.complex128))
    I = np.abs(F).reshape(len(pos), -1) ** 2 + ab[:, :1] * pdx + ab[:, 1:] * pdy
    I = np.maximum(I, 0)
    amps = np.sqrt(I).reshape(len(pos), S, S)
    ref = pos.astype(np.float64) - 0.5 * ab
    norm = np.hypot(*o.sampling)
    # update clip: max distance in A divided by |sampling| (pixels)
    out = o.position_correction(obj, pos.astype(np.float64), sp, overlap, amps, pos, 0.5,
                                max_position_update_distance=1e-3 * norm)
    step = pos - out
    np.testing.assert_allclose(np.linalg.norm(step, axis=1), 1e-3, rtol=1e-5)
    np.testing.assert_allclose(step / np.linalg.norm(step, axis=1)[:, None],
                               ab / np.linalg.norm(ab, axis=1)[:, None], atol=1e-3)
    # a generous clip changes nothing
    out = o.position_correction(obj, pos.astype(np.float64), sp, overlap, amps, pos, 0.5,
                                max_position_update_distance=1.0 * norm)
    np.testing.assert_allclose(out, ref, atol=1e-6)
    # total-distance limit: positions already 0.5 px from their initial
    # values move no further; those at the initial positions do move
    init = pos.astype(np.float64).copy()
    init[::2, 0] -= 0.5
    out = o.position_correction(obj, pos.astype(np.float64), sp, overlap, amps, init, 0.5,
                                max_position_total_distance=0.1 * norm)
    np.testing.assert_array_equal(out[::2], pos[::2].astype(np.float64))
    np.testing.assert_allclose(out[1::2], ref[1::2], atol=1e-6)


def test_position_correction_moves_displaced_positions_towards_truth():
    o, probe, obj, pos = _pos_problem(mode="c128", seed=8)
    S = 16
    rng = np.random.default_rng(9)
    true = pos.astype(np.float32)
    rows, cols = po.patch_indices(true, (S, S), obj.shape)
    amps = np.abs(np.fft.fft2(po.fft_shift(probe, true - np.round(true), "c128")
                              * obj[rows, cols])).astype(np.float32)
    shift = rng.choice([-0.3, 0.3], (len(pos), 2)).astype(np.float32)
    cur = true + shift
    sp = po.fft_shift(probe, cur - np.round(cur), "c128")
    rows, cols = po.patch_indices(cur, (S, S), obj.shape)
    out = o.position_correction(obj, cur.copy(), sp, sp * obj[rows, cols], amps, cur, 1.0)
    before = np.abs(cur - true)
    after = np.abs(out - true)
    assert np.mean(after) < 0.5 * np.mean(before)
    assert np.mean(np.sign(out - cur) == -np.sign(shift)) > 0.9


def test_positions_centre_of_mass_constraint():
    rng = np.random.default_rng(2)
    init = rng.uniform(10, 50, (40, 2)).astype(np.float32)
    cur = init + rng.normal(0, 0.5, init.shape).astype(np.float32) + np.float32(1.5)
    # flag as passed by the loop: False applies the centre-of-mass step
    out = po.OracleSingleslicePtychography.positions_constraints(cur.copy(), init, False, False,
                                                                 False)
    np.testing.assert_allclose(out.mean(0), init.mean(0), atol=1e-4)
    np.testing.assert_allclose(out - out.mean(0), cur - cur.mean(0), atol=1e-4)
    same = po.OracleSingleslicePtychography.positions_constraints(cur.copy(), init, False, True,
                                                                  False)
    np.testing.assert_array_equal(same, cur)
    fixed = po.OracleSingleslicePtychography.positions_constraints(cur.copy(), init, True, False,
                                                                   True)
    np.testing.assert_array_equal(fixed, cur)


def _planted_affine():
    g = np.arange(20, dtype=np.float32) * 4
    gx, gy = np.meshgrid(g, g, indexing="ij")
    p0 = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32) + 30
    tf = (1.02, 0.97, 0.015, 0.03, np.float32(0.6), np.float32(-0.4))
    origin = p0.mean(0)
    return p0, tf, origin


def test_affine_fit_recovers_planted_map_exactly():
    p0, tf, origin = _planted_affine()
    p1 = po.affine_apply(tf, p0, origin)
    # hand-derived matrix rows: s0 (cos, -sin); s1 (h cos + sin, cos - h sin)
    c, s = np.cos(0.03), np.sin(0.03)
    np.testing.assert_allclose(po.affine_matrix(tf),
                               [[1.02 * c, -1.02 * s],
                                [0.97 * (0.015 * c + s), 0.97 * (c - 0.015 * s)]], rtol=1e-6)
    est, err = po.affine_fit(p0, p1, origin)
    np.testing.assert_allclose(est[:4], tf[:4], atol=2e-5)
    np.testing.assert_allclose([est[4], est[5]], [0.6, -0.4], atol=2e-3)
    assert err < 0.05
    # degenerate input: identity (the reference catches the singular solve)
    est, _ = po.affine_fit(p0[:1], p1[:1], origin)
    assert est == po.AFFINE_IDENTITY


def test_affine_ransac_ignores_outliers():
    p0, tf, origin = _planted_affine()
    p1 = po.affine_apply(tf, p0, origin).astype(np.float32)
    rng = np.random.default_rng(6)
    bad = rng.choice(len(p0), len(p0) // 10, replace=False)
    p1[bad] += rng.uniform(30, 60, (len(bad), 2)).astype(np.float32)
    np.random.seed(0)
    est, fitness = po.affine_ransac(p0, p1, origin, min_sample=len(p0) // 10)
    np.testing.assert_allclose(est[:4], tf[:4], atol=1e-4)
    assert np.isfinite(fitness)
    # the plain least-squares fit is pulled off by the outliers
    est_ls, _ = po.affine_fit(p0, p1, origin)
    assert abs(est_ls[0] - tf[0]) > 1e-3 or abs(est_ls[4] - tf[4]) > 1.0


# ---------------------------------------------------------------------------
# projection sets
# ---------------------------------------------------------------------------
def test_projection_method_parameters_table():
    mp = po.OracleSingleslicePtychography.method_parameters
    assert mp("DM_AP", 0.3) == (True, -0.3, 1, 1.3)
    assert mp("difference-map_alternating-projections", 0.0) == (True, -0.0, 1, 1.0)
    assert mp("RAAR", 0.4) == (True, 1 - 0.8, 0.4, 2)
    assert mp("RRR", 1.5) == (True, -1.5, 1.5, 2)
    assert mp("SUPERFLIP") == (True, 0, 1, 2)
    assert mp("generalized-projections", None, 0.1, 0.2, 0.3) == (True, 0.1, 0.2, 0.3)
    assert mp("GD") == (False, None, None, None)
    for bad in (("DM_AP", 1.2), ("RAAR", -0.1), ("RRR", 2.5)):
        with pytest.raises(ValueError):
            mp(*bad)
    with pytest.raises(ValueError):
        mp("generalized-projections", None, 0.1, None, 0.3)


def _projector(psi, A):
    """Modulus projector with plain DFT matrices (float64): ifft2(A F/|F|)."""
    F = _dft2_loops(psi)
    S = psi.shape[-1]
    k = np.arange(S)
    Winv = np.exp(2j * np.pi * np.outer(k, k) / S) / S
    G = A * np.exp(1j * np.angle(F))
    return Winv @ G @ Winv.T


def _proj_case(tiny):
    o, probe, obj, pos = tiny
    rng = np.random.default_rng(12)
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    psi = po.fft_shift(probe, pos - np.round(pos)) * obj[rows, cols]
    A = (np.abs(np.fft.fft2(psi)) * (1 + 0.2 * rng.random(psi.shape))).astype(np.float32)
    e = (psi + 0.3 * (rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape))
         * np.abs(psi).mean()).astype(np.complex64)
    return o, probe, obj, pos, psi, A, e


def test_projection_ap_meets_the_modulus_constraint(tiny):
    o, probe, obj, pos, psi, A, e = _proj_case(tiny)
    _, _, ew, err = o.projection_forward(obj, probe, pos, A, e, 0, 1, 1)  # DM_AP(0) = AP
    np.testing.assert_allclose(np.abs(np.fft.fft2(ew.astype(np.complex128))), A,
                               rtol=2e-5, atol=2e-5 * A.max())
    for b in range(len(pos)):
        np.testing.assert_allclose(ew[b], _projector(psi[b], A[b]), rtol=0,
                                   atol=1e-5 * np.abs(psi).max())
    assert err == pytest.approx(np.sum((A - np.abs(np.fft.fft2(psi.astype(np.complex128)))) ** 2),
                                rel=1e-4)


def test_projection_dm_is_the_difference_map(tiny):
    o, probe, obj, pos, psi, A, e = _proj_case(tiny)
    _, _, ew, _ = o.projection_forward(obj, probe, pos, A, e, -1.0, 1, 2.0)  # DM_AP(1)
    for b in range(len(pos)):
        ref = e[b] - psi[b] + _projector(2 * psi[b] - e[b], A[b])
        np.testing.assert_allclose(ew[b], ref, rtol=0, atol=2e-5 * np.abs(psi).max())
    # first iteration (no exit waves yet): e = psi
    _, _, ew0, _ = o.projection_forward(obj, probe, pos, A, None, -1.0, 1, 2.0)
    for b in range(len(pos)):
        np.testing.assert_allclose(ew0[b], _projector(psi[b], A[b]), rtol=0,
                                   atol=2e-5 * np.abs(psi).max())


@pytest.mark.parametrize("abc", [(1 - 2 * 0.7, 0.7, 2), (-1.2, 1.2, 2), (0.3, 0.6, 1.7), (0, 1, 2)])
def test_projection_consistent_exit_wave_is_fixed_point(tiny, abc):
    o, probe, obj, pos, psi, _, _ = _proj_case(tiny)
    A = np.abs(np.fft.fft2(psi.astype(np.complex128))).astype(np.float32)
    _, _, ew, err = o.projection_forward(obj, probe, pos, A, psi.copy(), *abc)
    np.testing.assert_allclose(ew, psi, rtol=0, atol=1e-5 * np.abs(psi).max())
    assert err < 1e-8 * np.sum(A.astype(np.float64) ** 2)


def test_projection_adjoint_reproduces_consistent_truth(tiny):
    o, probe, obj, pos = tiny
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    sp = po.fft_shift(probe, pos - np.round(pos))
    op = obj[rows, cols]
    new_obj, new_probe = o.projection_adjoint(obj.copy(), probe.copy(), op, sp, pos, sp * op, 0.0,
                                              False)
    cover = po.scatter_bincount(np.abs(sp) ** 2, pos, (16, 16), obj.shape) > 1e-3
    np.testing.assert_allclose(new_obj[cover], obj[cover], rtol=2e-5, atol=2e-5)
    # the probe term sums conj(O_b) O_b P_shift_b over positions; with all
    # fractions zero it is the probe itself
    pos0 = np.round(pos)
    sp0 = po.fft_shift(probe, pos0 - np.round(pos0))
    rows, cols = po.patch_indices(pos0, (16, 16), obj.shape)
    op0 = obj[rows, cols]
    _, new_probe = o.projection_adjoint(obj.copy(), probe.copy(), op0, sp0, pos0, sp0 * op0, 0.0,
                                        False)
    np.testing.assert_allclose(new_probe, probe, rtol=0, atol=2e-5 * np.abs(probe).max())


# ---------------------------------------------------------------------------
# mixed-state probes
# ---------------------------------------------------------------------------
def test_mixed_probe_modes_are_phase_ramped_copies():
    probe = po.ComplexProbeOracle(300e3, (16, 16), (0.4, 0.4), 20.0, 2.0,
                                  parameters={"defocus": 50.0}).build()._array.astype(np.complex64)
    np.random.seed(3)
    modes = po.mixed_probe_modes(probe, 3)
    np.random.seed(3)
    r = [np.random.rand() for _ in range(4)]
    np.testing.assert_array_equal(modes[0], probe)
    for i in (1, 2):
        np.testing.assert_allclose(np.abs(modes[i]), np.abs(probe), rtol=1e-5, atol=1e-7)
        fx, fy = r[2 * (i - 1)] - 0.5, r[2 * (i - 1) + 1] - 0.5
        f = np.fft.fftfreq(16)
        ramp = np.exp(-2j * np.pi * (fx * f[:, None] + fy * f[None, :]))
        np.testing.assert_allclose(modes[i], modes[i - 1] * ramp, rtol=0, atol=1e-5 * np.abs(probe).max())


def test_probe_orthogonalization_invariants():
    rng = np.random.default_rng(8)
    p = (rng.standard_normal((3, 16, 16)) + 1j * rng.standard_normal((3, 16, 16))).astype(np.complex64)
    p[1] += 0.8 * p[0]
    p[2] *= 0.3
    q = po.probe_orthogonalization(p)
    G = np.einsum("ixy,jxy->ij", q.conj().astype(np.complex128), q.astype(np.complex128))
    off = np.abs(G - np.diag(np.diag(G))).max()
    assert off < 1e-4 * np.abs(np.diag(G)).max()
    inten = np.real(np.diag(G))
    assert np.all(np.diff(inten) <= 0)  # largest first
    np.testing.assert_allclose(inten.sum(), np.sum(np.abs(p.astype(np.complex128)) ** 2), rtol=1e-5)
    # same span: projecting the input onto the output modes loses nothing
    Q = q.reshape(3, -1).astype(np.complex128).T
    P = p.reshape(3, -1).astype(np.complex128).T
    coef, *_ = np.linalg.lstsq(Q, P, rcond=None)
    assert np.linalg.norm(Q @ coef - P) < 1e-5 * np.linalg.norm(P)


@pytest.fixture(scope="module")
def mixed_tiny(tiny):
    o, probe, obj, pos = tiny
    m = po.OracleMixedstatePtychography(energy=300e3, num_probes=2)
    m._region_of_interest_shape = np.array([16, 16])
    m._mode = "c64"
    rng = np.random.default_rng(2)
    second = (0.5 * np.roll(probe, 1, axis=0) * np.exp(1j * rng.uniform(0, 6.28))).astype(np.complex64)
    return m, o, np.stack([probe, second]), obj, pos


def test_mixed_forward_single_mode_equals_single_slice(mixed_tiny):
    m, o, probes, obj, pos = mixed_tiny
    rng = np.random.default_rng(4)
    A = rng.uniform(0, 60, (len(pos), 16, 16)).astype(np.float32)
    sp1, op1, chi1, e1 = o.forward(obj, probes[0], pos, A)
    sp, op, chi, e = m.forward(obj, probes[:1], pos, A)
    np.testing.assert_allclose(sp[:, 0], sp1, rtol=0, atol=1e-6 * np.abs(sp1).max())
    np.testing.assert_allclose(chi[:, 0], chi1, rtol=0, atol=1e-5 * np.abs(chi1).max())
    assert e == pytest.approx(e1, rel=1e-5)


def test_mixed_forward_fixed_point_and_incoherent_sum(mixed_tiny):
    m, _, probes, obj, pos = mixed_tiny
    frac = pos - np.round(pos)
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    I = 0
    for k in range(2):
        I = I + np.abs(np.fft.fft2((po.fft_shift(probes[k], frac) * obj[rows, cols])
                                   .astype(np.complex128))) ** 2
    A = np.sqrt(I).astype(np.float32)
    _, _, chi, e = m.forward(obj, probes, pos, A)
    assert np.max(np.abs(chi)) < 1e-4 * np.max(np.abs(probes))
    assert e < 1e-8 * np.sum(A.astype(np.float64) ** 2)
    # one mode alone does not explain the data
    _, _, chi1, e1 = m.forward(obj, probes[:1], pos, A)
    assert e1 > 1e-3 * np.sum(A.astype(np.float64) ** 2)


def test_mixed_object_term_is_loss_gradient_fp64(mixed_tiny):
    """L = sum (sqrt(sum_k |F_k|^2) - A)^2: dL/dRe O + i dL/dIm O =
    -2 S^2 sum_k scatter(conj(P_k) chi_k) (the un-normalised object term)."""
    m, _, probes, obj, pos = mixed_tiny
    m64 = po.OracleMixedstatePtychography(energy=300e3, num_probes=2, mode="c128")
    m64._region_of_interest_shape = np.array([16, 16])
    rng = np.random.default_rng(9)
    A = rng.uniform(5, 40, (len(pos), 16, 16))
    obj64 = obj.astype(np.complex128)
    pr64 = probes.astype(np.complex128)
    sp, op, chi, _ = m64.forward(obj64, pr64, pos, A)
    g = sum(po.scatter_bincount(np.conj(sp[:, k]) * chi[:, k], pos, (16, 16), obj.shape)
            for k in range(2))
    g = -2 * 256 * g

    def loss(o):
        return m64.forward(o, pr64, pos, A)[3]
    for (r, c) in [(20, 20), (18, 25), (24, 16)]:
        h = 1e-6
        d = np.zeros_like(obj64)
        d[r, c] = h
        gr = (loss(obj64 + d) - loss(obj64 - d)) / (2 * h)
        gi = (loss(obj64 + 1j * d) - loss(obj64 - 1j * d)) / (2 * h)
        assert abs(gr - g[r, c].real) < 1e-4 * np.abs(g).max()
        assert abs(gi - g[r, c].imag) < 1e-4 * np.abs(g).max()


def test_mixed_adjoint_keeps_consistent_truth(mixed_tiny):
    m, _, probes, obj, pos = mixed_tiny
    frac = pos - np.round(pos)
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    sp = po.fft_shift(probes, frac)
    op = obj[rows, cols]
    zero = np.zeros_like(sp)
    o2, p2 = m.adjoint(obj.copy(), probes.copy(), op, sp, pos, zero, 0.5, 1, False)
    np.testing.assert_array_equal(o2, obj)
    np.testing.assert_array_equal(p2, probes)


def test_mixed_adjoint_step_is_normalised_loss_gradient(mixed_tiny):
    """With normalization_min = 0 the object step times the summed mode
    overlap sum_k scatter(|P_k|^2) is -1/(2 S^2) times the FP64 finite
    difference gradient of the mixed loss."""
    _, _, probes, obj, pos = mixed_tiny
    m64 = po.OracleMixedstatePtychography(energy=300e3, num_probes=2, mode="c128")
    m64._region_of_interest_shape = np.array([16, 16])
    rng = np.random.default_rng(10)
    A = rng.uniform(5, 40, (len(pos), 16, 16))
    obj64 = obj.astype(np.complex128)
    pr64 = probes.astype(np.complex128)
    sp, op, chi, _ = m64.forward(obj64, pr64, pos, A)
    new, _ = m64.adjoint(obj64.copy(), pr64.copy(), op, sp, pos, chi, 1.0, 0.0, True)
    overlap = (po.scatter_bincount(np.abs(sp[:, 0]) ** 2, pos, (16, 16), obj.shape)
               + po.scatter_bincount(np.abs(sp[:, 1]) ** 2, pos, (16, 16), obj.shape))
    step = (new - obj64) * np.sqrt(1e-16 + overlap.astype(np.float64) ** 2)

    def loss(o):
        return m64.forward(o, pr64, pos, A)[3]
    scale = np.abs(step).max()
    for (r, c) in [(20, 20), (18, 25), (24, 16)]:
        h = 1e-6
        d = np.zeros_like(obj64)
        d[r, c] = h
        gr = (loss(obj64 + d) - loss(obj64 - d)) / (2 * h)
        gi = (loss(obj64 + 1j * d) - loss(obj64 - 1j * d)) / (2 * h)
        assert abs(-gr / 512 - step[r, c].real) < 1e-4 * scale
        assert abs(-gi / 512 - step[r, c].imag) < 1e-4 * scale


def test_detector_mask_closed_forms(tiny):
    o, probe, obj, pos = tiny
    rng = np.random.default_rng(13)
    A = rng.uniform(1, 50, (len(pos), 16, 16)).astype(np.float32)
    _, _, chi0, e0 = o.forward(obj, probe, pos, A)
    _, _, chi1, e1 = o.forward(obj, probe, pos, A, np.ones((16, 16), np.float32))
    np.testing.assert_array_equal(chi1, chi0)
    assert e1 == e0
    mask = (rng.random((16, 16)) > 0.3).astype(np.float32)
    _, _, chim, em = o.forward(obj, probe, pos, A, mask)
    G0 = np.fft.fft2(chi0.astype(np.complex128))
    Gm = np.fft.fft2(chim.astype(np.complex128))
    # masked pixels carry no update, the others the unmasked one
    assert np.abs(Gm[:, mask == 0]).max() < 1e-5 * np.abs(G0).max()
    np.testing.assert_allclose(Gm[:, mask == 1], G0[:, mask == 1], rtol=0,
                               atol=1e-5 * np.abs(G0).max())
    # and count A^2 in the error
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)
    F = np.abs(np.fft.fft2((po.fft_shift(probe, pos - np.round(pos)) * obj[rows, cols])
                           .astype(np.complex128)))
    ref = np.sum(np.where(mask == 1, (A - F) ** 2, A.astype(np.float64) ** 2))
    assert em == pytest.approx(ref, rel=1e-4)


def test_probe_aperture_and_amplitude_constraints():
    o = po.OracleSingleslicePtychography(energy=300e3)
    o._mode = "c64"
    rng = np.random.default_rng(17)
    p = po.ComplexProbeOracle(300e3, (32, 32), (0.3, 0.3), 20.0, 2.0,
                              parameters={"defocus": 80.0}).build()._array.astype(np.complex64)
    noisy = (p + 0.05 * np.abs(p).max() * (rng.standard_normal(p.shape)
                                           + 1j * rng.standard_normal(p.shape))).astype(np.complex64)
    aperture = np.abs(np.fft.fft2(p)).astype(np.float32)
    out = o.probe_aperture_constraint(noisy, aperture)
    Fo = np.fft.fft2(out.astype(np.complex128))
    scale = np.sum(np.abs(Fo)) / np.sum(aperture)
    np.testing.assert_allclose(np.abs(Fo), scale * aperture, rtol=0, atol=1e-4 * aperture.max() * scale)
    Fn = np.fft.fft2(noisy.astype(np.complex128))
    on = aperture > 1e-3 * aperture.max()
    np.testing.assert_allclose(np.angle(Fo[on] * np.conj(Fn[on])), 0, atol=2e-3)
    assert np.sum(np.abs(out) ** 2) == pytest.approx(np.sum(np.abs(noisy) ** 2), rel=1e-5)
    # amplitude top-hat: intensity kept, factor 1/2 on the radius, ~1 well inside
    q = po.OracleSingleslicePtychography.probe_amplitude_constraint(noisy, 0.25, 0.01)
    assert np.sum(np.abs(q) ** 2) == pytest.approx(np.sum(np.abs(noisy) ** 2), rel=1e-5)
    ratio = np.abs(q) / np.maximum(np.abs(noisy), 1e-30)
    c = ratio[0, 0]  # the normalisation factor (mask 1 at r = 0)
    np.testing.assert_allclose(ratio[8, 0], 0.5 * c, rtol=1e-4)  # r = 8/32 = 0.25
    assert ratio[0, 12] < 1e-6 * c and abs(ratio[2, 2] / c - 1) < 1e-6


def test_gaussian_smoothing_order_and_closed_forms():
    """The object smoothing goes through scipy.ndimage.gaussian_filter, as the
    reference's (ptychographic_constraints.py:119-151).  Pinned here: the
    float64 sum the GPU kernel evaluates (centre tap, then symmetric pairs
    from the outside in, reflect boundaries, float32 after each axis)
    reproduces scipy bit for bit, also when the radius exceeds the axis; a
    delta returns the kernel weights; a constant stays constant."""
    import scipy.ndimage as nd
    from py4dstem_amd.kernels import gaussian_kernel1d

    def refl(i, n):
        p = 2 * n
        m = i % p
        return m if m < n else p - 1 - m

    def axis_pass(a, w, r, axis):
        a = np.moveaxis(a, axis, 0).astype(np.float64)
        n = a.shape[0]
        out = a * w[r]
        for j in range(r, 0, -1):
            lo = [refl(i - j, n) for i in range(n)]
            hi = [refl(i + j, n) for i in range(n)]
            out = out + (a[lo] + a[hi]) * w[r - j]
        return np.moveaxis(out.astype(np.float32), 0, axis)

    rng = np.random.default_rng(0)
    for shape, sigma in [((37, 53), 1.5), ((64, 40), 0.7), ((9, 11), 3.3)]:
        z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        w, r = gaussian_kernel1d(sigma)
        assert r == int(4.0 * sigma + 0.5)
        mine = (axis_pass(axis_pass(z.real, w, r, 0), w, r, 1)
                + 1j * axis_pass(axis_pass(z.imag, w, r, 0), w, r, 1)).astype(np.complex64)
        assert np.array_equal(mine, nd.gaussian_filter(z, sigma))
    w, r = gaussian_kernel1d(2.2)
    delta = np.zeros(64)
    delta[32] = 1.0
    np.testing.assert_array_equal(nd.gaussian_filter1d(delta, 2.2)[32 - r:33 + r], w)
    assert abs(w.sum() - 1) < 1e-15 and np.array_equal(w, w[::-1])
    c = np.full((20, 30), 0.7 - 0.2j, np.complex64)
    np.testing.assert_allclose(nd.gaussian_filter(c, 1.9), c, rtol=1e-6)
    o = po.OracleSingleslicePtychography.object_constraints(
        np.exp(1j * rng.standard_normal((16, 16))).astype(np.complex64), None, 0.0, True, 1.2)
    np.testing.assert_allclose(np.abs(o), 1, rtol=1e-6)  # pure phase stays unit modulus


def test_butterworth_closed_forms():
    """ptychographic_constraints.py:153-201 restated in the oracle: a constant
    plus one plane wave comes back as the constant plus the wave scaled by
    the Butterworth envelope at its frequency (on the fftfreq grid in 1/A);
    the low-pass envelope is 1/2 at the cut-off; cut-offs of 0 leave the
    object unchanged."""
    f = po.OracleSingleslicePtychography.object_butterworth_constraint
    Px, Py, d = 48, 40, (0.2, 0.25)
    x = np.arange(Px)[:, None]
    y = np.arange(Py)[None, :]
    kx, ky = 5, -7
    wave = np.exp(2j * np.pi * (kx * x / Px + ky * y / Py))
    obj = (0.8 + 0.1j + 0.3 * wave).astype(np.complex64)
    q = np.hypot(kx / (Px * d[0]), ky / (Py * d[1]))
    for lo, hi, order in [(1.0, None, 2), (None, 0.9, 3), (2.0, 0.5, 1.5)]:
        env = 1.0
        if hi:
            env *= 1 - 1 / (1 + (q / hi) ** (2 * order))
        if lo:
            env *= 1 / (1 + (q / lo) ** (2 * order))
        out = f(obj.copy(), lo, hi, order, d)
        ref = 0.8 + 0.1j + 0.3 * env * wave
        np.testing.assert_allclose(out, ref, atol=2e-6)
    # the low-pass envelope is 1/2 at the cut-off: a wave exactly at q_c
    out = f(obj.copy(), q, None, 4, d)
    np.testing.assert_allclose(out, 0.8 + 0.1j + 0.15 * wave, atol=2e-6)
    np.testing.assert_allclose(f(obj.copy(), 0, 0, 2, d), obj, atol=2e-6)


def test_vacuum_probe_closed_form():
    """ptychographic_methods.py:1033-1064 restated in the oracle.  A centred,
    point-symmetric, integer-valued vacuum intensity V (float32 sums exact, so
    the centre of mass is exactly the grid centre and the bilinear shift an
    exact roll) gives the probe ifft2(sqrt(V at the corner) exp(-i chi)),
    normalised to the mean diffraction intensity, with exp(-i chi) the
    spectral phase of the default probe; the semiangle cut-off becomes inf.
    The host helpers of the drop-in class equal the oracle's.  (With
    non-exact sums the reference's centre of mass is a few 1e-6 px off the
    grid, and the shift leaks that fraction of the aperture edge into the
    zero pixels next to it, which the square root raises to ~1e-3.)"""
    from py4dstem_amd.preprocess import bilinear_shift, center_of_mass
    from py4dstem_amd.synthetic import small_workload
    w = small_workload(seed=1, scan=(8, 8), object_px=128)
    a = po.OracleSingleslicePtychography(datacube=w["datacube"], **w["init"])
    a.preprocess(**w["preprocess"])
    F = np.fft.fft2(a._probe.astype(np.complex128))
    V = np.abs(F) ** 2
    V = 0.5 * (V + np.roll(V[::-1, ::-1], 1, axis=(0, 1)))  # V(k) = V(-k)
    V = np.round(V / V.max() * 1000).astype(np.float32)
    S = V.shape[0]
    b = po.OracleSingleslicePtychography(datacube=w["datacube"],
                                         vacuum_probe_intensity=np.fft.fftshift(V), **w["init"])
    b.preprocess(**w["preprocess"])
    assert b._semiangle_cutoff == np.inf
    assert po.get_CoM(np.fft.fftshift(V)) == (S // 2, S // 2)
    phase = np.where(np.abs(F) > 0, F / np.maximum(np.abs(F), 1e-300), 0)
    p = np.fft.ifft2(np.sqrt(V) * phase)
    p = p / np.sqrt(np.sum(np.abs(p) ** 2))
    p = p * np.sqrt(b._mean_diffraction_intensity / np.sum(np.abs(np.fft.fft2(p)) ** 2))
    assert np.linalg.norm(b._probe - p) / np.linalg.norm(p) < 2e-6
    rng = np.random.default_rng(2)
    ar = rng.uniform(0, 1, (24, 30)).astype(np.float32)
    assert center_of_mass(ar) == po.get_CoM(ar)
    for dx, dy in [(2.3, -1.7), (-5.0, 3.0), (0.49, -0.51)]:
        assert np.array_equal(bilinear_shift(ar, dx, dy), po.get_shifted_ar(ar, dx, dy))


def test_fit_origin_models_match_curve_fit():
    """fit_origin's models (origin.py:76-190, fit.py:127-162): noiseless
    surfaces are recovered exactly; on noisy masked data the least-squares
    solution equals scipy.optimize.curve_fit (the reference's call) on the
    model functions written out term by term; 'constant' is the mean; the
    drop-in class's host fit equals the oracle's."""
    from scipy.optimize import curve_fit

    from py4dstem_amd.preprocess import fit_origin

    def parabola(xy, c0, cx1, cx2, cy1, cy2, cxy):
        return c0 + cx1 * xy[0] + cy1 * xy[1] + cx2 * xy[0] ** 2 + cy2 * xy[1] ** 2 + cxy * xy[0] * xy[1]

    def bezier_two(xy, c00, c01, c02, c10, c11, c12, c20, c21, c22):
        x, y = xy
        return (c00 * (1 - x) ** 2 * (1 - y) ** 2 + c10 * 2 * (1 - x) * x * (1 - y) ** 2
                + c20 * x ** 2 * (1 - y) ** 2 + c01 * 2 * (1 - x) ** 2 * (1 - y) * y
                + c11 * 4 * (1 - x) * x * (1 - y) * y + c21 * 2 * x ** 2 * (1 - y) * y
                + c02 * (1 - x) ** 2 * y ** 2 + c12 * 2 * (1 - x) * x * y ** 2 + c22 * x ** 2 * y ** 2)

    shape = (13, 11)
    ry, rx = np.meshgrid(np.linspace(0, 1, shape[1]), np.linspace(0, 1, shape[0]))
    rng = np.random.default_rng(4)
    mask = rng.uniform(size=shape) > 0.2
    for name, f, npar in (("parabola", parabola, 6), ("bezier_two", bezier_two, 9)):
        p = rng.standard_normal(npar)
        clean = f((rx, ry), *p)
        np.testing.assert_allclose(po.fit_origin_2d(clean, None, name), clean, atol=1e-10)
        noisy = clean + 0.05 * rng.standard_normal(shape)
        popt, _ = curve_fit(f, np.vstack((rx[mask], ry[mask])), noisy[mask], p0=np.zeros(npar))
        np.testing.assert_allclose(po.fit_origin_2d(noisy, mask, name), f((rx, ry), *popt),
                                   atol=1e-6)
        np.testing.assert_allclose(fit_origin(noisy, mask, name), po.fit_origin_2d(noisy, mask, name),
                                   atol=1e-12)
    d = rng.standard_normal(shape).astype(np.float32)
    assert np.all(po.fit_origin_2d(d, mask, "constant") == np.mean(d))
    assert np.array_equal(fit_origin(d, mask, "constant"), po.fit_origin_2d(d, mask, "constant"))
    with pytest.raises(AssertionError):
        fit_origin(d, mask, "cubic")


def test_potential_object_closed_forms(tiny):
    """Potential objects (ptychographic_methods.py:40-79, 1522-1523,
    1868-1880; ptychographic_constraints.py:60-118): the forward model uses
    exp(i V), so a potential V and the complex object exp(i V) give the same
    exit-wave update and error; the potential's gradient step is the real
    part of -i conj(exp(i V)) times the complex step (for a pixel covered by
    one position it is exactly that product); shrinkage subtracts the
    masked mean, positivity clips at zero."""
    o, probe, obj, pos = tiny
    rng = np.random.default_rng(6)
    V = (0.3 * rng.standard_normal(obj.shape)).astype(np.float32)
    Oc = np.exp(1j * V).astype(np.complex64)
    amps = np.abs(np.fft.fft2(o.forward(Oc, probe, pos, np.zeros((len(pos), 16, 16),
                                                                 np.float32))[0] * 1.1))
    amps = amps.astype(np.float32)
    o._object_type = "complex"
    spc, opc, ewc, errc = o.forward(Oc, probe, pos, amps)
    o._object_type = "potential"
    spv, opv, ewv, errv = o.forward(V, probe, pos, amps)
    np.testing.assert_allclose(opv, opc, atol=1e-6)
    np.testing.assert_allclose(ewv, ewc, atol=1e-5 * np.abs(ewc).max())
    assert errv == pytest.approx(errc, rel=1e-5)
    # one position: the potential update is Re(-i conj(O) g) * n * step per pixel
    p1 = pos[:1]
    sp, op, ew, _ = o.forward(V, probe, p1, amps[:1])
    Vn, _ = o.adjoint(V.copy(), probe.copy(), op, sp, p1, ew, 0.5, 1.0, True)
    o._object_type = "complex"
    Oc2, _ = o.adjoint(Oc.copy(), probe.copy(), op, sp, p1, ew, 0.5, 1.0, True)
    dO = (Oc2 - Oc).astype(np.complex128)
    np.testing.assert_allclose(Vn - V, np.real(-1j * np.conj(Oc) * dO), atol=2e-6)
    mask = np.zeros(V.shape, bool)
    mask[:5] = True
    out = po.OracleSingleslicePtychography.object_constraints(
        V.copy(), mask, 0.1, False, None, None, "potential", False)
    np.testing.assert_allclose(out, V - (0.1 + V[mask].mean()), atol=1e-6)
    assert abs(out[mask].mean() + 0.1) < 1e-6
    clipped = po.OracleSingleslicePtychography.object_constraints(
        V.copy(), None, 0.0, False, None, None, "potential", True)
    assert np.array_equal(clipped, np.clip(V, 0, None))


@pytest.mark.parametrize("dis,method,pad", [((64, 64), "fourier", None), ((48, 48), "fourier", (64, 64)),
                                            ((128, 128), "fourier", None), ((48, 48), "bilinear", None),
                                            ((32, 32), "bin", None), (None, "bilinear", (112, 112))])
def test_resampling_maps_match_the_reference_stack_operations(dis, method, pad):
    """Resampling and padding (phase_base_class.py:223-420): the drop-in
    class's per-axis maps, applied as Re(Mx X My^T), equal the oracle's
    restatement of the reference's per-pattern operations (FFT quadrant copy,
    scipy zoom, block sums, np.pad) to float32 rounding."""
    from py4dstem_amd.resample import axis_maps
    rng = np.random.default_rng(8)
    data = rng.uniform(0, 5, (2, 3, 96, 96)).astype(np.float32)
    ref, fac = po.resample_datacube(data, dis, method, pad)
    Mx, My, fac2 = axis_maps((96, 96), dis, method, pad)
    mine = np.real(np.einsum("ij,abjk,lk->abil", Mx, data.astype(np.float64), My))
    assert mine.shape == ref.shape and fac == fac2
    assert np.max(np.abs(mine - ref)) <= 2e-6 * np.max(np.abs(ref))


def test_resampling_closed_forms():
    """Fourier resampling of a band-limited periodic pattern gives its samples
    on the new grid times (Q/S)^2 (sums conserved); binning a constant c by b
    gives b^2 c; zero padding keeps the sum and puts the pattern (S-Q)//2 in."""
    Q, S = 48, 64
    n = np.arange(Q)
    f = lambda x, N: 1 + np.cos(2 * np.pi * 3 * x / N)[:, None] * np.sin(2 * np.pi * 5 * x / N)[None, :]
    data = f(n, Q)[None, None].astype(np.float32)
    out, fac = po.resample_datacube(data, (S, S), "fourier")
    np.testing.assert_allclose(out[0, 0], f(np.arange(S), S) * (Q / S) ** 2, atol=1e-5)
    assert fac == Q / S and out.sum() == pytest.approx(data.sum(), rel=1e-5)
    c = np.full((1, 1, 9, 9), 0.5, np.float32)
    b, fac = po.resample_datacube(c, (4, 4), "bin")
    assert b.shape == (1, 1, 4, 4) and np.all(b == 2.0) and fac == 2.0
    p, _ = po.resample_datacube(data, None, "bilinear", (S, S + 1))
    assert p.shape == (1, 1, S, S + 1) and p.sum() == pytest.approx(data.sum(), rel=1e-6)
    assert np.array_equal(p[0, 0, 8:8 + Q, 8:8 + Q], data[0, 0])


def test_mixed_position_partials_reduce_to_single_mode(tiny):
    """Mixed-state position partials (ptychographic_methods.py:2187-2189,
    2736-2740): with a second mode of zeros they equal the single-mode
    partials (the intensity sqrt(|F|^2)^2 differs from |F|^2 by rounding
    only); with two modes they are the sums of each mode's own partials."""
    o, probe, obj, pos = tiny
    rng = np.random.default_rng(9)
    sp1, op, _, _ = o.forward(obj, probe, pos, np.ones((len(pos), 16, 16), np.float32))
    amps = (np.abs(np.fft.fft2(sp1 * op)) * rng.uniform(0.9, 1.1, sp1.shape)).astype(np.float32)
    obj2 = (obj * np.exp(0.05j * rng.standard_normal(obj.shape))).astype(np.complex64)
    d1, px1, py1 = o.position_partials(obj2, pos, sp1, sp1 * op, amps)
    sp = np.stack([sp1, np.zeros_like(sp1)], axis=1)
    d, px, py = o.position_partials(obj2, pos, sp, sp * op[:, None], amps)
    np.testing.assert_allclose(px, px1, rtol=1e-6, atol=1e-6 * np.abs(px1).max())
    np.testing.assert_allclose(py, py1, rtol=1e-6, atol=1e-6 * np.abs(py1).max())
    np.testing.assert_allclose(d, d1, atol=1e-5 * np.abs(d1).max())
    sp2 = (0.5 * sp1 * np.exp(1j * 0.3)).astype(np.complex64)
    sp = np.stack([sp1, sp2], axis=1)
    d, px, py = o.position_partials(obj2, pos, sp, sp * op[:, None], amps)
    _, px2, py2 = o.position_partials(obj2, pos, sp2, sp2 * op, amps)
    np.testing.assert_allclose(px, px1 + px2, atol=1e-5 * np.abs(px1).max())
    np.testing.assert_allclose(py, py1 + py2, atol=1e-5 * np.abs(py1).max())
