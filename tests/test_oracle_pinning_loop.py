"""Pins of the oracle's reconstruction loop, step normalisation and rotation
search by what the mathematics fixes.

* Error normalisation (phase/singleslice_ptychography.py:929-932): with a
  zero object the far field vanishes, so the first error is
  sum A^2 / (mean_I N) = 1; with the object c * truth (true probe, integer
  positions, noise-free data) it is (1 - c)^2.
* Loop sequence (:847-983): noise-free truth with a fixed probe and |O| = 1 of
  zero phase outside the scanned field is a fixed point of the whole
  iteration, constraints included; one mini-batch iteration equals the
  reference's sequence (seeded shuffle, near-equal batches, forward, adjoint,
  error, constraints) composed by hand from the separately pinned parts.
* Step normalisation (phase/ptychographic_methods.py:1859-1906): for one
  position, normalization_min = 0 divides the step by the probe (object
  step step * chi / P) and by the patch (probe step step * chi / O_patch);
  normalization_min = 1 divides by the maximum overlap; for several
  positions at normalization_min = 1 the object step is the loss gradient
  (FP64 central differences of |fft2(P O)|) times -step / (2 S^2 max overlap).
* Rotation search (phase/phase_base_class.py:769-1345): a curl-free field
  rotated by a known angle (plain or transposed) is recovered exactly, and a
  radial field for the divergence branch.
"""

import numpy as np
import pytest
import scipy.ndimage

import oracle.ptycho_oracle as po
from py4dstem_amd.datacube import Calibration, DataCube
from tests.conftest import rel

S = 32
SAMPLING = 0.2            # Angstrom per object pixel
OBJ = 96
STEP = 4                  # scan step in pixels (integer positions)
SCAN = (12, 12)


def _centres():
    ext = (np.array(SCAN) - 1) * STEP
    gx = OBJ // 2 - ext[0] // 2 + np.arange(SCAN[0]) * STEP
    gy = OBJ // 2 - ext[1] // 2 + np.arange(SCAN[1]) * STEP
    cx, cy = np.meshgrid(gx, gy, indexing="ij")
    return np.stack([cx.ravel(), cy.ravel()], axis=1).astype(np.float64)


def _truth(seed=3, phase_outside_zero=True):
    """|O| = 1 phase object; phase only well inside the scanned field."""
    rng = np.random.default_rng(seed)
    phi = scipy.ndimage.gaussian_filter(rng.standard_normal((OBJ, OBJ)), 2.0)
    phi *= 0.3 / np.abs(phi).max()
    if phase_outside_zero:
        inner = np.zeros((OBJ, OBJ), bool)
        inner[30:66, 30:66] = True
        phi[~inner] = 0.0
    return np.exp(1j * phi).astype(np.complex64)


def _probe():
    p = po.ComplexProbeOracle(300e3, (S, S), (SAMPLING, SAMPLING), 20.0, 2.0,
                              parameters={"C10": -60.0}).build()._array
    return (p * 200.0).astype(np.complex64)


def _noise_free_case(obj, probe):
    """Datacube of |fft2(P O_patch)|^2 at integer positions, centred as a
    detector records it; positions in Angstrom for initial_scan_positions."""
    c = _centres()
    idx = np.fft.fftfreq(S, 1 / S).astype(int)
    frames = []
    for x0, y0 in c.astype(int):
        patch = obj.astype(np.complex128)[np.ix_((x0 + idx) % OBJ, (y0 + idx) % OBJ)]
        frames.append(np.abs(np.fft.fft2(probe.astype(np.complex128) * patch)) ** 2)
    data = np.fft.fftshift(np.array(frames, np.float32), axes=(-2, -1)).reshape(SCAN + (S, S))
    lam = po.electron_wavelength_angstrom(300e3)
    dk = 1.0 / (S * SAMPLING)
    cal = Calibration(R_pixel_size=STEP * SAMPLING, R_pixel_units="A", Q_pixel_size=dk,
                      Q_pixel_units="A^-1")
    assert lam > 0
    return DataCube(data, cal), (c - S / 2) * SAMPLING


PRE = dict(force_com_rotation=0, force_com_transpose=False, force_com_shifts=(S / 2, S / 2),
           center_positions_in_fov=False)


def _oracle(obj_guess, probe, dc, pos_ang, mode="c64"):
    o = po.OracleSingleslicePtychography(
        energy=300e3, datacube=dc, semiangle_cutoff=20.0, initial_object_guess=obj_guess.copy(),
        initial_probe_guess=probe.copy(), initial_scan_positions=pos_ang, mode=mode)
    o.preprocess(**PRE)
    return o


@pytest.fixture(scope="module")
def case():
    obj, probe = _truth(), _probe()
    dc, pos_ang = _noise_free_case(obj, probe)
    return obj, probe, dc, pos_ang


def test_case_positions_are_the_integer_centres(case):
    obj, probe, dc, pos_ang = case
    o = _oracle(obj, probe, dc, pos_ang)
    np.testing.assert_allclose(o._positions_px, _centres(), atol=1e-4)
    # amplitudes are the corner-centred sqrt of the simulated intensities
    assert o._amplitudes.shape == (SCAN[0] * SCAN[1], S, S)


@pytest.mark.parametrize("c", [0.0, 0.6])
def test_first_error_is_one_minus_c_squared(case, c):
    """error_1 = sum_b sum (A - |F|)^2 / (mean_I N) with F = c F_true, A =
    |F_true|: (1 - c)^2 sum A^2 / sum I = (1 - c)^2."""
    obj, probe, dc, pos_ang = case
    o = _oracle(c * obj, probe, dc, pos_ang)
    o.reconstruct(num_iter=1, fix_probe=True, max_batch_size=None)
    assert o.error_iterations[0] == pytest.approx((1 - c) ** 2, rel=2e-5, abs=1e-6)


def test_noise_free_truth_is_a_fixed_point_of_the_loop(case):
    """Truth object and probe, probe fixed: every error ~0 and the object
    passes forward, adjoint, shrinkage and threshold unchanged."""
    obj, probe, dc, pos_ang = case
    o = _oracle(obj, probe, dc, pos_ang)
    assert o._object_fov_mask[48, 48] and not o._object_fov_mask[2, 2]
    o.reconstruct(num_iter=3, fix_probe=True, max_batch_size=50, seed_random=1)
    assert max(o.error_iterations) < 1e-10
    assert rel(o.object, obj) < 2e-6


def test_one_iteration_equals_the_hand_composed_sequence(case):
    """reconstruct(num_iter=1, mini-batches, probe update) against the loop
    of singleslice_ptychography.py:847-983 written out from the pinned
    pieces: np.random.seed, shuffle, near-equal batches (larger first),
    forward/adjoint per batch, error / (mean_I N), object constraints with
    the FOV mask, probe recentring."""
    obj, probe, dc, pos_ang = case
    start = _truth(seed=9, phase_outside_zero=False)
    o = _oracle(start, probe, dc, pos_ang)
    h = _oracle(start, probe, dc, pos_ang)
    o.reconstruct(num_iter=1, fix_probe=False, max_batch_size=30, seed_random=5)

    N = h._num_diffraction_patterns
    np.random.seed(5)
    order = np.arange(N)
    np.random.shuffle(order)
    sizes = [29, 29, 29, 29, 28]         # 144 positions, at most 30: 5 batches, larger first
    assert sum(sizes) == N
    objh, probeh, err, s0 = h._object.copy(), h._probe.copy(), 0.0, 0
    for n in sizes:
        idx = order[s0:s0 + n]
        pos = h._positions_px[idx]
        sp, op, chi, e = h.forward(objh, probeh, pos, h._amplitudes[idx])
        objh, probeh = h.adjoint(objh, probeh, op, sp, pos, chi, 0.5, 1, False)
        err += e
        s0 += n
    err /= h._mean_diffraction_intensity * N
    objh = h.object_constraints(objh, h._object_fov_mask_inverse, 0.0, False, None, None,
                                "complex", True)
    probeh = h.probe_com_constraint(probeh)
    assert o.error_iterations[0] == pytest.approx(float(err), rel=1e-12)
    np.testing.assert_array_equal(o._object, objh)
    np.testing.assert_array_equal(o._probe, probeh)


# ---------------------------------------------------------------------------
# step normalisation
# ---------------------------------------------------------------------------
def _one_position(seed=4, frac=(0.3, -0.2)):
    rng = np.random.default_rng(seed)
    o = po.OracleSingleslicePtychography(energy=300e3)
    o._region_of_interest_shape = np.array([16, 16])
    probe = (po.ComplexProbeOracle(300e3, (16, 16), (0.4, 0.4), 20.0, 2.0,
                                   parameters={"defocus": 40.0}).build()._array * 25
             ).astype(np.complex64)
    obj = (np.exp(1j * 0.2 * rng.standard_normal((40, 40)))
           * rng.uniform(0.7, 1.0, (40, 40))).astype(np.complex64)
    pos = np.array([[20 + frac[0], 19 + frac[1]]], np.float32)
    amps = rng.uniform(0, 30, (1, 16, 16)).astype(np.float32)
    return o, probe, obj, pos, amps


def _window(pos):
    idx = np.fft.fftfreq(16, 1 / 16).astype(int)
    x0, y0 = np.round(pos[0]).astype(int)
    return np.ix_((x0 + idx) % 40, (y0 + idx) % 40)


@pytest.mark.parametrize("n", [0.0, 1.0])
def test_one_position_object_and_probe_steps(n):
    o, probe, obj, pos, amps = _one_position()
    sp, op, chi, _ = o.forward(obj, probe, pos, amps)
    obj2, probe2 = o.adjoint(obj.copy(), probe.copy(), op, sp, pos, chi, 0.5, n, False)
    win = _window(pos)
    dobj = (obj2 - obj)[win].astype(np.complex128)
    dprobe = (probe2 - probe).astype(np.complex128)
    P, O, X = sp[0].astype(np.complex128), op[0].astype(np.complex128), chi[0].astype(
        np.complex128)
    outside = np.ones(obj.shape, bool)
    outside[win] = False
    assert np.all(obj2[outside] == obj[outside])
    if n == 0.0:
        big = np.abs(P) ** 2 > 1e-3 * np.max(np.abs(P) ** 2)
        np.testing.assert_allclose(dobj[big], 0.5 * X[big] / P[big], rtol=2e-4,
                                   atol=1e-6 * np.abs(dobj).max())
        np.testing.assert_allclose(dprobe, 0.5 * X / O, rtol=2e-4, atol=1e-6 * np.abs(dprobe).max())
    else:
        np.testing.assert_allclose(dobj, 0.5 * np.conj(P) * X / np.max(np.abs(P) ** 2),
                                   rtol=2e-4, atol=1e-6 * np.abs(dobj).max())
        np.testing.assert_allclose(dprobe, 0.5 * np.conj(O) * X / np.max(np.abs(O) ** 2),
                                   rtol=2e-4, atol=1e-6 * np.abs(dprobe).max())


def test_object_step_is_normalised_fp64_loss_gradient(tiny_many):
    """normalization_min = 1, several positions: dO = -step / (2 S^2 m) dL/dO*,
    m = max over pixels of sum_b |P_shift_b|^2 placed at the windows, dL/dO*
    by central differences of L = sum (|fft2(P O)| - A)^2 in float64."""
    o, probe, obj, pos, amps = tiny_many
    sp, op, chi, _ = o.forward(obj, probe, pos, amps)
    obj2, _ = o.adjoint(obj.copy(), probe.copy(), op, sp, pos, chi, 0.5, 1, True)
    idx = np.fft.fftfreq(16, 1 / 16).astype(int)
    overlap = np.zeros(obj.shape)
    for b in range(len(pos)):
        x0, y0 = np.round(pos[b]).astype(int)
        for i in range(16):
            for j in range(16):
                overlap[(x0 + idx[i]) % 40, (y0 + idx[j]) % 40] += abs(complex(sp[b, i, j])) ** 2
    m = overlap.max()
    probe64 = probe.astype(np.complex128)
    frac = pos - np.round(pos)
    sp64 = po.fft_shift(probe64, frac, "c128")
    rows, cols = po.patch_indices(pos, (16, 16), obj.shape)

    def loss(objv):
        return np.sum((np.abs(np.fft.fft2(sp64 * objv[rows, cols])) - amps) ** 2)

    o64 = obj.astype(np.complex128)
    h = 1e-6
    for (i, j) in [(20, 20), (17, 24), (24, 15)]:
        g = 0j
        for unit in (1.0, 1j):
            e = np.zeros_like(o64)
            e[i, j] = unit * h
            g += unit * (loss(o64 + e) - loss(o64 - e)) / (2 * h)
        want = -0.5 / (2 * 256 * m) * g
        got = complex(obj2[i, j]) - complex(obj[i, j])
        assert abs(got - want) <= 2e-4 * abs(want) + 1e-7


@pytest.fixture(scope="module")
def tiny_many():
    rng = np.random.default_rng(6)
    o = po.OracleSingleslicePtychography(energy=300e3)
    o._region_of_interest_shape = np.array([16, 16])
    probe = (po.ComplexProbeOracle(300e3, (16, 16), (0.4, 0.4), 20.0, 2.0,
                                   parameters={"defocus": 50.0}).build()._array * 30
             ).astype(np.complex64)
    obj = np.exp(1j * 0.3 * scipy.ndimage.gaussian_filter(rng.standard_normal((40, 40)), 2)
                 / 0.1).astype(np.complex64)
    pos = rng.uniform(14, 26, (10, 2)).astype(np.float32)
    amps = rng.uniform(0, 40, (10, 16, 16)).astype(np.float32)
    return o, probe, obj, pos, amps


# ---------------------------------------------------------------------------
# rotation / transpose search
# ---------------------------------------------------------------------------
def _grad_field(n=24):
    """Gradient of x^2 y + x^3 / 3 + y^3 (x along axis 0): the central
    difference curl of this field is exactly zero."""
    x, y = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n), indexing="ij")
    return 2 * x * y + x ** 2, x ** 2 + 3 * y ** 2


def _rot(gx, gy, th):
    return np.cos(th) * gx - np.sin(th) * gy, np.sin(th) * gx + np.cos(th) * gy


@pytest.mark.parametrize("deg", [30.0, -47.0, 0.0])
def test_rotation_search_recovers_rotated_gradient(deg):
    gx, gy = _grad_field()
    nx, ny = _rot(gx, gy, -np.deg2rad(deg))
    rot, transpose, cx, cy = po.solve_com_relative_rotation(nx.astype(np.float32),
                                                            ny.astype(np.float32))
    assert not transpose
    assert np.rad2deg(rot) == pytest.approx(deg, abs=1e-4)
    np.testing.assert_allclose(cx, gx, atol=1e-5)
    np.testing.assert_allclose(cy, gy, atol=1e-5)
    assert po._mean_abs_curl(cx, cy) < 1e-5


@pytest.mark.parametrize("deg", [12.0, -71.0])
def test_rotation_search_detects_transpose(deg):
    gx, gy = _grad_field()
    ax, ay = _rot(gx, gy, -np.deg2rad(deg))
    nx, ny = ay.astype(np.float32), ax.astype(np.float32)   # detector axes swapped
    rot, transpose, cx, cy = po.solve_com_relative_rotation(nx, ny)
    assert transpose
    assert np.rad2deg(rot) == pytest.approx(deg, abs=1e-4)
    np.testing.assert_allclose(cx, gx, atol=1e-5)
    # forced angle: the transpose is decided at that angle
    assert po.solve_com_relative_rotation(nx, ny, force_com_rotation=deg)[1]
    assert not po.solve_com_relative_rotation(ay.astype(np.float32), ax.astype(np.float32)[::-1],
                                              force_com_rotation=deg,
                                              force_com_transpose=False)[1]
    # forced transpose, searched angle
    r2, t2, _, _ = po.solve_com_relative_rotation(nx, ny, force_com_transpose=True)
    assert t2 and np.rad2deg(r2) == pytest.approx(deg, abs=1e-4)


def test_rotation_search_divergence_branch():
    """Radial field (x, y): divergence 2 cos(theta - theta0) after rotating by
    theta, maximal at theta0."""
    n = 20
    x, y = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n), indexing="ij")
    th0 = np.deg2rad(-25.0)
    nx, ny = _rot(x, y, -th0)
    rot, transpose, _, _ = po.solve_com_relative_rotation(
        nx.astype(np.float32), ny.astype(np.float32), maximize_divergence=True)
    assert np.rad2deg(rot) == pytest.approx(-25.0, abs=1e-4)
    # swapped axes make the field a reflection of (x, y), whose divergence is
    # zero at every angle: the plain field wins
    assert not transpose


def test_product_rotation_search_equals_oracle():
    """The product's host search against the oracle's restatement, all four
    branches, curl and divergence, on random fields."""
    from py4dstem_amd.preprocess import solve_com_rotation
    rng = np.random.default_rng(12)
    for trial in range(12):
        nx = scipy.ndimage.gaussian_filter(rng.standard_normal((16, 18)), 2).astype(np.float32)
        ny = scipy.ndimage.gaussian_filter(rng.standard_normal((16, 18)), 2).astype(np.float32)
        for kw in ({}, {"force_com_rotation": 17.0}, {"force_com_transpose": True},
                   {"force_com_transpose": False}, {"force_com_rotation": -3.0,
                                                    "force_com_transpose": True}):
            md = bool(trial % 2)
            a = solve_com_rotation(nx, ny, maximize_divergence=md, **kw)
            b = po.solve_com_relative_rotation(nx, ny, maximize_divergence=md, **kw)
            assert a[0] == b[0] and a[1] == b[1]
            np.testing.assert_array_equal(a[2], b[2])
            np.testing.assert_array_equal(a[3], b[3])


def test_oracle_preprocess_runs_the_search(case):
    """Without forced rotation/transpose the oracle preprocess now runs the
    search (it raised before): the simulated patterns have no CoM shift, so
    every CoM is the pattern centre and the normalised CoM field is ~0."""
    obj, probe, dc, pos_ang = case
    o = po.OracleSingleslicePtychography(
        energy=300e3, datacube=dc, semiangle_cutoff=20.0, initial_object_guess=obj.copy(),
        initial_probe_guess=probe.copy(), initial_scan_positions=pos_ang)
    o.preprocess(force_com_shifts=(S / 2, S / 2), center_positions_in_fov=False)
    assert np.isfinite(o._rotation_best_rad)
    assert o._com_x.shape == SCAN


# ---------------------------------------------------------------------------
# planted mistakes: each variant of the oracle must fail one of the pins above
# ---------------------------------------------------------------------------
class _DroppedSquare(po.OracleSingleslicePtychography):
    """Overlap normalisation from |P_shift| instead of |P_shift|^2."""

    def adjoint(self, obj, probe, object_patches, shifted_probes, positions_px, exit_waves,
                step_size, normalization_min, fix_probe):
        roi, shape = self._region_of_interest_shape, obj.shape
        pn = po.scatter_bincount(np.abs(shifted_probes), positions_px, roi, shape)
        pn = 1 / np.sqrt(1e-16 + ((1 - normalization_min) * pn) ** 2
                         + (normalization_min * np.max(pn)) ** 2)
        obj += step_size * po.scatter_bincount(np.conj(shifted_probes) * exit_waves,
                                               positions_px, roi, shape) * pn
        return obj, probe


class _PerPositionMax(po.OracleSingleslicePtychography):
    """Maximum of one position's |P_shift|^2 instead of the summed overlap."""

    def adjoint(self, obj, probe, object_patches, shifted_probes, positions_px, exit_waves,
                step_size, normalization_min, fix_probe):
        _, probe = super().adjoint(obj.copy(), probe, object_patches, shifted_probes,
                                   positions_px, exit_waves, step_size, normalization_min,
                                   fix_probe)
        roi, shape = self._region_of_interest_shape, obj.shape
        pn = po.scatter_bincount(np.abs(shifted_probes) ** 2, positions_px, roi, shape)
        m = np.max(np.abs(shifted_probes) ** 2)
        pn = 1 / np.sqrt(1e-16 + ((1 - normalization_min) * pn) ** 2 + (normalization_min * m) ** 2)
        obj += step_size * po.scatter_bincount(np.conj(shifted_probes) * exit_waves,
                                               positions_px, roi, shape) * pn
        return obj, probe


class _ProbeStepByProbe(po.OracleSingleslicePtychography):
    """Probe step normalised by the probe intensity instead of the patches'."""

    def adjoint(self, obj, probe, object_patches, shifted_probes, positions_px, exit_waves,
                step_size, normalization_min, fix_probe):
        obj, _ = super().adjoint(obj, probe.copy(), object_patches, shifted_probes, positions_px,
                                 exit_waves, step_size, normalization_min, True)
        if not fix_probe:
            on = np.sum(np.abs(shifted_probes) ** 2, axis=0)
            on = 1 / np.sqrt(1e-16 + ((1 - normalization_min) * on) ** 2
                             + (normalization_min * np.max(on)) ** 2)
            probe += step_size * np.sum(np.conj(object_patches) * exit_waves, axis=0) * on
        return obj, probe


def _one_position_ok(cls, n):
    o0, probe, obj, pos, amps = _one_position()
    o = cls(energy=300e3)
    o._region_of_interest_shape = o0._region_of_interest_shape
    sp, op, chi, _ = o.forward(obj, probe, pos, amps)
    obj2, probe2 = o.adjoint(obj.copy(), probe.copy(), op, sp, pos, chi, 0.5, n, False)
    P, O, X = sp[0].astype(np.complex128), op[0].astype(np.complex128), chi[0].astype(
        np.complex128)
    dobj = (obj2 - obj)[_window(pos)]
    want_o = 0.5 * np.conj(P) * X / np.max(np.abs(P) ** 2)
    want_p = 0.5 * np.conj(O) * X / np.max(np.abs(O) ** 2)
    return (np.allclose(dobj, want_o, rtol=2e-4, atol=1e-6 * np.abs(want_o).max())
            and np.allclose(probe2 - probe, want_p, rtol=2e-4, atol=1e-6 * np.abs(want_p).max()))


def test_planted_mistakes_fail_the_pins(case, tiny_many):
    # wrong error divisor: the total intensity in place of the mean
    obj, probe, dc, pos_ang = case
    o = _oracle(0 * obj, probe, dc, pos_ang)
    o._mean_diffraction_intensity *= o._num_diffraction_patterns
    o.reconstruct(num_iter=1, fix_probe=True)
    assert abs(o.error_iterations[0] - 1.0) > 0.1
    # the true oracle passes the one-position check, the planted ones fail it
    assert _one_position_ok(po.OracleSingleslicePtychography, 1.0)
    assert not _one_position_ok(_DroppedSquare, 1.0)
    assert not _one_position_ok(_ProbeStepByProbe, 1.0)
    # a per-position maximum is the summed overlap when there is one position;
    # with several it no longer gives the finite-difference-pinned step
    assert _one_position_ok(_PerPositionMax, 1.0)
    o0, probe, obj, pos, amps = tiny_many
    bad = _PerPositionMax(energy=300e3)
    bad._region_of_interest_shape = o0._region_of_interest_shape
    sp, op, chi, _ = bad.forward(obj, probe, pos, amps)
    obj_bad, _ = bad.adjoint(obj.copy(), probe.copy(), op, sp, pos, chi, 0.5, 1, True)
    obj_good, _ = o0.adjoint(obj.copy(), probe.copy(), op, sp, pos, chi, 0.5, 1, True)
    d_bad, d_good = obj_bad - obj, obj_good - obj
    assert rel(d_bad, d_good) > 0.1   # the good one is the FD-pinned step
