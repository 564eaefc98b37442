"""Host side of the position constraints (no GPU): the affine model of the
scan used by global_affine_transformation, against planted maps and the
oracle's restatement."""

import numpy as np
import pytest

import oracle.ptycho_oracle as po
from py4dstem_amd.positions import (
    AffineTransform,
    estimate_global_transformation,
    estimate_global_transformation_ransac,
)


def _scan(n=20, step=4.0):
    g = np.arange(n, dtype=np.float32) * step
    gx, gy = np.meshgrid(g, g, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32) + 30


def test_identity_and_planted_map_round_trip():
    p0 = _scan()
    origin = p0.mean(0)
    np.testing.assert_array_equal(AffineTransform()(p0, origin=origin), p0)
    tf = AffineTransform(scale0=1.03, scale1=0.98, shear1=-0.02, angle=0.05,
                         t0=np.float32(1.5), t1=np.float32(-0.7))
    p1 = tf(p0, origin=origin)
    assert p1.dtype == np.float32
    est, err = estimate_global_transformation(p0, p1, origin=origin)
    np.testing.assert_allclose(est.astuple()[:4], tf.astuple()[:4], atol=2e-5)
    np.testing.assert_allclose(est.astuple()[4:], (1.5, -0.7), atol=2e-3)
    assert err < 0.05
    # 3x2 matrix form: rows of the 2x2 map, then the translation
    np.testing.assert_allclose(tf.asarray3()[:2], tf.asarray(), rtol=0)
    np.testing.assert_allclose(tf.asarray3()[2], (1.5, -0.7), rtol=1e-6)


def test_degenerate_fit_is_identity():
    p0 = _scan()
    est, _ = estimate_global_transformation(p0[:2], p0[:2] + 1, origin=p0.mean(0))
    assert est.astuple() == AffineTransform().astuple()


@pytest.mark.parametrize("seed", [0, 3])
def test_ransac_matches_oracle_and_rejects_outliers(seed):
    p0 = _scan()
    origin = p0.mean(0)
    tf = AffineTransform(scale0=1.01, scale1=1.02, shear1=0.01, angle=0.02,
                         t0=np.float32(0.4), t1=np.float32(0.2))
    p1 = tf(p0, origin=origin).astype(np.float32)
    rng = np.random.default_rng(seed)
    bad = rng.choice(len(p0), len(p0) // 10, replace=False)
    p1[bad] += rng.uniform(30, 60, (len(bad), 2)).astype(np.float32)
    np.random.seed(seed)
    est, fit = estimate_global_transformation_ransac(p0, p1, origin=origin,
                                                     min_sample=len(p0) // 10)
    state = np.random.get_state()[1].copy()
    np.random.seed(seed)
    ref, ref_fit = po.affine_ransac(p0, p1, origin, min_sample=len(p0) // 10)
    # same draws from the global generator, same arithmetic
    np.testing.assert_array_equal(np.random.get_state()[1], state)
    assert est.astuple()[:4] == ref[:4] and fit == ref_fit
    np.testing.assert_array_equal(est(p0, origin=origin), po.affine_apply(ref, p0, origin))
    np.testing.assert_allclose(est.astuple()[:4], tf.astuple()[:4], atol=1e-4)
