"""CPU-side checks: the C-ABI library exports every declared symbol, and the
product's host-side set-up code agrees with the oracle."""

import os
import re

import numpy as np
import pytest

import oracle.ptycho_oracle as po
from tests.conftest import ROOT, rel


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pty_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pty_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes

    import torch  # noqa: F401  (HIP runtime first)

    lib_path = os.path.join(ROOT, "py4dstem_amd", "libpty_hip.so")
    if not os.path.exists(lib_path):
        from py4dstem_amd import _build
        _build.build()
    lib = ctypes.CDLL(lib_path)
    syms = declared_symbols()
    assert len(syms) >= 15
    for name in syms:
        assert hasattr(lib, name), name
    from py4dstem_amd import _lib
    assert sorted(_lib.EXPORTED) == syms


def test_product_subdivide_matches_oracle():
    from py4dstem_amd.ptychography import subdivide_into_batches
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(1, 5000))
        b = int(rng.integers(1, n + 1))
        assert subdivide_into_batches(n, b) == po.subdivide_into_batches(n, max_batch=b)


def test_product_probe_matches_oracle():
    from py4dstem_amd.probe import ComplexProbe
    params = {"defocus": 150.0, "C12": 20.0, "phi12": 0.4, "Cs": 1e5, "C23": 300.0,
              "phi23": 0.1, "C45": 1e7, "phi45": 0.2}
    for S in (64, 128):
        a = ComplexProbe(300e3, (S, S), (0.15, 0.15), 22.0, 2.0, parameters=dict(params)).build()
        b = po.ComplexProbeOracle(300e3, (S, S), (0.15, 0.15), 22.0, 2.0,
                                  parameters=dict(params)).build()
        np.testing.assert_array_equal(a._array, b._array)


def test_product_positions_match_oracle(small_case):
    from py4dstem_amd.preprocess import scan_positions_in_pixels
    w, o = small_case
    pos, _ = scan_positions_in_pixels(w["positions_ang"], (32, 32), (1.0, 1.0), 0.0, False,
                                      o.sampling, (64, 64))
    np.testing.assert_array_equal(pos.astype(np.float32), o._positions_px)
    np.testing.assert_array_equal(pos.astype(np.float32), w["positions_px"])
    # rotated and transposed raster as well
    o._rotation_best_rad = np.deg2rad(17.0)
    o._rotation_best_transpose = True
    ref, _ = o._positions_in_pixels(None, None)
    got, _ = scan_positions_in_pixels(None, (32, 32), (1.0, 1.0), np.deg2rad(17.0), True,
                                      o.sampling, np.array([64, 64]))
    np.testing.assert_array_equal(got, ref)
    o._rotation_best_rad, o._rotation_best_transpose = 0.0, False


def test_product_rotation_search_matches_closed_form():
    """A CoM field that is a pure gradient rotated by 30 degrees has zero
    curl at -30 degrees after derotation."""
    from py4dstem_amd.preprocess import solve_com_rotation
    x, y = np.meshgrid(np.linspace(-1, 1, 24), np.linspace(-1, 1, 24), indexing="ij")
    gx, gy = 2 * x * y + x**2, x**2 + 3 * y**2   # gradient of x^2 y + x^3/3 + y^3
    th = np.deg2rad(30.0)
    cx = np.cos(-th) * gx - np.sin(-th) * gy
    cy = np.sin(-th) * gx + np.cos(-th) * gy
    rot, transpose, _, _ = solve_com_rotation(cx.astype(np.float32), cy.astype(np.float32))
    assert not transpose
    assert np.rad2deg(rot) == pytest.approx(30.0, abs=0.5)


def test_workload_generators_shapes():
    from py4dstem_amd.synthetic import small_workload
    w = small_workload(seed=0, scan=(4, 5), S=64, object_px=128)
    assert w["datacube"].data.shape == (4, 5, 64, 64)
    assert w["datacube"].data.dtype == np.float32
    total = w["datacube"].data.reshape(20, -1).sum(1)
    assert np.all(np.abs(total / 1e5 - 1) < 0.02)   # ~1e5 counts per pattern
