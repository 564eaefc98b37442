import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libpty_hip.so")
    config.addinivalue_line("markers", "slow: long-running test")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def emu():
    """ctypes handle on the CPU emulator of the kernels (built on demand)."""
    from __graft_entry__ import EMU_LIB, EMU_SRC, build_emulator

    if not os.path.exists(EMU_LIB) or os.path.getmtime(EMU_LIB) < max(
            os.path.getmtime(EMU_SRC),
            os.path.getmtime(os.path.join(ROOT, "py4dstem_amd", "csrc", "pty_core.h"))):
        build_emulator()
    lib = ctypes.CDLL(EMU_LIB)
    lib.emu_violations.restype = ctypes.c_long
    lib.emu_first_violation.restype = ctypes.c_char_p
    return lib


def cptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


@pytest.fixture(scope="session")
def small_case():
    """Small seeded workload (32x32 scan, 64 px frames, 256^2 object) plus
    a preprocessed oracle instance."""
    from oracle import OracleSingleslicePtychography
    from py4dstem_amd.synthetic import small_workload

    w = small_workload(seed=1)
    init = dict(w["init"], initial_object_guess=w["init"]["initial_object_guess"].copy())
    o = OracleSingleslicePtychography(datacube=w["datacube"], **init)
    o.preprocess(**w["preprocess"])
    return w, o


@pytest.fixture(scope="session")
def small_case32():
    """The same kind of workload with 32 px frames (16x16 scan, 128^2 object,
    0.4 A pixels) plus a preprocessed oracle instance."""
    from oracle import OracleSingleslicePtychography
    from py4dstem_amd.synthetic import small_workload

    w = small_workload(seed=2, scan=(16, 16), S=32, object_px=128)
    init = dict(w["init"], initial_object_guess=w["init"]["initial_object_guess"].copy())
    o = OracleSingleslicePtychography(datacube=w["datacube"], **init)
    o.preprocess(**w["preprocess"])
    return w, o


@pytest.fixture(params=[64, 32], ids=["S64", "S32"])
def frame_case(request):
    """The 64 px and the 32 px small workloads."""
    return request.getfixturevalue("small_case" if request.param == 64 else "small_case32")


def far_staging_to_kernel(far):
    """(B, S, S) far field from the fused pass's staging layout (element j of
    row-pass thread t at [j][t]) to the kernel layout (NumPy)."""
    S = far.shape[-1]
    N1 = {32: 8, 64: 8, 128: 16}[S]
    x = far.reshape(far.shape[0], N1, S, S // N1)
    return np.ascontiguousarray(x.transpose(0, 2, 3, 1).reshape(far.shape[0], S, S))
