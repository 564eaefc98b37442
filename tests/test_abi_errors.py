"""C-ABI error paths on the host: argument validation returns a negative
PTY_E* code before anything reaches the GPU, and the Python binding turns it
into RuntimeError carrying pty_last_error().

The calls below are rejected by the library's host-side checks (frame size,
null pointers, workspace size, sizes), so the pointers they carry are never
dereferenced and no device is needed.
"""

import ctypes

import pytest

from tests.conftest import ROOT  # noqa: F401  (puts the repo on sys.path)

PTY_E_ARG, PTY_E_UNSUPPORTED, PTY_E_WORKSPACE = -1, -2, -4
FAKE = 0x10000  # an aligned, never dereferenced stand-in for a device pointer


@pytest.fixture(scope="module")
def lib():
    import os

    from py4dstem_amd import _build, _lib
    if not os.path.exists(_build.LIB_PATH):
        _build.build()
    return _lib


def _problem(L, S=64, Px=256, Py=256, N=100, **over):
    p = L.PtyProblem(S, Px, Py, N, FAKE, FAKE, FAKE, FAKE, None, 1.0, None, 0)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def _params(L, max_batch=50, num_iter=1):
    q = L.PtyGdParams()
    q.num_iter, q.max_batch, q.step_size, q.normalization_min = num_iter, max_batch, 0.5, 1.0
    return q


def _gd(L, p, q, ws_bytes=1 << 40):
    return L.load().pty_gd_run(ctypes.byref(p), ctypes.byref(q), ctypes.c_void_p(FAKE),
                               ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), ws_bytes, None)


def test_unsupported_frame_size_is_rejected(lib):
    rc = _gd(lib, _problem(lib, S=48), _params(lib))
    assert rc == PTY_E_UNSUPPORTED
    assert "S=48" in lib.last_error()
    with pytest.raises(RuntimeError, match="not supported"):
        lib.check(rc, "pty_gd_run")
    with pytest.raises(RuntimeError, match="S=96"):
        lib.call("pty_fft2_c64", ctypes.c_void_p(FAKE), ctypes.c_void_p(2 * FAKE), 1, 96, 0, None)


def test_null_pointers_are_rejected(lib):
    rc = _gd(lib, _problem(lib, object=None), _params(lib))
    assert rc == PTY_E_ARG and "null" in lib.last_error()
    rc = lib.load().pty_gd_run(None, None, None, None, None, 0, None)
    assert rc == PTY_E_ARG
    with pytest.raises(RuntimeError, match="null buffer"):
        lib.call("pty_forward_adjoint_batch", None, ctypes.c_void_p(FAKE), 256, 256,
                 ctypes.c_void_p(FAKE), 10, ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), 10, 64,
                 ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), None, None)
    with pytest.raises(RuntimeError, match="in place"):
        lib.call("pty_fft2_c64", ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), 1, 64, 0, None)


def test_misaligned_and_bad_sizes_are_rejected(lib):
    assert _gd(lib, _problem(lib, probe=FAKE + 4), _params(lib)) == PTY_E_ARG
    assert "aligned" in lib.last_error()
    assert _gd(lib, _problem(lib, Px=32), _params(lib)) == PTY_E_ARG        # object < frame
    assert _gd(lib, _problem(lib), _params(lib, max_batch=101)) == PTY_E_ARG  # batch > N
    assert _gd(lib, _problem(lib), _params(lib, max_batch=0)) == PTY_E_ARG
    assert _gd(lib, _problem(lib, N=0), _params(lib)) == PTY_E_ARG
    assert _gd(lib, _problem(lib, object_type=7), _params(lib)) == PTY_E_ARG


def test_small_workspace_is_rejected(lib):
    p, q = _problem(lib), _params(lib)
    need = lib.load().pty_gd_workspace_bytes(ctypes.byref(p), ctypes.byref(q))
    assert need > 0
    rc = _gd(lib, p, q, ws_bytes=need - 1)
    assert rc == PTY_E_WORKSPACE
    assert f"required {need}" in lib.last_error()
    with pytest.raises(RuntimeError, match="workspace"):
        lib.check(rc, "pty_gd_run")
    # the loss-gradient entry checks its workspace the same way
    L = lib.load()
    need = L.pty_loss_workspace_bytes(64, 256, 256, 100, 50)
    rc = L.pty_loss_gradients(ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), 256, 256,
                              ctypes.c_void_p(FAKE), 100, ctypes.c_void_p(FAKE),
                              ctypes.c_void_p(FAKE), 100, 50, 64, 1.0, ctypes.c_void_p(FAKE),
                              ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE),
                              need - 8, None)
    assert rc == PTY_E_WORKSPACE


def test_error_message_is_per_call(lib):
    _gd(lib, _problem(lib, S=48), _params(lib))
    first = lib.last_error()
    _gd(lib, _problem(lib, N=0), _params(lib))
    assert lib.last_error() != first


def test_sizes_beyond_32_bit_element_offsets_are_rejected(lib):
    # the kernels index the object and the staged frames with 32-bit offsets
    rc = _gd(lib, _problem(lib, Px=16384, Py=16385), _params(lib))
    assert rc == PTY_E_UNSUPPORTED and "2^28" in lib.last_error()
    rc = _gd(lib, _problem(lib, S=128, N=140000), _params(lib, max_batch=1000))
    assert rc == PTY_E_UNSUPPORTED and "frames" in lib.last_error()
    with pytest.raises(RuntimeError, match="2\\^31"):
        lib.call("pty_scatter_patches", ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE),
                 ctypes.c_void_p(FAKE), 200000, ctypes.c_void_p(FAKE), 140000, 128, 720, 720,
                 ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), ctypes.c_void_p(FAKE), None)


def test_pipelined_workspace_doubles_the_staging(lib):
    # P_shift and chi (8 bytes per pixel each) are double-buffered when the
    # stages of a mini-batch are pipelined; without stages nothing changes
    L = lib.load()
    p = _problem(lib, S=128, Px=720, Py=720, N=4096)
    q = _params(lib, max_batch=1024)
    q.stage = 256
    plain = L.pty_gd_workspace_bytes(ctypes.byref(p), ctypes.byref(q))
    q.pipeline = 1
    piped = L.pty_gd_workspace_bytes(ctypes.byref(p), ctypes.byref(q))
    assert piped - plain == 2 * 256 * 128 * 128 * 8
    q.stage = 0
    whole = L.pty_gd_workspace_bytes(ctypes.byref(p), ctypes.byref(q))
    q.pipeline = 0
    assert whole == L.pty_gd_workspace_bytes(ctypes.byref(p), ctypes.byref(q))
