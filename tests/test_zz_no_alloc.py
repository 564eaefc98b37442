"""pty_gd_run_noalloc (static per-GPU workspace) against pty_gd_run (caller
workspace), both through the C-ABI, on two calls in a row: the first at
128 px frames with a mini-batch whose staging does not fit the 64 MiB
scratch (the entry point lowers the positions per launch, which leaves the
update bitwise unchanged), the second a smaller 64 px problem that reuses
the scratch left by the first.  Object, probe and every error must be
bit-identical; a problem too large for the scratch is rejected with
PTY_E_WORKSPACE."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PTY_E_WORKSPACE = -4


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    from py4dstem_amd import _lib
    _lib.load()
    return torch.device("cuda", 0)


def _prepared(w, max_batch, num_iter):
    from py4dstem_amd.ptychography import SingleslicePtychography
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    run = p._prepare_run(num_iter=num_iter, max_batch_size=max_batch, fix_probe=False,
                         seed_random=4, progress_bar=False, _allocate=False)
    return p, run


def _both(dev, w, max_batch, num_iter):
    """(object, probe, errors) from pty_gd_run and from pty_gd_run_noalloc on
    identical copies of one prepared state."""
    from py4dstem_amd import _lib
    from py4dstem_amd import kernels as K
    p, run = _prepared(w, max_batch, num_iter)
    params = run["params"]
    params.num_iter = num_iter
    out = []
    for entry in ("pty_gd_run", "pty_gd_run_noalloc"):
        obj = p._object_dev.clone()
        probe = p._probe_dev.clone()
        problem = K.make_problem(obj, probe, p._positions_dev, p._amplitudes_dev, run["mask"],
                                 p._mean_diffraction_intensity, None)
        errors = torch.zeros(num_iter, dtype=torch.float64, device=dev)
        if entry == "pty_gd_run":
            nbytes = K.gd_workspace_bytes(problem, params)
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            _lib.call(entry, ctypes.byref(problem), ctypes.byref(params), _lib.ptr(run["orders"]),
                      _lib.ptr(errors), _lib.ptr(ws), nbytes, _lib.stream_handle())
        else:
            _lib.call(entry, ctypes.byref(problem), ctypes.byref(params), _lib.ptr(run["orders"]),
                      _lib.ptr(errors), _lib.stream_handle())
        torch.cuda.synchronize()
        out.append((obj.cpu().numpy(), probe.cpu().numpy(), errors.cpu().numpy()))
    return out, params, problem


def _check(pair, what):
    (o1, p1, e1), (o2, p2, e2) = pair
    say(f"{what}: object max |diff| {np.abs(o1 - o2).max():.2e}, errors {e1[-1]:.6e} {e2[-1]:.6e}")
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(e1, e2)
    assert np.all(np.isfinite(e1))


def test_noalloc_equals_caller_workspace_two_calls(dev):
    from py4dstem_amd import kernels as K
    from py4dstem_amd.synthetic import production_workload, small_workload
    say("no-alloc: 128 px, 40x40 scan, 360^2 object, mini-batch 800")
    w1 = production_workload(seed=3, scan=(40, 40), object_px=360, num_iter=2)
    pair, params, problem = _both(dev, w1, 800, 2)
    need_whole = K.gd_workspace_bytes(problem, params)
    assert need_whole > (64 << 20)   # the first call had to stage
    _check(pair, "first call")
    say("no-alloc: 64 px, 12x12 scan, 128^2 object, mini-batch 50 (second, smaller call)")
    w2 = small_workload(seed=4, scan=(12, 12), object_px=128)
    pair, _, _ = _both(dev, w2, 50, 3)
    _check(pair, "second call")


def test_noalloc_rejects_problems_beyond_the_scratch(dev):
    from py4dstem_amd import _lib
    from py4dstem_amd import kernels as K
    # a 2000^2 object needs 96 MB of running sums alone; nothing is launched
    obj = torch.ones((2000, 2000), dtype=torch.complex64, device=dev)
    probe = torch.ones((128, 128), dtype=torch.complex64, device=dev)
    pos = torch.full((2000, 2), 500.0, dtype=torch.float32, device=dev)
    amps = torch.zeros((1, 128, 128), dtype=torch.float32, device=dev)
    problem = K.make_problem(obj, probe, pos, amps, None, 1.0, None)
    params = _lib.PtyGdParams()
    params.num_iter, params.max_batch, params.step_size, params.normalization_min = 1, 1024, 0.5, 1.0
    orders = torch.zeros(2000, dtype=torch.int32, device=dev)
    errors = torch.zeros(1, dtype=torch.float64, device=dev)
    rc = _lib.load().pty_gd_run_noalloc(ctypes.byref(problem), ctypes.byref(params),
                                        _lib.ptr(orders), _lib.ptr(errors), _lib.stream_handle())
    assert rc == PTY_E_WORKSPACE
    assert "static scratch" in _lib.last_error()
