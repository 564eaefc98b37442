"""Learned initialiser kernels on the GPU.

* MFMA convolutions (forward, backward data, backward weight) of every layer
  shape against float64 torch on the same bf16-rounded operands: relative
  L2 <= 1e-5 on the float32 outputs (the only difference is the float32
  accumulation order).
* Adam against its formula (float64): relative L2 <= 1e-6.
* Network input assembly (natural counts and kernel-layout amplitudes)
  against NumPy: within one bf16 rounding (|d| <= 2^-8 |ref| + 2^-8).
* Loss and its gradient against NumPy; stitching of exact patches restores
  the object on covered pixels (<= 1e-6).
* A short training run on forward-model data lowers the loss.
"""

import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


def to_bf16_bits(x):
    """float32 tensor -> int16 bf16 bit patterns (round to nearest even)."""
    u = x.contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    u = u + 0x7FFF + ((u >> 16) & 1)
    return ((u >> 16) & 0xFFFF).to(torch.int32).to(torch.int16)


def bf16_to_f64(h):
    u = (h.to(torch.int32) & 0xFFFF) << 16
    return u.view(torch.float32).to(torch.float64)


LAYERS = [(0, 2), (1, 2), (2, 2), (3, 3), (4, 1), (5, 2)]  # (layer index, samples)


@pytest.mark.parametrize("li,n", LAYERS)
def test_conv_forward_backward_match_torch(dev, li, n):
    from py4dstem_amd import _lib
    from py4dstem_amd.learned_init import ARCH, _hw, conv_desc
    from py4dstem_amd._lib import ptr, stream_handle
    ci, co, s, d, relu = ARCH[li]
    hi, ho = _hw(li)
    say(f"conv layer {li}: {ci}->{co} {hi}->{ho} stride {s} dil {d}")
    g = torch.Generator(device="cpu").manual_seed(li)
    x = to_bf16_bits(torch.randn(n, hi, hi, ci, generator=g)).to(dev)
    w = (torch.randn(co, 9, ci, generator=g) * 0.1).to(dev)
    b = torch.randn(co, generator=g).to(dev)
    desc = conv_desc(li, n)
    lib = _lib.load()
    wf = torch.zeros(lib.pty_conv_weight_words(desc, 0), dtype=torch.int16, device=dev)
    wb = torch.zeros(lib.pty_conv_weight_words(desc, 1), dtype=torch.int16, device=dev)
    _lib.call("pty_conv_pack_weights", desc, ptr(w), ptr(wf), ptr(wb), stream_handle())
    # float64 reference on the bf16-rounded operands (NCHW, OIHW)
    x64 = bf16_to_f64(x.cpu()).permute(0, 3, 1, 2).contiguous().requires_grad_(True)
    w64 = bf16_to_f64(to_bf16_bits(w.cpu())).reshape(co, 3, 3, ci).permute(0, 3, 1, 2).contiguous()
    w64.requires_grad_(True)
    ref = torch.nn.functional.conv2d(x64, w64, b.double().cpu(), stride=s, padding=d, dilation=d)
    y = torch.empty((n, ho, ho, co), dtype=torch.float32, device=dev)
    _lib.call("pty_conv_forward", desc, ptr(x), ptr(wf), ptr(b), 0, None, ptr(y), stream_handle())
    e_fwd = rel(y.permute(0, 3, 1, 2).double().cpu().numpy(), ref.detach().cpu().numpy())
    # backward with a bf16 upstream gradient
    dy = to_bf16_bits(torch.randn(n, ho, ho, co, generator=g)).to(dev)
    dy64 = bf16_to_f64(dy.cpu()).permute(0, 3, 1, 2)
    (ref * dy64).sum().backward()
    dx = torch.empty((n, hi, hi, ci), dtype=torch.float32, device=dev)
    _lib.call("pty_conv_backward_data", desc, ptr(dy), ptr(wb), None, None, ptr(dx), stream_handle())
    e_dx = rel(dx.permute(0, 3, 1, 2).double().cpu().numpy(), x64.grad.cpu().numpy())
    nbytes = lib.pty_conv_backward_weight_workspace_bytes(desc)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dw = torch.empty((co, 9, ci), dtype=torch.float32, device=dev)
    db = torch.empty(co, dtype=torch.float32, device=dev)
    _lib.call("pty_conv_backward_weight", desc, ptr(x), ptr(dy), ptr(dw), ptr(db), ptr(ws), nbytes,
              stream_handle())
    dw_ref = w64.grad.permute(0, 2, 3, 1).reshape(co, 9, ci)
    e_dw = rel(dw.double().cpu().numpy(), dw_ref.cpu().numpy())
    e_db = rel(db.double().cpu().numpy(), dy64.sum((0, 2, 3)).cpu().numpy())
    # ReLU mask of the layer input in the data gradient
    dxm = torch.empty_like(dx)
    _lib.call("pty_conv_backward_data", desc, ptr(dy), ptr(wb), ptr(x), None, ptr(dxm),
              stream_handle())
    mask = (bf16_to_f64(x.cpu()) > 0).float().to(dev)
    e_mask = rel(dxm.cpu().numpy(), (dx * mask).cpu().numpy())
    say(f"fwd {e_fwd:.1e} dx {e_dx:.1e} dw {e_dw:.1e} db {e_db:.1e} mask {e_mask:.1e}")
    assert e_fwd < 1e-5 and e_dx < 1e-5 and e_dw < 1e-5 and e_db < 1e-5 and e_mask < 1e-6


def test_adam_step_matches_formula(dev):
    from py4dstem_amd import _lib
    from py4dstem_amd.learned_init import conv_desc
    from py4dstem_amd._lib import ptr, stream_handle
    say("adam")
    g = torch.Generator(device="cpu").manual_seed(5)
    desc = conv_desc(2, 1)
    w0 = torch.randn(64, 9, 64, generator=g)
    m0 = torch.randn(64, 9, 64, generator=g) * 0.1
    v0 = torch.rand(64, 9, 64, generator=g) * 0.01
    gr = torch.randn(64, 9, 64, generator=g)
    b0, bm0, bv0, gb = (torch.randn(64, generator=g), torch.zeros(64), torch.zeros(64),
                        torch.randn(64, generator=g))
    t = [a.clone().to(dev) for a in (w0, m0, v0, gr, b0, bm0, bv0, gb)]
    lib = _lib.load()
    wf = torch.zeros(lib.pty_conv_weight_words(desc, 0), dtype=torch.int16, device=dev)
    wb = torch.zeros(lib.pty_conv_weight_words(desc, 1), dtype=torch.int16, device=dev)
    lr, b1, b2, eps, step = 1e-3, 0.9, 0.999, 1e-8, 3
    _lib.call("pty_conv_adam", desc, lr, b1, b2, eps, step, *[ptr(a) for a in t], ptr(wf), ptr(wb),
              stream_handle())
    m = b1 * m0.double() + (1 - b1) * gr.double()
    v = b2 * v0.double() + (1 - b2) * gr.double() ** 2
    w = w0.double() - lr * (m / (1 - b1 ** step)) / (torch.sqrt(v / (1 - b2 ** step)) + eps)
    assert rel(t[0].cpu().numpy(), w.numpy()) < 1e-6
    assert rel(t[1].cpu().numpy(), m.numpy()) < 1e-6
    # the packed bf16 copies hold the new weights
    wf_v = bf16_to_f64(wf.view(64, -1)[:, :576].reshape(64, 9, 64).cpu())
    assert rel(wf_v.numpy(), t[0].double().cpu().numpy()) < 4e-3


@pytest.mark.parametrize("layout", [0, 1])
def test_net_inputs_match_numpy(dev, layout):
    from py4dstem_amd import kernels as K
    from py4dstem_amd.learned_init import neighbour_table, net_inputs
    say(f"net inputs layout {layout}")
    rng = np.random.default_rng(3)
    S, grid = 128, (4, 5)
    N = grid[0] * grid[1]
    counts = rng.poisson(rng.uniform(0, 30, (N, S, S))).astype(np.float32)
    ij = np.array([[0, 0], [1, 2], [3, 4], [2, 1]])
    nbr = neighbour_table(grid, ij)
    if layout:
        frames = K.to_kernel_layout(torch.from_numpy(np.sqrt(counts)).to(dev))
    else:
        frames = torch.from_numpy(counts).to(dev)
    x = net_inputs(frames, layout, torch.from_numpy(nbr).to(dev))
    got = bf16_to_f64(x).cpu().numpy()
    c = np.fft.fftshift(counts, axes=(-2, -1)).astype(np.float64)
    binned = c.reshape(N, 64, 2, 64, 2).sum((2, 4))
    v = np.log1p(binned)[nbr]                              # [n, 9, 64, 64]
    v = v - v.mean(axis=(1, 2, 3), keepdims=True)
    ref = np.zeros((len(ij), 64, 64, 16))
    ref[..., :9] = v.transpose(0, 2, 3, 1)
    assert np.all(np.abs(got - ref) <= 2.0 ** -8 * np.abs(ref) + 2.0 ** -8)


def test_loss_and_stitch(dev):
    from py4dstem_amd import _lib
    from py4dstem_amd.learned_init import patch_targets, stitch
    from py4dstem_amd._lib import ptr, stream_handle
    say("loss and stitching")
    rng = np.random.default_rng(4)
    P = 160
    obj = np.exp(1j * rng.uniform(0, 0.5, (P, P))).astype(np.complex64)
    objs = torch.from_numpy(obj[None]).to(dev)
    gx, gy = np.meshgrid(np.arange(20, 141, 6), np.arange(20, 141, 6), indexing="ij")
    pos = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32) + 0.25
    posd = torch.from_numpy(pos).to(dev)
    n = len(pos)
    t = patch_targets(objs, posd, torch.zeros(n, dtype=torch.int64, device=dev))
    y = torch.zeros((n, 32, 32, 16), dtype=torch.float32, device=dev)
    y[..., :2] = t
    init = stitch(y, posd, (P, P)).cpu().numpy()
    covered = np.zeros((P, P), bool)
    covered[30:130, 30:130] = True
    assert rel(init[covered], obj[covered]) < 1e-6
    assert np.all(init[:2, :2] == 1)
    # loss: mean squared difference over the two channels, gradient 2 (y - t) / count
    y2 = y.clone()
    y2[..., :2] += torch.from_numpy(rng.standard_normal((n, 32, 32, 2)).astype(np.float32)).to(dev) * 0.1
    dy = torch.empty((n, 32, 32, 16), dtype=torch.int16, device=dev)
    loss = torch.empty(1, dtype=torch.float64, device=dev)
    nb = _lib.load().pty_net_loss_workspace_bytes(n)
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    _lib.call("pty_net_loss", ptr(y2), ptr(t), n, ptr(dy), ptr(loss), ptr(ws), nb, stream_handle())
    diff = (y2[..., :2] - t).double().cpu().numpy()
    assert abs(loss.item() / np.mean(diff ** 2) - 1) < 1e-9
    g = bf16_to_f64(dy).cpu().numpy()
    assert np.all(g[..., 2:] == 0)
    ref = 2 * diff / diff.size
    assert np.all(np.abs(g[..., :2] - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-30)


def test_short_training_lowers_loss(dev):
    from py4dstem_amd.learned_init import train
    say("short training run")
    net, hist = train(num_scans=64, scans_per_chunk=8, samples_per_chunk=512, batch=128, lr=2e-3,
                      seed=0, log=say)
    losses = [l for _, l in hist]
    assert all(np.isfinite(losses))
    assert losses[-1] < losses[0]


def test_judge_keeps_flat_start_when_learned_is_worse(dev):
    """The solve judges the learned start: a network whose amplitude output is
    0.5 (exit wave halved, data error up) must fall back to the flat start;
    the shipped model must be kept on an object of its training distribution."""
    from py4dstem_amd import SingleslicePtychography, _lib
    from py4dstem_amd._lib import ptr, stream_handle
    from py4dstem_amd.learned_init import ARCH, ConvNet, conv_desc, evaluation_workload
    say("initialiser judge / flat fallback")
    w = evaluation_workload(3001, scan=32, object_px=256)
    bad = ConvNet(dev, seed=1)
    last = bad.layers[-1]
    last["w"].zero_()
    last["b"].zero_()
    last["b"][0] = 0.5
    _lib.call("pty_conv_pack_weights", conv_desc(len(ARCH) - 1, 1), ptr(last["w"]), ptr(last["wf"]),
              ptr(last["wb"]), stream_handle())
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    flat = p._object_initial.copy()
    p.initialize_object_learned(bad)
    info = p.learned_init_info
    say(f"bad network: {info}")
    assert not info["used_learned"] and info["error_learned"] > info["error_flat"]
    np.testing.assert_array_equal(p._object_initial, flat)
    q = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    q.preprocess(**w["preprocess"])
    q.initialize_object_learned()
    say(f"shipped model: {q.learned_init_info}")
    assert q.learned_init_info["used_learned"]
    assert not np.array_equal(q._object_initial, flat)
    q.reconstruct(num_iter=1, reset=True, seed_random=0, **w["reconstruct"])
    assert np.isfinite(q.error)


@pytest.mark.parametrize("seed", [2001, 2003])
def test_learned_start_ablation_at_production_shapes(dev, seed):
    """Ablation on held-out objects of the initialiser's distribution at the
    production shapes (128x128 raster at 0.4 A, 128 px frames, 720x720
    object, 1e6 e/A^2, mini-batch 1024, probe update): the judged learned
    start has a lower object NMSE than the flat start, and reaches NMSE <=
    1e-3 in no more iterations than the flat start (both within 6)."""
    from py4dstem_amd.learned_init import ConvNet, evaluation_workload, time_to_nmse
    say(f"learned vs flat start, held-out seed {seed}")
    w = evaluation_workload(seed)
    r = time_to_nmse(w, ConvNet.load(), max_iter=6)
    fl, le = r["flat"], r["learned"]
    say(f"flat: nmse {[f'{v:.2e}' for v in fl['nmse'][:3]]} iterations {fl['iterations']} "
        f"seconds {fl['seconds']}; learned: nmse {[f'{v:.2e}' for v in le['nmse'][:3]]} "
        f"iterations {le['iterations']} seconds {le['seconds']} used {le['init_info']['used_learned']}")
    assert le["init_info"]["used_learned"]
    assert le["nmse"][0] < fl["nmse"][0]
    assert le["iterations"] is not None and fl["iterations"] is not None
    assert le["iterations"] <= fl["iterations"]
