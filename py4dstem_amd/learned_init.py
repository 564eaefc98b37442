"""Learned object initialiser on MFMA convolution kernels.

A six-layer 3x3 convolutional network (64 channels, BF16 weights and
activations on the MFMA units, FP32 accumulation and master weights) maps the
3x3 scan neighbourhood of a position's diffraction patterns (log counts,
binned 2x2 to 64x64, one channel per neighbour) to the 32x32 complex object
patch (amplitude, phase) centred on the position.  Predicted patches are
overlap-stitched (sin^2-weighted average, the deterministic scatter of the
reconstruction) into a starting object.  The iterative solve remains the
judge: ``initialize_object`` keeps the learned start only when its data error
is below the flat object's, otherwise the flat start is used.

Layers (Ci -> Co, stride, dilation): 9(16) -> 64 /2 d1 (64x64 -> 32x32),
64 -> 64 d1, d2, d4, d8 (receptive field spans the whole 32x32 patch),
64 -> 2(16) d1 linear.  ReLU after layers 1-5.

Training data come from the forward model itself (the HIP simulation kernel):
random Gaussian-atom objects (sigma 0.3 A) on random 2-4 A square lattices
with 10 % jitter, 10-40 % vacancies and peak phase 0.05-0.5 rad, scanned
32x32 at 0.4 A (4 px at 0.1 A), Poisson noise at 1e5-1e6 e/A^2.

There is no reference counterpart; the reference starts from a flat object
(py4DSTEM/process/phase/ptychographic_methods.py:40-79).
"""

from __future__ import annotations

import math
import os
import time

import numpy as np
import torch

from . import _lib
from . import kernels as K
from ._lib import PtyConvDesc, ptr, require_gpu, stream_handle

# (Ci, Co, stride, dilation, relu)
ARCH = ((16, 64, 2, 1, True), (64, 64, 1, 1, True), (64, 64, 1, 2, True),
        (64, 64, 1, 4, True), (64, 64, 1, 8, True), (64, 16, 1, 1, False))
IN_HW, OUT_HW = 64, 32
REAL_IN, REAL_OUT = 9, 2
PATCH = 32
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                             "learned_init_s128.npz")


def _hw(li):
    """(Hi, Ho) of layer li: the first layer halves 64 -> 32."""
    return (IN_HW, OUT_HW) if li == 0 else (OUT_HW, OUT_HW)


def conv_desc(li, n):
    ci, co, s, d, _ = ARCH[li]
    hi, ho = _hw(li)
    return PtyConvDesc(n, hi, hi, ci, ho, ho, co, s, d)


def _bf16(shape, dev):
    return torch.empty(shape, dtype=torch.int16, device=dev)


def _call(name, *args):
    _lib.call(name, *args)


class ConvNet:
    """Weights, Adam state and the kernel calls of the network."""

    def __init__(self, device=None, seed=0, state=None):
        self.dev = device if device is not None else require_gpu()
        self.layers = []
        g = torch.Generator().manual_seed(int(seed))
        for li, (ci, co, _, _, _) in enumerate(ARCH):
            ci_real = REAL_IN if li == 0 else ci
            co_real = REAL_OUT if li == len(ARCH) - 1 else co
            if state is not None:
                w = torch.from_numpy(np.asarray(state[f"w{li}"], np.float32))
                b = torch.from_numpy(np.asarray(state[f"b{li}"], np.float32))
            else:
                w = torch.zeros(co, 9, ci)
                std = math.sqrt(2.0 / (9 * ci_real)) * (0.1 if li == len(ARCH) - 1 else 1.0)
                w[:co_real, :, :ci_real] = torch.randn(co_real, 9, ci_real, generator=g) * std
                b = torch.zeros(co)
                if li == len(ARCH) - 1:
                    b[0] = 1.0  # amplitude of a pure phase object
            L = {"w": w.to(self.dev).contiguous(), "b": b.to(self.dev).contiguous()}
            for k in ("m", "v"):
                L[k] = torch.zeros_like(L["w"])
                L["b" + k] = torch.zeros_like(L["b"])
            d1 = conv_desc(li, 1)
            L["wf"] = torch.zeros(_lib.load().pty_conv_weight_words(d1, 0), dtype=torch.int16,
                                  device=self.dev)
            L["wb"] = torch.zeros(_lib.load().pty_conv_weight_words(d1, 1), dtype=torch.int16,
                                  device=self.dev)
            _call("pty_conv_pack_weights", d1, ptr(L["w"]), ptr(L["wf"]), ptr(L["wb"]),
                  stream_handle())
            self.layers.append(L)
        self.step = 0
        self._ws = None

    # -- persistence ----------------------------------------------------------
    def state(self):
        out = {}
        for li, L in enumerate(self.layers):
            out[f"w{li}"] = L["w"].cpu().numpy()
            out[f"b{li}"] = L["b"].cpu().numpy()
        return out

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, **self.state(), **{f"meta_{k}": np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=DEFAULT_MODEL, device=None):
        with np.load(path, allow_pickle=False) as z:
            state = {k: z[k] for k in z.files if not k.startswith("meta_")}
        return cls(device=device, state=state)

    # -- forward / backward -----------------------------------------------------
    def forward(self, x, keep=True):
        """x: bf16 [n, 64, 64, 16].  Returns the list of layer inputs/outputs
        (keep=True; last entry the f32 output [n, 32, 32, 16]) or just the
        output."""
        n = x.shape[0]
        acts = [x]
        h = x
        last = len(ARCH) - 1
        for li, L in enumerate(self.layers):
            ci, co, _, _, relu = ARCH[li]
            _, ho = _hw(li)
            d = conv_desc(li, n)
            if li == last:
                y = torch.empty((n, ho, ho, co), dtype=torch.float32, device=self.dev)
                _call("pty_conv_forward", d, ptr(h), ptr(L["wf"]), ptr(L["b"]), int(relu), None,
                      ptr(y), stream_handle())
            else:
                y = _bf16((n, ho, ho, co), self.dev)
                _call("pty_conv_forward", d, ptr(h), ptr(L["wf"]), ptr(L["b"]), int(relu), ptr(y),
                      None, stream_handle())
            if keep:
                acts.append(y)
            h = y
        return acts if keep else h

    def _workspace(self, n):
        need = max(_lib.load().pty_conv_backward_weight_workspace_bytes(conv_desc(li, n))
                   for li in range(len(ARCH)))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.dev)
        return self._ws

    def backward(self, acts, dy):
        """Gradients [(dw, db)] per layer from dy (bf16 [n, 32, 32, 16])."""
        n = dy.shape[0]
        ws = self._workspace(n)
        grads = [None] * len(ARCH)
        for li in range(len(ARCH) - 1, -1, -1):
            L = self.layers[li]
            d = conv_desc(li, n)
            dw = torch.empty_like(L["w"])
            db = torch.empty_like(L["b"])
            _call("pty_conv_backward_weight", d, ptr(acts[li]), ptr(dy), ptr(dw), ptr(db), ptr(ws),
                  ws.numel(), stream_handle())
            grads[li] = (dw, db)
            if li > 0:
                ci = ARCH[li][0]
                hi, _ = _hw(li)
                dx = _bf16((n, hi, hi, ci), self.dev)
                _call("pty_conv_backward_data", d, ptr(dy), ptr(L["wb"]), ptr(acts[li]), ptr(dx),
                      None, stream_handle())
                dy = dx
        return grads

    def adam(self, grads, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        for li, (L, (dw, db)) in enumerate(zip(self.layers, grads)):
            _call("pty_conv_adam", conv_desc(li, 1), float(lr), float(beta1), float(beta2),
                  float(eps), self.step, ptr(L["w"]), ptr(L["m"]), ptr(L["v"]), ptr(dw), ptr(L["b"]),
                  ptr(L["bm"]), ptr(L["bv"]), ptr(db), ptr(L["wf"]), ptr(L["wb"]), stream_handle())

    def loss_and_grad(self, y, target):
        n = y.shape[0]
        dy = _bf16((n, OUT_HW, OUT_HW, 16), self.dev)
        loss = torch.empty(1, dtype=torch.float64, device=self.dev)
        nbytes = _lib.load().pty_net_loss_workspace_bytes(n)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        _call("pty_net_loss", ptr(y), ptr(target), n, ptr(dy), ptr(loss), ptr(ws), nbytes,
              stream_handle())
        return loss, dy

    def train_step(self, x, target, lr):
        acts = self.forward(x, keep=True)
        loss, dy = self.loss_and_grad(acts[-1], target)
        self.adam(self.backward(acts, dy), lr)
        return loss

    def predict(self, x, chunk=1024):
        """f32 [n, 32, 32, 16] in chunks (inference keeps no activations)."""
        outs = [self.forward(x[s:s + chunk], keep=False) for s in range(0, x.shape[0], chunk)]
        return outs[0] if len(outs) == 1 else torch.cat(outs)


# -- inputs, targets, stitching ------------------------------------------------
def neighbour_table(grid_shape, centres_ij, base=0):
    """int32 [n, 9]: raster indices of the 3x3 neighbourhood (clamped at the
    scan edges) of grid positions centres_ij [n, 2]; raster index = i*ny + j."""
    nx, ny = grid_shape
    ci = np.asarray(centres_ij)
    base = np.asarray(base)
    out = np.empty((len(ci), 9), np.int32)
    k = 0
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            ii = np.clip(ci[:, 0] + di, 0, nx - 1)
            jj = np.clip(ci[:, 1] + dj, 0, ny - 1)
            out[:, k] = base + ii * ny + jj
            k += 1
    return out


def net_inputs(frames, kernel_layout, nbr):
    """bf16 [n, 64, 64, 16] from frames [N, S, S] (counts natural order, or
    kernel-layout amplitudes) and the neighbour table (device int32 [n, 9])."""
    N, S, _ = frames.shape
    n = nbr.shape[0]
    x = _bf16((n, IN_HW, IN_HW, 16), frames.device)
    nbytes = _lib.load().pty_net_inputs_workspace_bytes(n)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=frames.device)
    _call("pty_net_inputs", ptr(frames), N, S, int(kernel_layout), ptr(nbr), n, ptr(x), ptr(ws),
          nbytes, stream_handle())
    return x


def patch_targets(objects, positions, obj_index):
    """(amplitude, phase) [n, 32, 32, 2] of the object patch centred on each
    rounded position; objects [M, Px, Py] complex64, obj_index [n]."""
    n = positions.shape[0]
    Px, Py = objects.shape[-2:]
    r = torch.round(positions).to(torch.int64)  # round half to even, like the kernels
    off = torch.arange(PATCH, device=positions.device) - PATCH // 2
    rows = (r[:, 0, None] + off[None, :]) % Px
    cols = (r[:, 1, None] + off[None, :]) % Py
    o = objects[obj_index[:, None, None], rows[:, :, None], cols[:, None, :]]
    return torch.stack((o.abs(), torch.angle(o)), dim=-1).to(torch.float32).contiguous()


def stitch(y, positions, object_shape, floor_weight=0.5):
    """Overlap-stitch predicted patches at `positions` into an object
    (complex64 device); uncovered pixels are 1."""
    n = y.shape[0]
    dev = y.device
    ps = torch.empty((n, PATCH, PATCH), dtype=torch.complex64, device=dev)
    chi = torch.empty_like(ps)
    _call("pty_net_patches", ptr(y), n, ptr(ps), ptr(chi), stream_handle())
    order = torch.arange(n, dtype=torch.int32, device=dev)
    grad, wsum, _ = K.scatter(ps, chi, positions.contiguous(), order, object_shape)
    obj = torch.empty(object_shape, dtype=torch.complex64, device=dev)
    _call("pty_net_finish", ptr(grad), ptr(wsum), object_shape[0], object_shape[1],
          float(floor_weight), ptr(obj), stream_handle())
    return obj


# -- training data from the forward model ----------------------------------------
class TrainingScans:
    """Random Gaussian-atom phase objects and their simulated, Poisson-noisy
    32x32 scans (128 px frames at 0.1 A, 300 keV, 20 mrad, defocus 200 A)."""

    def __init__(self, device=None, seed=0, S=128, scan=32, step_px=4, object_px=256,
                 sampling=0.1, energy=300e3, semiangle=20.0, defocus=200.0,
                 dose_range=(1e5, 1e6), lattice_range=(2.0, 4.0), vacancy_range=(0.1, 0.4),
                 phase_range=(0.05, 0.5), sigma_ang=0.3):
        from .synthetic import scaled_probe
        self.dev = device if device is not None else require_gpu()
        self.rng = np.random.default_rng(seed)
        self.S, self.scan, self.step_px, self.P = S, scan, step_px, object_px
        self.sampling = sampling
        self.dose_range, self.lattice_range = dose_range, lattice_range
        self.vacancy_range, self.phase_range = vacancy_range, phase_range
        self.sigma_px = sigma_ang / sampling
        probe = scaled_probe(energy, S, (sampling, sampling), semiangle, 2.0, 1.0, defocus=defocus)
        self.probe = torch.from_numpy(probe.astype(np.complex64)).to(self.dev)
        self.phat = K.probe_spectrum(self.probe)
        ext = (scan - 1) * step_px
        g = np.arange(scan) * step_px + (object_px - ext) / 2.0
        gx, gy = np.meshgrid(g, g, indexing="ij")
        self.grid = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32)
        k = np.fft.fftfreq(object_px)
        self.gauss_k = torch.from_numpy(
            np.exp(-2 * (np.pi * self.sigma_px) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
            * 2 * np.pi * self.sigma_px ** 2).to(self.dev)

    def objects(self, n):
        """n random objects [n, P, P] complex64 (|O| = 1)."""
        P = self.P
        splat = np.zeros((n, P, P), np.float64)
        for s in range(n):
            a = self.rng.uniform(*self.lattice_range) / self.sampling
            off = self.rng.uniform(0, a, 2)
            xs = np.arange(off[0], P, a)
            ys = np.arange(off[1], P, a)
            cx, cy = np.meshgrid(xs, ys, indexing="ij")
            c = np.stack([cx.ravel(), cy.ravel()], 1)
            c += self.rng.uniform(-0.1, 0.1, c.shape) * a
            c = c[self.rng.uniform(size=len(c)) >= self.rng.uniform(*self.vacancy_range)]
            peak = self.rng.uniform(*self.phase_range)
            i0 = np.floor(c).astype(int)
            f = c - i0
            for di in (0, 1):
                for dj in (0, 1):
                    wgt = (f[:, 0] if di else 1 - f[:, 0]) * (f[:, 1] if dj else 1 - f[:, 1])
                    np.add.at(splat[s], ((i0[:, 0] + di) % P, (i0[:, 1] + dj) % P), peak * wgt)
        sp = torch.from_numpy(splat).to(self.dev)
        # gauss_k is the transform of exp(-r^2 / 2 sigma^2) (peak 1): each
        # atom adds `peak` rad at its centre
        phase = torch.fft.ifft2(torch.fft.fft2(sp) * self.gauss_k).real
        return torch.polar(torch.ones_like(phase), phase).to(torch.complex64)

    def chunk(self, n_scans):
        """(counts [n*scan^2, S, S] natural order, positions [n, scan^2, 2],
        objects [n, P, P])."""
        objs = self.objects(n_scans)
        npos = self.scan * self.scan
        frames = torch.empty((n_scans * npos, self.S, self.S), dtype=torch.float32, device=self.dev)
        positions = np.empty((n_scans, npos, 2), np.float32)
        for s in range(n_scans):
            positions[s] = self.grid + self.rng.uniform(0, 1, 2).astype(np.float32)
            pos = torch.from_numpy(positions[s]).to(self.dev)
            out = frames[s * npos:(s + 1) * npos]
            _call("pty_simulate_intensities", ptr(self.phat), ptr(objs[s].contiguous()), self.P,
                  self.P, ptr(pos), npos, self.S, ptr(out), stream_handle())
            dose = self.rng.uniform(*self.dose_range) * (self.step_px * self.sampling) ** 2
            out.mul_(float(dose))
        frames = torch.poisson(frames)
        return frames, torch.from_numpy(positions).to(self.dev), objs

    def samples(self, frames, positions, objs, count):
        """Network inputs and targets of `count` random positions of a chunk."""
        n_scans, npos = positions.shape[:2]
        pick = self.rng.integers(0, n_scans * npos, count)
        scan_of = pick // npos
        local = pick % npos
        ij = np.stack([local // self.scan, local % self.scan], 1)
        nbr = neighbour_table((self.scan, self.scan), ij, base=scan_of * npos)
        x = net_inputs(frames, 0, torch.from_numpy(nbr).to(self.dev))
        pos = positions.reshape(-1, 2)[torch.from_numpy(pick).to(self.dev)]
        t = patch_targets(objs, pos, torch.from_numpy(scan_of).to(self.dev))
        return x, t


def train(num_scans=20000, scans_per_chunk=16, samples_per_chunk=1024, batch=256, lr=1e-3,
          seed=0, log=None, time_limit=None, device=None):
    """Train a network from scratch on `num_scans` simulated scans; returns
    (net, history [(step, loss)])."""
    dev = device if device is not None else require_gpu()
    net = ConvNet(dev, seed=seed)
    data = TrainingScans(dev, seed=seed + 1)
    n_chunks = max(1, num_scans // scans_per_chunk)
    total_steps = n_chunks * max(1, samples_per_chunk // batch)
    hist = []
    t0 = time.perf_counter()
    step = 0
    for c in range(n_chunks):
        frames, pos, objs = data.chunk(scans_per_chunk)
        x, t = data.samples(frames, pos, objs, samples_per_chunk)
        del frames
        for b0 in range(0, samples_per_chunk - batch + 1, batch):
            rate = lr * 0.5 * (1 + math.cos(math.pi * step / total_steps))
            loss = net.train_step(x[b0:b0 + batch], t[b0:b0 + batch], rate)
            step += 1
        if c % 50 == 0 or c == n_chunks - 1:
            hist.append((step, float(loss.item())))
            if log:
                log(f"chunk {c + 1}/{n_chunks} step {step} loss {hist[-1][1]:.4e} "
                    f"{time.perf_counter() - t0:.1f} s")
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            if log:
                log(f"time limit reached after {c + 1} chunks")
            break
    torch.cuda.synchronize()
    return net, hist


# -- initialising a reconstruction -------------------------------------------------
def _data_errors(p, objects, probe, order):
    """Data error sum_b sum (|F_b| - A_b)^2 over `order` for each object (one
    workspace, one synchronisation)."""
    S = probe.shape[-1]
    Px, Py = objects[0].shape
    n = order.numel()
    ws_bytes = _lib.load().pty_loss_workspace_bytes(S, Px, Py, n, n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=probe.device)
    losses = torch.empty(len(objects), dtype=torch.float64, device=probe.device)
    for i, obj in enumerate(objects):
        _call("pty_loss_gradients", ptr(obj), ptr(probe), Px, Py, ptr(p._positions_dev),
              p._positions_dev.shape[0], ptr(p._amplitudes_dev), ptr(order), n, n, S, 1.0,
              ptr(losses[i:i + 1]), None, None, ptr(ws), ws.numel(), stream_handle())
    return losses.cpu().numpy().tolist()


_TABLES = {}


def _grid_tables(nx, ny, stride, n_judge, seed, dev):
    """Device tables of a raster scan, kept per (scan, stride, judge, device):
    raster indices of the network's patch centres (every `stride`-th grid
    point per axis plus the last), their 3x3 neighbour tables, and the
    judge's seeded positions."""
    key = (nx, ny, stride, n_judge, seed, str(dev))
    hit = _TABLES.get(key)
    if hit is None:
        gi = np.unique(np.r_[np.arange(0, nx, stride), nx - 1])
        gj = np.unique(np.r_[np.arange(0, ny, stride), ny - 1])
        ii, jj = np.meshgrid(gi, gj, indexing="ij")
        ij = np.stack([ii.ravel(), jj.ravel()], 1)
        centres = torch.from_numpy((ij[:, 0] * ny + ij[:, 1]).astype(np.int64)).to(dev)
        nbr = torch.from_numpy(neighbour_table((nx, ny), ij)).to(dev)
        rng = np.random.default_rng(seed)
        order = torch.from_numpy(rng.choice(nx * ny, n_judge, replace=False).astype(np.int32)).to(dev)
        hit = (centres, nbr, order, int(len(ij)))
        _TABLES[key] = hit
    return hit


def initialize_object(p, net=None, stride=4, judge_positions=512, seed=0):
    """Learned starting object for a preprocessed SingleslicePtychography
    with a raster scan: network on every `stride`-th position per scan axis,
    stitched; kept only if its data error over `judge_positions` seeded
    positions is below the flat object's.  Returns (object (complex64
    device), used_learned, info)."""
    if net is None:
        net = ConvNet.load()
    dev = p._positions_dev.device
    S = int(p._region_of_interest_shape[0])
    if S not in (64, 128):
        raise ValueError(f"learned initialiser needs 64 or 128 px frames, not {S}")
    nx, ny = (int(v) for v in p._grid_scan_shape)
    if nx * ny != p._num_diffraction_patterns:
        raise ValueError("learned initialiser needs a raster scan (one pattern per grid point)")
    t0 = time.perf_counter()
    n = min(judge_positions, p._num_diffraction_patterns)
    centres, nbr, order, npatch = _grid_tables(nx, ny, stride, n, seed, dev)
    x = net_inputs(p._amplitudes_dev, 1, nbr)
    y = net.predict(x, chunk=8192)
    learned = stitch(y, p._positions_dev[centres], tuple(p._object_shape))
    probe = torch.from_numpy(np.ascontiguousarray(p._probe_initial, dtype=np.complex64)).to(dev)
    flat = getattr(p, "_object_initial_dev", None)
    if flat is None:
        flat = torch.from_numpy(np.ascontiguousarray(p._object_initial, dtype=np.complex64)).to(dev)
    e_learned, e_flat = _data_errors(p, (learned, flat), probe, order)
    use = e_learned < e_flat
    info = {"seconds": time.perf_counter() - t0, "error_learned": e_learned, "error_flat": e_flat,
            "patches": npatch, "used_learned": bool(use)}
    return (learned if use else flat), bool(use), info


# -- evaluation at production shapes -------------------------------------------------
def evaluation_workload(seed, device=None, scan=128, object_px=720, dose=1e6, step_px=4):
    """A held-out object from the training distribution at production shapes
    (128x128 raster at 0.4 A, 128 px frames at 0.1 A, 720^2 object), simulated
    on the GPU with Poisson noise at `dose` e/A^2.  Same dictionary layout as
    py4dstem_amd.synthetic.production_workload."""
    from .datacube import Calibration, DataCube
    from .preprocess import center_positions

    dev = device if device is not None else require_gpu()
    gen = TrainingScans(dev, seed=seed, object_px=object_px)
    obj = gen.objects(1)[0].contiguous()
    g = np.arange(scan) * step_px
    gx, gy = np.meshgrid(g, g, indexing="ij")
    pos = center_positions(np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32) + 64.0,
                           (object_px, object_px))
    posd = torch.from_numpy(pos).to(dev)
    N, S = len(pos), gen.S
    counts = torch.empty((N, S, S), dtype=torch.float32, device=dev)
    for s0 in range(0, N, 4096):
        n = min(4096, N - s0)
        _call("pty_simulate_intensities", ptr(gen.phat), ptr(obj), object_px, object_px,
              ptr(posd[s0:s0 + n]), n, S, ptr(counts[s0:s0 + n]), stream_handle())
    counts.mul_(float(dose * (step_px * gen.sampling) ** 2))
    torch.manual_seed(seed)
    counts = torch.poisson(counts)
    data = torch.fft.fftshift(counts, dim=(-2, -1)).cpu().numpy().reshape(scan, scan, S, S)
    del counts
    dk = 1.0 / (S * gen.sampling)
    cal = Calibration(R_pixel_size=step_px * gen.sampling, R_pixel_units="A",
                      Q_pixel_size=dk, Q_pixel_units="A^-1")
    return {
        "datacube": DataCube(data, cal), "object_true": obj.cpu().numpy(),
        "init": dict(energy=300e3, semiangle_cutoff=20.0, rolloff=2.0, defocus=200.0,
                     initial_object_guess=np.ones((object_px, object_px), np.complex64),
                     verbose=False),
        "preprocess": dict(force_com_rotation=0, force_com_transpose=False,
                           force_com_shifts=(S / 2, S / 2), plot_center_of_mass=False,
                           plot_rotation=False, plot_probe_overlaps=False),
        "reconstruct": dict(max_batch_size=1024, fix_probe=False, step_size=0.5,
                            normalization_min=1, progress_bar=False),
    }


def time_to_nmse(w, net=None, max_iter=6, target=1e-3, seed=0):
    """Flat against learned start on one workload.

    For each start: the object NMSE (scanned field) before the first
    iteration and after each of `max_iter` iterations of one reconstruct()
    call (store_iterations keeps the per-iteration objects), the first
    iteration count at which NMSE <= target, and the wall time to reach it:
    initialiser time + that many iterations at the per-iteration time of one
    uninterrupted `max_iter`-iteration reconstruct() call (no per-iteration
    downloads; the call's fixed cost spread over its iterations).  The
    learned start includes its whole initialiser (network inputs, network,
    stitching and the data-error judge)."""
    from .ptychography import SingleslicePtychography
    from .synthetic import object_nmse

    out = {}
    # one untimed pass so that kernel code objects are loaded before timing
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    p.initialize_object_learned(net)
    p.reconstruct(num_iter=2, reset=True, seed_random=seed, **w["reconstruct"])
    torch.cuda.synchronize()
    for kind in ("flat", "learned"):
        p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
        p.preprocess(**w["preprocess"])
        t_init = 0.0
        info = None
        if kind == "learned":
            torch.cuda.synchronize()
            t = time.perf_counter()
            p.initialize_object_learned(net)
            torch.cuda.synchronize()
            t_init = time.perf_counter() - t
            info = p.learned_init_info
        mask = p._object_fov_mask
        nmse = [object_nmse(p._object, w["object_true"], mask)]
        p.reconstruct(num_iter=max_iter, reset=True, seed_random=seed, store_iterations=True,
                      **w["reconstruct"])
        nmse += [object_nmse(o, w["object_true"], mask) for o in p.object_iterations]
        torch.cuda.synchronize()
        t = time.perf_counter()
        p.reconstruct(num_iter=max_iter, reset=True, seed_random=seed, **w["reconstruct"])
        torch.cuda.synchronize()
        t_iter = (time.perf_counter() - t) / max_iter
        hit = next((k for k, v in enumerate(nmse) if v <= target), None)
        out[kind] = {"nmse": nmse, "iterations": hit, "s_per_iteration": t_iter,
                     "init_seconds": t_init,
                     "seconds": None if hit is None else t_init + hit * t_iter, "init_info": info}
    return out


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train / evaluate the learned object initialiser")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scans", type=int, default=20000)
    t.add_argument("--scans-per-chunk", type=int, default=16)
    t.add_argument("--samples-per-chunk", type=int, default=1024)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--time-limit", type=float, default=None)
    t.add_argument("--out", default=DEFAULT_MODEL)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--seeds", type=int, nargs="+", default=[1001, 1002])
    e.add_argument("--max-iter", type=int, default=12)
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    def log(msg):
        print(f"[learned-init] {msg}", flush=True)

    if a.cmd == "train":
        t0 = time.perf_counter()
        net, hist = train(num_scans=a.scans, scans_per_chunk=a.scans_per_chunk,
                          samples_per_chunk=a.samples_per_chunk, batch=a.batch, lr=a.lr,
                          seed=a.seed, log=log, time_limit=a.time_limit)
        net.save(a.out, scans=a.scans, seed=a.seed, steps=net.step,
                 seconds=time.perf_counter() - t0, loss_history=np.asarray(hist))
        log(f"saved {a.out} after {net.step} steps, {time.perf_counter() - t0:.1f} s")
    else:
        net = ConvNet.load(a.model)
        res = {}
        for s in a.seeds:
            log(f"evaluation seed {s}")
            w = evaluation_workload(s)
            res[s] = time_to_nmse(w, net, max_iter=a.max_iter)
            for kind, r in res[s].items():
                log(f"  {kind}: iterations {r['iterations']} seconds {r['seconds']} "
                    f"nmse {[f'{v:.2e}' for v in r['nmse'][:6]]} info {r['init_info']}")
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1, default=float)


if __name__ == "__main__":
    main()
