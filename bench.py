"""Benchmark: single-slice GD ptychography position-updates/s on one MI355X.

Workload (synthetic, seeded): 128x128 scan at 0.4 A, 128x128-pixel float32
diffraction patterns (Poisson, 1.6e5 counts each), 720x720 complex64 object,
300 keV / 20 mrad / defocus 200 A probe, mini-batches of 1024 positions in
seeded order, probe update on.  One step = one full GD iteration (16,384
position updates + constraints) of ``reconstruct()``, amplitudes resident in
HBM.  Prints one JSON line.

``--workload many`` instead times 64 independent seeded reconstructions
(square scans 32..96 positions per side, half 64 px and half 128 px frames,
defocus 100..400 A, 30 iterations, mini-batch and stage size from the
autotuner) advanced together by MultiPtychography (one launch per stage for
all problems, one graph replay per iteration) against the same solves run one
after another with reconstruct(); prints aggregate position-updates/s.

``--workload positions`` times the production workload with position
correction on (reconstruct(fix_positions=False)); ``--workload projection``
with the DM_AP(0.5) projection-set method (all positions per update);
``--workload mixed`` as a mixed-state reconstruction with ``--probe-modes``
probe modes.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds T]
                       [--workload production|many|kernels|learned-init|positions|projection|
                                   mixed] [--probe-modes K]
                       [--problems P]
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_update(S, Px, Py, B):
    """20 S^2 (amplitude read, object-patch gather, object-gradient scatter)
    + object and probe read+written once per mini-batch, amortised."""
    return 20 * S * S + 16 * (Px * Py + S * S) / B


COUNTERS = os.path.join(ROOT, "profiles", "counters_gd_minibatch_scan128_frame128.md")


def measured_traffic(key):
    """HBM bytes of `key` (a kernel name: per launch; "minibatch": all kernels
    of one 1024-position mini-batch) from the committed counter summary
    (FETCH_SIZE / WRITE_SIZE passes scaled by a copy-kernel calibration), or
    None when the summary is absent."""
    try:
        for line in open(COUNTERS):
            if line.startswith(f"traffic_bytes {key} = "):
                return int(line.rsplit("=", 1)[1])
    except OSError:
        pass
    return None


def learned_init_summary(seeds, max_iter):
    """Time to object NMSE <= 1e-3 from the learned and from the flat start
    (learned_init.time_to_nmse) on held-out seeds at the production shapes;
    value = total flat seconds / total learned seconds."""
    from py4dstem_amd import learned_init as li

    net = li.ConvNet.load()
    per = {}
    for s in seeds:
        t = time.perf_counter()
        r = li.time_to_nmse(li.evaluation_workload(s), net, max_iter=max_iter)
        per[s] = r
        log(f"learned-init seed {s} ({time.perf_counter() - t:.1f} s): "
            + "; ".join(f"{k} nmse0 {r[k]['nmse'][0]:.2e} iterations {r[k]['iterations']} "
                        f"seconds {r[k]['seconds']}" for k in ("flat", "learned")))
    def total(kind):
        v = [per[s][kind]["seconds"] for s in seeds]
        return None if any(x is None for x in v) else sum(v)
    t_flat, t_learned = total("flat"), total("learned")
    # seeds whose flat start is not yet at the target (there the initialiser
    # can save iterations; elsewhere it can only add its own time)
    miss = [s for s in seeds if per[s]["flat"]["iterations"] and per[s]["learned"]["seconds"]]
    ratios = sorted(per[s]["flat"]["seconds"] / per[s]["learned"]["seconds"] for s in miss)
    return {
        "metric": "time to object NMSE<=1e-3, flat start / learned start",
        "value": round(t_flat / t_learned, 3) if (t_flat and t_learned) else None,
        "unit": "x", "target": 2.0,
        "flat_seconds": t_flat, "learned_seconds": t_learned,
        "where_flat_start_misses_target": {"seeds": miss,
                                           "ratios": [round(r, 3) for r in ratios]},
        "workload": "held-out Gaussian-atom objects of the initialiser's distribution, 128x128 "
                    "scan at 0.4 A, 128 px frames, 720x720 object, 1e6 e/A^2, mini-batch 1024, "
                    "probe update; seconds = initialiser + iterations to target x per-iteration "
                    "time of one uninterrupted multi-iteration reconstruct()",
        "per_seed": {str(s): {k: {"iterations": per[s][k]["iterations"],
                                  "seconds": per[s][k]["seconds"],
                                  "init_seconds": round(per[s][k]["init_seconds"], 5),
                                  "s_per_iteration": round(per[s][k]["s_per_iteration"], 5),
                                  "nmse_start": per[s][k]["nmse"][0],
                                  "nmse_after_1": per[s][k]["nmse"][1]}
                              for k in ("flat", "learned")} for s in seeds},
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(p, batch, seconds):
    """Oracle (NumPy, one core) forward+adjoint on mini-batches of the same
    workload until ~`seconds` of work; returns (pos-updates/s, sample text)."""
    from oracle import OracleSingleslicePtychography

    o = OracleSingleslicePtychography(energy=p._energy)
    o._region_of_interest_shape = p._region_of_interest_shape
    obj = p._object_initial.copy()
    probe = p._probe_initial.copy()
    rng = np.random.default_rng(0)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        idx = rng.permutation(p._num_diffraction_patterns)[:batch]
        pos = p._positions_px[idx]
        sp, op, ew, _ = o.forward(obj, probe, pos, p._amplitudes[idx])
        o.adjoint(obj, probe, op, sp, pos, ew, 0.5, 1, False)
        done += batch
    dt = time.perf_counter() - t0
    sample = (f"{done // batch} mini-batches of {batch} positions (forward+adjoint, probe update) "
              f"of the same workload, NumPy complex64 oracle, {dt:.1f} s, single process "
              f"(numpy FFT/bincount single-threaded); host {cpu_model()}, "
              f"os.cpu_count()={os.cpu_count()}")
    return done / dt, sample


def bench_many(args):
    """Batched uneven reconstructions against sequential reconstruct() calls."""
    import torch

    from py4dstem_amd.multi import MultiPtychography
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import many_workloads

    torch.cuda.set_device(0)
    t = time.perf_counter()
    ws = many_workloads(count=args.problems, seed=0, num_iter=30)
    log(f"{len(ws)} problems simulated in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    items = []
    for w in ws:
        p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
        p.preprocess(**w["preprocess"])
        items.append(p)
        del w["datacube"]
    torch.cuda.synchronize()
    log(f"preprocess in {time.perf_counter() - t:.1f} s")
    num_iter = ws[0]["reconstruct"]["num_iter"]
    npos = sum(p._num_diffraction_patterns for p in items)
    frames = sorted({int(p._region_of_interest_shape[0]) for p in items})
    m = MultiPtychography(items)
    t = time.perf_counter()
    batches, stage, timings = m.tune(use_cache=False)
    log(f"autotuner: mini-batches {sorted(set(batches))}, stage {stage} "
        f"({time.perf_counter() - t:.1f} s) {timings}")
    kws = [dict(seed_random=w["reconstruct"]["seed_random"], max_batch_size=b, fix_probe=False,
                step_size=0.5, normalization_min=1, reset=True) for w, b in zip(ws, batches)]

    def sequential():
        np.random.seed(7)
        for p, kw in zip(items, kws):
            p.reconstruct(num_iter=num_iter, progress_bar=False, **kw)

    def batched():
        np.random.seed(7)
        m.reconstruct(num_iter, kws, use_graph=True)

    for fn in (sequential, batched)[: 2 if args.warmup else 0]:
        fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    sequential()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t
    seq = [(p.object.copy(), p.probe.copy(), list(p.error_iterations)) for p in items]
    t = time.perf_counter()
    batched()
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t
    same = all(np.array_equal(o, p.object) and np.array_equal(pr, p.probe) and e == p.error_iterations
               for (o, pr, e), p in zip(seq, items))
    rate_bat = npos * num_iter / t_bat
    rate_seq = npos * num_iter / t_seq
    log(f"sequential {t_seq:.3f} s ({rate_seq:,.0f}/s), batched {t_bat:.3f} s ({rate_bat:,.0f}/s), "
        f"bitwise identical: {same}")
    out = {
        "metric": "aggregate position-updates/s (batched uneven reconstructions)",
        "value": round(rate_bat, 1),
        "unit": "position-updates/s",
        "n_gpus": 1,
        "steps": num_iter,
        "warmup": args.warmup,
        "ms_per_step": round(t_bat / num_iter * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded Gaussian-atom objects, GPU-simulated Poisson patterns)",
        "config": {
            "workload": f"{len(items)} independent GD ptychography reconstructions, square scans "
                        "32..96 per side, frames " + "/".join(map(str, frames)) +
                        f" px, {num_iter} iterations, batched in one graph replay per iteration",
            "problems": len(items), "total_positions": npos,
            "mini_batch": {"min": min(batches), "max": max(batches)}, "stage": stage,
            "autotune_s_per_candidate": timings,
        },
        "sequential": {"value": round(rate_seq, 1), "seconds": round(t_seq, 3)},
        "speedup_vs_sequential": round(t_seq / t_bat, 3),
        "bitwise_identical_to_sequential": bool(same),
    }
    print(json.dumps(out), flush=True)


def bench_kernels(args):
    """The production gradient-descent loop on a 1024-position-per-batch
    problem (128x128 raster at 4 px, 128 px frames, 720x720 object, uniform
    random amplitudes, mini-batches of 1024 in seeded order, probe update),
    `--kernel-reps` iterations of pty_gd_run, plus a plain copy of 256 MiB:
    the workload for hardware-counter passes.  Every kernel of a mini-batch
    (fused pass, scatter + probe partial sums, probe reduction and step,
    object step with the probe's transform) appears 16 times per iteration;
    the copy calibrates FETCH_SIZE / WRITE_SIZE against a known byte
    count."""
    import torch

    from py4dstem_amd import _lib
    from py4dstem_amd import kernels as K
    from py4dstem_amd.preprocess import center_positions, real_space_sampling, scan_positions_in_pixels
    from py4dstem_amd.probe import electron_wavelength_angstrom
    from py4dstem_amd.synthetic import scaled_probe

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    S, P, B, n = 128, 720, 1024, 128
    rng = np.random.default_rng(0)
    lam = electron_wavelength_angstrom(300e3)
    sampling = real_space_sampling(300e3, (lam * 1e3 / (S * 0.1),) * 2, (S, S))
    pos, _ = scan_positions_in_pixels(None, (n, n), (0.4, 0.4), 0.0, False, sampling, (S, S))
    pos = center_positions(pos.astype(np.float32), (P, P))
    probe = scaled_probe(300e3, S, sampling, 20.0, 2.0, 1.6e5, defocus=200.0).astype(np.complex64)
    obj = np.exp(1j * rng.uniform(0, 0.3, (P, P))).astype(np.complex64)
    o = torch.from_numpy(obj).to(dev)
    pr = torch.from_numpy(probe).to(dev)
    posd = torch.from_numpy(pos).to(dev)
    ampd = torch.rand((n * n, S, S), generator=torch.Generator(device=dev).manual_seed(0),
                      device=dev) * 20.0
    N = n * n
    reps = args.kernel_reps
    orders = torch.from_numpy(np.stack([rng.permutation(N) for _ in range(reps)]).astype(np.int32)).to(dev)
    errors = torch.zeros(reps, dtype=torch.float64, device=dev)
    problem = K.make_problem(o, pr, posd, ampd, None, 20.0 * 20.0 / 3 * S * S, None)
    params = _lib.PtyGdParams()
    params.num_iter, params.max_batch, params.fix_probe, params.fix_probe_com = reps, B, 0, 1
    params.step_size, params.normalization_min = 0.5, 1.0
    ws = torch.empty(K.gd_workspace_bytes(problem, params), dtype=torch.uint8, device=dev)
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    K.gd_run(problem, params, orders, errors, ws)
    torch.cuda.synchronize()
    print(json.dumps({
        "copy_bytes": 2 * src.numel() * 4,
        "minibatches": reps * (N // B),
        "k_fused_alg_bytes": B * (4 + 8) * S * S + 8 * S * S,
        "minibatch_alg_bytes": round(algorithmic_bytes_per_update(S, P, P, B) * B),
        "errors_finite": bool(torch.isfinite(errors).all().item()),
        "note": "k_fused algorithmic: amplitudes and object patch read, probe spectrum; "
                "mini-batch algorithmic: 20 S^2 per position plus the object and probe "
                "read and written once",
    }), flush=True)


def bench_learned_init(args):
    """Time to object NMSE <= 1e-3, learned start against flat start, on
    held-out objects of the initialiser's distribution at production shapes
    (128x128 scan, 128 px frames, 720x720 object)."""
    from py4dstem_amd import learned_init as li

    net = li.ConvNet.load()
    seeds = list(range(2001, 2001 + args.problems))
    per = {}
    for s in seeds:
        log(f"learned-init evaluation seed {s}")
        w = li.evaluation_workload(s)
        per[s] = li.time_to_nmse(w, net, max_iter=args.steps)
        for kind in ("flat", "learned"):
            r = per[s][kind]
            log(f"  {kind}: iterations {r['iterations']} seconds {r['seconds']} nmse0 {r['nmse'][0]:.3e}")
    def total(kind):
        v = [per[s][kind]["seconds"] for s in seeds]
        return None if any(x is None for x in v) else sum(v)
    t_flat, t_learned = total("flat"), total("learned")
    speedup = (t_flat / t_learned) if (t_flat and t_learned) else None
    # objects whose flat start is already at the target need no iteration
    # (flat time 0); the ratio over the others isolates the initialiser
    hard = [s for s in seeds if per[s]["flat"]["iterations"] and per[s]["learned"]["seconds"]]
    ratios = sorted(per[s]["flat"]["seconds"] / per[s]["learned"]["seconds"] for s in hard)
    out = {
        "metric": "time-to-NMSE<=1e-3, learned vs flat start",
        "value": round(speedup, 3) if speedup else None,
        "unit": "x (flat seconds / learned seconds)",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 network (fp32 accumulation), fp32 (complex64) reconstruction",
        "data": "synthetic (held-out Gaussian-atom objects, GPU-simulated Poisson patterns)",
        "config": {"workload": "learned initialiser vs flat start, 128x128 scan, 128 px frames, "
                               "720x720 object, mini-batch 1024, probe update",
                   "seeds": seeds},
        "per_seed": {str(s): {k: {kk: per[s][k][kk] for kk in ("iterations", "seconds", "init_seconds",
                                                              "s_per_iteration")}
                              | {"nmse_start": per[s][k]["nmse"][0], "nmse_1": per[s][k]["nmse"][1],
                                 "used_learned": (per[s][k]["init_info"] or {}).get("used_learned")}
                              for k in ("flat", "learned")} for s in seeds},
        "flat_seconds": t_flat, "learned_seconds": t_learned,
        "speedup_where_flat_start_misses_target": {
            "seeds": hard, "median": ratios[len(ratios) // 2] if ratios else None,
            "min": ratios[0] if ratios else None},
    }
    print(json.dumps(out, default=float), flush=True)


def bench_positions(args):
    """Production workload with position correction on (fix_positions=False):
    GD iterations with the per-batch position step, and the position kernel
    timed live on one mini-batch."""
    import torch

    from py4dstem_amd import kernels as K
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import production_workload

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    w = production_workload(seed=0, scan=(args.scan, args.scan))
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    S = int(p._region_of_interest_shape[0])
    Px, Py = p._object_shape
    N = p._num_diffraction_patterns
    B = w["reconstruct"]["max_batch_size"]
    recon = dict(w["reconstruct"], progress_bar=False, fix_positions=False,
                 positions_step_size=0.5)
    if args.warmup:
        p.reconstruct(**dict(recon, num_iter=args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.reconstruct(**dict(recon, num_iter=args.steps, reset=False))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pos_per_s = N * args.steps / dt
    log(f"{args.steps} iterations with position correction in {dt:.3f} s -> "
        f"{pos_per_s:,.0f} pos-updates/s")

    dev = torch.device("cuda", 0)
    order = torch.from_numpy(np.random.default_rng(1).permutation(N)[:B].astype(np.int32)).to(dev)
    phat = K.probe_spectrum(p._probe_dev)
    ps, chi, err, far = K.forward_adjoint(phat, p._object_dev, p._positions_dev,
                                          p._amplitudes_dev, order, with_far=True)
    positions = p._positions_dev.clone()
    initial = p._positions_initial_dev
    pp = K.position_params(positions, initial, 0.5)
    stream = torch.cuda.current_stream()
    K.position_correction(pp, ps, far, p._object_dev, p._amplitudes_dev, order)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.kernel_reps):
        K.position_correction(pp, ps, far, p._object_dev, p._amplitudes_dev, order)
    e1.record(stream)
    e1.synchronize()
    k_avg = e0.elapsed_time(e1) * 1e-3 / args.kernel_reps
    # shifted probe + far field + amplitudes + two shifted object patches
    per_pos = 36 * S * S
    achieved = per_pos * B / k_avg / 1e9
    log(f"k_position_correction avg {k_avg * 1e6:.1f} us per launch of {B} positions")

    cpu_rate = sample = None
    if args.cpu_seconds > 0:
        from oracle import OracleSingleslicePtychography
        o = OracleSingleslicePtychography(energy=p._energy)
        o._region_of_interest_shape = p._region_of_interest_shape
        o._amplitudes_shape = p._amplitudes_shape
        o._angular_sampling = p._angular_sampling
        obj = p._object_initial.copy()
        probe = p._probe_initial.copy()
        pos_h = p._positions_px.copy()
        rng = np.random.default_rng(0)
        done, b = 0, 256
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            idx = rng.permutation(N)[:b]
            pos = pos_h[idx]
            sp, op, ew, _ = o.forward(obj, probe, pos, p._amplitudes[idx])
            o.adjoint(obj, probe, op, sp, pos, ew, 0.5, 1, False)
            pos_h[idx] = o.position_correction(obj, pos, sp, sp * op, p._amplitudes[idx],
                                               p._positions_px_initial[idx], 0.5)
            done += b
        cdt = time.perf_counter() - t0
        cpu_rate = done / cdt
        sample = (f"{done // b} mini-batches of {b} positions (forward, adjoint with probe "
                  f"update, position correction), NumPy complex64 oracle, {cdt:.1f} s, one "
                  f"process; host {cpu_model()}")
        log(f"cpu baseline {cpu_rate:,.1f} pos-updates/s")

    out = {
        "metric": "position-updates/s",
        "value": round(pos_per_s, 1),
        "unit": "position-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded Gaussian-atom object, simulated Poisson patterns)",
        "config": {
            "workload": f"single-slice GD ptychography with position correction, "
                        f"{args.scan}x{args.scan} scan, {S}x{S} px patterns, {Px}x{Py} object, "
                        f"mini-batch {B}, probe update",
            "scan_positions": N, "frame_px": S, "object_px": [Px, Py], "mini_batch": B,
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_position_correction", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "kernel_avg_us": round(k_avg * 1e6, 2), "bytes_per_update": per_pos,
        },
        "cpu_baseline": {
            "value": round(cpu_rate, 2) if cpu_rate is not None else None,
            "unit": "position-updates/s", "cores": 1, "kind": "port", "sample": sample,
        },
    }
    print(json.dumps(out), flush=True)


def bench_projection(args):
    """Production workload with a projection-set method (DM_AP, alpha 0.5;
    all positions per update, exit waves resident in HBM); k_projection timed
    live on 1024 positions."""
    import torch

    from py4dstem_amd import kernels as K
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import production_workload

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    w = production_workload(seed=0, scan=(args.scan, args.scan))
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    S = int(p._region_of_interest_shape[0])
    Px, Py = p._object_shape
    N = p._num_diffraction_patterns
    recon = dict(w["reconstruct"], progress_bar=False, max_batch_size=None,
                 reconstruction_method="DM_AP", reconstruction_parameter=0.5)
    if args.correct_positions:
        # position correction after every projection-set update; the
        # corrected positions need the forward overlap's far field, one more
        # whole-batch fused pass per iteration
        recon.update(fix_positions=False, positions_step_size=0.1)
    if args.warmup:
        p.reconstruct(**dict(recon, num_iter=args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.reconstruct(**dict(recon, num_iter=args.steps, reset=False))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pos_per_s = N * args.steps / dt
    log(f"{args.steps} DM_AP iterations in {dt:.3f} s -> {pos_per_s:,.0f} pos-updates/s; "
        f"errors {p.error_iterations[-3:]}")

    dev = torch.device("cuda", 0)
    B = 1024
    order = torch.arange(B, dtype=torch.int32, device=dev)
    exits = p._exit_waves[:B].clone()
    jp = K.projection_params(exits, -0.5, 1, 1.5, exit_valid=True)
    phat = K.probe_spectrum(p._probe_dev)
    K.projection_batch(phat, p._object_dev, p._positions_dev, p._amplitudes_dev, order, jp)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.kernel_reps):
        K.projection_batch(phat, p._object_dev, p._positions_dev, p._amplitudes_dev, order, jp)
    e1.record(stream)
    e1.synchronize()
    k_avg = e0.elapsed_time(e1) * 1e-3 / args.kernel_reps
    # amplitudes 4, object patch 8, exit wave read 8 + partial write 8 + read 8 + write 8,
    # P_shift write 8 (x S^2 bytes)
    per_pos = 52 * S * S
    achieved = per_pos * B / k_avg / 1e9
    log(f"k_projection avg {k_avg * 1e6:.1f} us per launch of {B} positions")
    out = {
        "metric": "position-updates/s",
        "value": round(pos_per_s, 1),
        "unit": "position-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded Gaussian-atom object, simulated Poisson patterns)",
        "config": {
            "workload": f"single-slice ptychography, DM_AP(0.5) projection sets, "
                        f"{args.scan}x{args.scan} scan, {S}x{S} px patterns, {Px}x{Py} object, "
                        f"all positions per update, probe update"
                        + (", position correction" if args.correct_positions else ""),
            "scan_positions": N, "frame_px": S, "object_px": [Px, Py],
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_projection", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "kernel_avg_us": round(k_avg * 1e6, 2), "bytes_per_update": per_pos,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_mixed(args):
    """Production workload as a mixed-state reconstruction with
    ``--probe-modes`` incoherent probe modes (gradient descent, mini-batch
    1024, probe update and mode orthogonalisation on)."""
    import torch

    from py4dstem_amd import MixedstatePtychography
    from py4dstem_amd.synthetic import production_workload

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    w = production_workload(seed=0, scan=(args.scan, args.scan))
    np.random.seed(0)
    p = MixedstatePtychography(datacube=w["datacube"], num_probes=args.probe_modes, **w["init"])
    p.preprocess(**w["preprocess"])
    S = int(p._region_of_interest_shape[0])
    Px, Py = p._object_shape
    N = p._num_diffraction_patterns
    B = w["reconstruct"]["max_batch_size"]
    recon = dict(w["reconstruct"], progress_bar=False)
    if args.warmup:
        p.reconstruct(**dict(recon, num_iter=args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.reconstruct(**dict(recon, num_iter=args.steps, reset=False))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pos_per_s = N * args.steps / dt
    log(f"{args.steps} mixed-state iterations (K={args.probe_modes}) in {dt:.3f} s -> "
        f"{pos_per_s:,.0f} pos-updates/s; errors {p.error_iterations[-3:]}")
    out = {
        "metric": "position-updates/s", "value": round(pos_per_s, 1), "unit": "position-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded Gaussian-atom object, simulated Poisson patterns)",
        "config": {
            "workload": f"mixed-state GD ptychography, {args.probe_modes} probe modes, "
                        f"{args.scan}x{args.scan} scan, {S}x{S} px patterns, {Px}x{Py} object, "
                        f"mini-batch {B}, probe update, mode orthogonalisation",
            "scan_positions": N, "frame_px": S, "object_px": [Px, Py], "mini_batch": B,
            "probe_modes": args.probe_modes,
        },
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--scan", type=int, default=128)
    ap.add_argument("--workload", choices=("production", "many", "kernels", "learned-init",
                                           "positions", "projection", "mixed"),
                    default="production")
    ap.add_argument("--probe-modes", type=int, default=4)
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--correct-positions", action="store_true",
                    help="--workload projection: with position correction (fix_positions=False)")
    ap.add_argument("--learned-seeds", type=int, default=4,
                    help="held-out seeds for the time-to-NMSE sub-object of the default line "
                         "(0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); the multi-GPU "
              "path is not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "many":
        return bench_many(args)
    if args.workload == "kernels":
        return bench_kernels(args)
    if args.workload == "learned-init":
        return bench_learned_init(args)
    if args.workload == "positions":
        return bench_positions(args)
    if args.workload == "projection":
        return bench_projection(args)
    if args.workload == "mixed":
        return bench_mixed(args)

    import torch

    from py4dstem_amd import kernels as K
    from py4dstem_amd.ptychography import SingleslicePtychography
    from py4dstem_amd.synthetic import production_workload

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    t = time.perf_counter()
    w = production_workload(seed=0, scan=(args.scan, args.scan))
    log(f"workload generated in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    p = SingleslicePtychography(datacube=w["datacube"], **w["init"])
    p.preprocess(**w["preprocess"])
    torch.cuda.synchronize()
    log(f"preprocess in {time.perf_counter() - t:.1f} s")
    S = int(p._region_of_interest_shape[0])
    Px, Py = p._object_shape
    N = p._num_diffraction_patterns
    B = w["reconstruct"]["max_batch_size"]
    recon = dict(w["reconstruct"], progress_bar=False)

    if args.warmup:
        p.reconstruct(**dict(recon, num_iter=args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.reconstruct(**dict(recon, num_iter=args.steps, reset=False))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pos_per_s = N * args.steps / dt
    log(f"{args.steps} iterations in {dt:.3f} s -> {pos_per_s:,.0f} pos-updates/s; "
        f"errors {p.error_iterations[-3:]}")

    # live timing of the dominant kernel (fused forward/adjoint, one launch = B
    # positions) with HIP events on the stream it is launched on
    dev = torch.device("cuda", 0)
    obj = p._object_dev.clone()
    phat = K.probe_spectrum(p._probe_dev)
    order = torch.from_numpy(np.random.default_rng(1).permutation(N)[:B].astype(np.int32)).to(dev)
    stream = torch.cuda.current_stream()
    ps = torch.empty((B, S, S), dtype=torch.complex64, device=dev)
    K.forward_adjoint(phat, obj, p._positions_dev, p._amplitudes_dev, order)
    torch.cuda.synchronize()
    # back-to-back launches between two events: average launch duration
    bufs = [K.forward_adjoint(phat, obj, p._positions_dev, p._amplitudes_dev, order)
            for _ in range(2)]
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.kernel_reps):
        K.forward_adjoint(phat, obj, p._positions_dev, p._amplitudes_dev, order)
    e1.record(stream)
    e1.synchronize()
    del ps, bufs
    k_avg = e0.elapsed_time(e1) * 1e-3 / args.kernel_reps
    per_update = algorithmic_bytes_per_update(S, Px, Py, B)
    # the fused pass's own algorithmic bytes: amplitudes and object patch
    # read per position, the probe spectrum once (its P_shift / chi staging
    # is not algorithmic: the scatter and probe sums read it back)
    fused_bytes = B * (4 + 8) * S * S + 8 * S * S
    achieved = fused_bytes / k_avg / 1e9
    iter_achieved = pos_per_s * per_update / 1e9
    log(f"k_fused avg {k_avg * 1e6:.1f} us per launch of {B} positions: "
        f"{achieved:,.0f} GB/s; iteration level {iter_achieved:,.0f} GB/s")
    production = (S, B, Px, Py) == (128, 1024, 720, 720)

    if args.cpu_seconds > 0:
        cpu_rate, sample = cpu_baseline(p, B, args.cpu_seconds)
        log(f"cpu baseline {cpu_rate:,.1f} pos-updates/s")
    else:
        cpu_rate, sample = None, "skipped (--cpu-seconds 0)"

    out = {
        "metric": "position-updates/s",
        "value": round(pos_per_s, 1),
        "unit": "position-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded Gaussian-atom object, simulated Poisson patterns)",
        "config": {
            "workload": f"single-slice GD ptychography, {args.scan}x{args.scan} scan, "
                        f"{S}x{S} px patterns, {Px}x{Py} object, mini-batch {B}, probe update",
            "scan_positions": N, "frame_px": S, "object_px": [Px, Py], "mini_batch": B,
            "iterations_timed": args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_fused",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("k_fused") if production else None,
            "kernel_avg_us": round(k_avg * 1e6, 2),
            "bytes_per_launch": fused_bytes,
            "iteration": {
                "achieved": round(iter_achieved, 1),
                "frac": round(iter_achieved / HBM_PEAK_GBS, 4),
                "bytes_per_update": per_update,
                "traffic_per_minibatch": measured_traffic("minibatch") if production else None,
                "algorithmic_per_minibatch": round(per_update * B),
            },
        },
        "cpu_baseline": {
            "value": round(cpu_rate, 2) if cpu_rate is not None else None,
            "unit": "position-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": sample,
        },
    }
    if args.learned_seeds > 0:
        try:
            out["learned_init"] = learned_init_summary(list(range(2001, 2001 + args.learned_seeds)),
                                                       max_iter=4)
        except Exception as exc:  # reported, never hides the main line
            out["learned_init"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(out, default=float), flush=True)


if __name__ == "__main__":
    main()
