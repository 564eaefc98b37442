"""Staged and pipelined gradient descent against the whole-batch launch.

Splitting a mini-batch into several fused launches ("stages") carries each
pixel's float64 scatter sums from stage to stage in batch order, and
pipelining runs the scatter and probe sums of stage k on a second stream
beside the fused pass of stage k + 1 with P_shift and chi double-buffered.
Neither changes what is summed or in which order, so object, probe and every
per-iteration error must be bitwise equal to the whole-batch run: probe
update (scatter and probe sums in one launch), fixed probe (scatter alone),
a potential object, uneven stage splits and 128 px frames.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    from py4dstem_amd import _lib
    _lib.load()
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def case64():
    from py4dstem_amd.synthetic import small_workload
    return small_workload(seed=41, scan=(24, 24), S=64, object_px=224)


@pytest.fixture(scope="module")
def case128():
    from py4dstem_amd.synthetic import small_workload
    return small_workload(seed=42, scan=(16, 16), S=128, object_px=256)


def _run(w, stage, pipeline, object_type="complex", **recon):
    from py4dstem_amd.ptychography import SingleslicePtychography
    init = dict(w["init"], initial_object_guess=w["init"]["initial_object_guess"].copy(),
                object_type=object_type)
    if object_type == "potential":
        init["initial_object_guess"] = np.zeros(init["initial_object_guess"].shape, np.float32)
    p = SingleslicePtychography(datacube=w["datacube"], **init)
    p.preprocess(**w["preprocess"])
    if stage is not None:
        p.set_stage(stage, pipeline=pipeline)
    kw = dict(w["reconstruct"], progress_bar=False)
    kw.update(recon)
    p.reconstruct(**kw)
    return p.object.copy(), p.probe.copy(), np.array(p.error_iterations)


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("variant", ["probe-update", "fixed-probe", "potential"])
def test_pipelined_stages_equal_whole_batch_64(dev, case64, variant):
    recon = dict(num_iter=3, max_batch_size=576, seed_random=7,
                 fix_probe=variant == "fixed-probe")
    object_type = "potential" if variant == "potential" else "complex"
    say(f"S=64, 576 positions, {variant}: whole batch, staged 128, pipelined 128 and 192")
    whole = _run(case64, 0, False, object_type, **recon)
    staged = _run(case64, 128, False, object_type, **recon)
    piped = _run(case64, 128, True, object_type, **recon)
    piped_uneven = _run(case64, 192, True, object_type, **recon)  # 192 + 192 + 192
    assert np.isfinite(whole[2]).all() and whole[2][-1] < whole[2][0]
    _same(whole, staged)
    _same(whole, piped)
    _same(whole, piped_uneven)


def test_pipelined_stages_equal_whole_batch_128_minibatches(dev, case128):
    # mini-batches of 100 (subdivided 86 + 85 + 85) in stages of 64: every
    # mini-batch has two stages, the second shorter
    recon = dict(num_iter=2, max_batch_size=100, seed_random=9, fix_probe=False)
    say("S=128, 256 positions, mini-batches of ~86: whole batch vs pipelined stages of 64")
    whole = _run(case128, 0, False, **recon)
    piped = _run(case128, 64, True, **recon)
    _same(whole, piped)


def test_pipeline_without_stages_is_the_plain_run(dev, case64):
    # pipeline with stage >= mini-batch: one launch per mini-batch, one stream
    recon = dict(num_iter=2, max_batch_size=200, seed_random=3, fix_probe=False)
    _same(_run(case64, 0, False, **recon), _run(case64, 0, True, **recon))
