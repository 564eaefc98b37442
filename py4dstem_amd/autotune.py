"""Execution-parameter autotuner with a JSON cache.

Tunes what does not change results (positions per fused launch, "stage") and,
for the batched multi-problem solver, also the mini-batch size when the
caller leaves it free.  Candidates are timed on the GPU with the real
problem; the winner is cached in a JSON file keyed by the problem signature.
"""

from __future__ import annotations

import json
import os
import time

CACHE_PATH = os.environ.get(
    "PTY_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "py4dstem_amd_tune.json"))


def default_stage(S, batch):
    """Stage size used when nothing was tuned: the whole mini-batch in one
    fused launch (measured fastest at 64 and 128 px frames), unless staging
    P_shift and chi (16 bytes per pixel per position) would exceed 2 GiB."""
    per_pos = 16 * S * S
    if per_pos * batch <= (2 << 30):
        return 0
    return max(64, (2 << 30) // per_pos // 64 * 64)


def default_execution(S, batch):
    """(stage, pipeline) used when nothing was tuned: default_stage's
    positions per fused launch, not pipelined."""
    return default_stage(S, batch), False


def _load():
    try:
        with open(CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        tmp = CACHE_PATH + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE_PATH)
    except OSError:
        pass


def lookup(key):
    return _load().get(key)


def remember(key, value):
    cache = _load()
    cache[key] = value
    _store(cache)


def best_of(candidates, run, repeats=2):
    """Time run(candidate) (which must synchronise) and return
    (best candidate, {candidate: seconds})."""
    timings = {}
    for c in candidates:
        run(c)  # warm-up / first-touch
        t0 = time.perf_counter()
        for _ in range(repeats):
            run(c)
        timings[c] = (time.perf_counter() - t0) / repeats
    best = min(timings, key=timings.get)
    return best, timings
