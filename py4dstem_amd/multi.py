"""Many independent reconstructions advanced together on one GPU.

``MultiPtychography`` takes preprocessed ``SingleslicePtychography`` objects
(any mix of 64 and 128 pixel frames, any scan and object sizes) and runs
their GD iterations with ``pty_multi_run``: one launch per stage for all
problems and, by default, one captured hipGraph replayed per iteration.
Each problem's batch order is generated exactly as its own ``reconstruct()``
would, and its result is bitwise identical to that sequential solve.

The reference reconstructs one dataset at a time
(py4DSTEM/process/phase/singleslice_ptychography.py:606-996); this runner has
no reference counterpart.
"""

from __future__ import annotations

import ctypes
import json

import numpy as np
import torch

from . import _lib
from . import autotune
from ._lib import PtyMultiProblem, stream_handle


class MultiPtychography:
    def __init__(self, reconstructions):
        self.items = list(reconstructions)
        if not self.items:
            raise ValueError("no reconstructions given")
        for p in self.items:
            if not getattr(p, "_preprocessed", False):
                raise ValueError("every reconstruction must be preprocessed")

    def _runs(self, num_iter, kwargs):
        runs = []
        for p, kw in zip(self.items, kwargs):
            if int(p._region_of_interest_shape[0]) not in (64, 128):
                raise NotImplementedError("the batched runner takes 64 or 128 px frames; "
                                          f"{int(p._region_of_interest_shape[0])} px frames run "
                                          "through SingleslicePtychography.reconstruct()")
            if getattr(p, "_num_probes", 1) > 1:
                raise NotImplementedError("mixed-state reconstructions run through their own "
                                          "reconstruct(), not the batched runner")
            kw = dict(kw, num_iter=num_iter, progress_bar=False, store_iterations=False)
            run = p._prepare_run(_allocate=False, **kw)
            if run["projection_params"] is not None:
                raise NotImplementedError("projection-set methods run through "
                                          "SingleslicePtychography.reconstruct(), not the batched "
                                          "runner")
            if run["fourier_mask"] is not None:
                raise NotImplementedError("detector Fourier masks run through "
                                          "SingleslicePtychography.reconstruct(), not the batched "
                                          "runner")
            if run["position_params"] is not None:
                raise NotImplementedError("position correction (fix_positions=False) runs through "
                                          "SingleslicePtychography.reconstruct(), not the batched "
                                          "runner")
            if (run["params"].fix_probe_aperture or run["params"].constrain_probe_amplitude
                    or run["params"].gaussian_weights):
                raise NotImplementedError("probe aperture/amplitude constraints and object "
                                          "smoothing run through "
                                          "SingleslicePtychography.reconstruct(), not the batched "
                                          "runner")
            runs.append(run)
        return runs

    def reconstruct(self, num_iter, reconstruct_kwargs=None, use_graph=True):
        """num_iter GD iterations of every problem.  reconstruct_kwargs: one
        dict (shared) or a list of dicts with reconstruct() keywords."""
        if reconstruct_kwargs is None:
            reconstruct_kwargs = {}
        if isinstance(reconstruct_kwargs, dict):
            reconstruct_kwargs = [reconstruct_kwargs] * len(self.items)
        runs = self._runs(num_iter, reconstruct_kwargs)
        self.run_prepared(runs, num_iter, use_graph)
        for p, run in zip(self.items, runs):
            p._finish_run(run)
        return self

    @staticmethod
    def run_prepared(runs, num_iter, use_graph=True):
        P = len(runs)
        arr = (PtyMultiProblem * P)()
        for i, run in enumerate(runs):
            arr[i].problem = run["problem"]
            arr[i].params = run["params"]
            arr[i].orders = ctypes.c_void_p(run["orders"].data_ptr())
            arr[i].errors = ctypes.c_void_p(run["errors"].data_ptr())
        lib = _lib.load()
        nbytes = lib.pty_multi_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), P)
        if nbytes == 0:
            raise RuntimeError(f"pty_multi_workspace_bytes failed: {_lib.last_error()}")
        ws = torch.empty(nbytes, dtype=torch.uint8, device=runs[0]["orders"].device)
        _lib.call("pty_multi_run", ctypes.cast(arr, ctypes.c_void_p), P, int(num_iter),
                  int(bool(use_graph)), ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle())

    # -- autotuning ------------------------------------------------------------
    def signature(self):
        return json.dumps(sorted((int(p._region_of_interest_shape[0]), int(p._num_diffraction_patterns),
                                  tuple(int(v) for v in p._object_shape)) for p in self.items))

    def tune(self, batch_candidates=(256, 512, 1024), split_candidates=(2, 4, 8),
             stage_candidates=(0, 256), iters=2, reconstruct_kwargs=None, use_cache=True):
        """Pick every problem's mini-batch size (when the caller leaves it
        free) and positions per fused launch by timing `iters` iterations of
        each candidate on copies of the problems' current state, all problems
        together as reconstruct() runs them.  Mini-batch candidates are either
        one size for every problem (batch_candidates) or a split of every
        problem into the same number of mini-batches (split_candidates: a
        problem of N positions gets ceil(N / k)), which keeps the problems'
        launch sequences the same length.  Returns (per-problem batch sizes,
        stage, timings); the choice is cached in a JSON file keyed by the
        problems' signature."""
        key = "multi2:" + self.signature()
        if use_cache:
            hit = autotune.lookup(key)
            if hit:
                self._apply(hit["batch"], hit["stage"])
                return list(hit["batch"]), hit["stage"], hit.get("timings", {})
        kw = dict(reconstruct_kwargs or {})
        saved = [(p._object_dev.clone() if hasattr(p, "_object_dev") else None,
                  p._probe_dev.clone() if hasattr(p, "_probe_dev") else None,
                  list(getattr(p, "error_iterations", [])), hasattr(p, "error")) for p in self.items]
        sizes = [int(p._num_diffraction_patterns) for p in self.items]
        batches = {f"b{b}": [min(b, n) for n in sizes] for b in batch_candidates}
        batches.update({f"k{k}": [-(-n // k) for n in sizes] for k in split_candidates})

        def trial(cand):
            name, stage = cand
            self._apply(batches[name], stage)
            runs = self._runs(iters, [dict(kw, max_batch_size=b, reset=True) for b in self.batch])
            self.run_prepared(runs, iters)
            torch.cuda.synchronize()

        cands = [(name, s) for name in batches for s in stage_candidates
                 if s == 0 or s < max(batches[name])]
        best, timings = autotune.best_of(cands, trial, repeats=1)
        for p, (o, pr, errs, had) in zip(self.items, saved):
            for attr, val in (("_object_dev", o), ("_probe_dev", pr)):
                if val is not None:
                    setattr(p, attr, val)
                elif hasattr(p, attr):
                    delattr(p, attr)
            p.error_iterations = errs
            if not had and hasattr(p, "error"):
                del p.error
        self._apply(batches[best[0]], best[1])
        tim = {f"{n}/{s}": t for (n, s), t in timings.items()}
        autotune.remember(key, {"batch": self.batch, "stage": best[1], "choice": best[0],
                                "timings": tim})
        return list(self.batch), best[1], tim

    def _apply(self, batch, stage):
        """batch: one size for every problem or a list, one per problem."""
        self.batch = list(batch) if isinstance(batch, (list, tuple)) else [batch] * len(self.items)
        for p in self.items:
            p.set_stage(stage)
