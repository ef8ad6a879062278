"""Launch-shape autotuner for the batched MFMA local-inverse apply.

`dds_mfma_apply` is a persistent kernel: `grid` workgroups of `waves` wavefronts walk the
(subdomain, row-group) work list. The best shape depends on the subdomain size mix and on
how many work items there are per compute unit, so it is measured rather than guessed:
every candidate is timed with HIP events on the current stream over a few repetitions on
the preconditioner's own matrices and a random block of right-hand sides, and the fastest
one is written into the preconditioner. Every candidate computes the same result bit for
bit (one wavefront owns one work item whatever the shape), so tuning cannot change the
solve. Results are cached per (device, storage dtype, subdomain-size multiset).
"""

from __future__ import annotations

import hashlib

import numpy as np
import torch

from . import _lib

__all__ = ["autotune_mfma", "mfma_candidates"]

_CACHE: dict = {}


def mfma_candidates(n_items: int, n_cu: int):
    """(grid, waves) pairs: grid 0 means one wavefront per work item (no persistence)."""
    out = []
    for waves in (1, 2, 4, 8):
        out.append((0, waves))
        for per_cu in (1, 2, 4, 8):
            g = n_cu * per_cu
            if g * waves < n_items:
                out.append((g, waves))
    return out


def _key(pc) -> tuple:
    sizes = np.diff(pc.sub_ptr.cpu().numpy())
    h = hashlib.sha1(np.sort(sizes).astype(np.int64).tobytes()).hexdigest()
    return (torch.cuda.get_device_name(pc.device), str(pc.local.mat_dtype), h)


def autotune_mfma(pc, reps: int = 5, verbose: bool = False) -> dict:
    """Pick (grid, waves) for `pc` (a Schwarz preconditioner set up with
    Inv(..., layout='mfma')); returns {'grid', 'waves', 'ms', 'timings'}."""
    mf = getattr(pc, "_mfma", None)
    if mf is None:
        raise ValueError("autotune_mfma needs a preconditioner set up with layout='mfma'")
    key = _key(pc)
    if key in _CACHE:
        best = _CACHE[key]
        mf["grid"], mf["waves"] = best["grid"], best["waves"]
        return best
    dev = pc.device
    with torch.cuda.device(dev):
        n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
        gen = torch.Generator(device=dev).manual_seed(0)
        X = torch.rand((pc.n, _lib.NRHS), dtype=torch.float64, device=dev, generator=gen) - 0.5
        Y = mf["Ym"]
        pc._enqueue_local_multi(X, Y, None)  # pack once; the apply reads the pack
        L = pc.local
        dt = _lib.dtype_code(L.mat_dtype)
        s = _lib.stream()

        def run(grid, waves):
            _lib.call("dds_mfma_apply", dt, L.n_items, L.items.data_ptr(), pc.sub_ptr.data_ptr(),
                      L.mat_off.data_ptr(), mf["b_off"].data_ptr(), _lib.ptr(L.scale),
                      mf["sigma"].data_ptr(), L.mats.data_ptr(), mf["bp"].data_ptr(), Y.data_ptr(),
                      int(grid), int(waves), None, s)

        timings = {}
        for grid, waves in mfma_candidates(L.n_items, n_cu):
            run(grid, waves)  # warm
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(reps):
                run(grid, waves)
            t1.record()
            t1.synchronize()
            timings[(grid, waves)] = t0.elapsed_time(t1) / reps
            if verbose:
                print(f"mfma grid={grid} waves={waves}: {timings[(grid, waves)] * 1e3:.1f} us")
        (grid, waves), ms = min(timings.items(), key=lambda kv: kv[1])
    best = {"grid": int(grid), "waves": int(waves), "ms": float(ms), "timings": timings}
    _CACHE[key] = best
    mf["grid"], mf["waves"] = best["grid"], best["waves"]
    return best
