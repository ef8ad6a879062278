"""Launch-shape autotuner for the batched MFMA local-inverse apply.

Two kernels compute the product, in shapes (grid, waves, rpw):
  grid = -1   dds_mfma_apply_sub: one workgroup of `waves` (4, 8, 16) wavefronts per
              subdomain, each wavefront taking `rpw` (2, 4, 8) row blocks of 16 rows at a
              time (the tile height of the product: rpw A operands share each load of the
              packed right-hand sides);
  grid >= 0   dds_mfma_apply: a persistent grid of `grid` workgroups (0: one wavefront per
              item) of `waves` (1, 2, 4, 8) wavefronts walking the (subdomain, 4 row blocks)
              work list.
The best shape depends on the subdomain size mix and on how much work there is per compute
unit, so it is measured rather than guessed: every candidate is timed with HIP events on
the current stream over a few repetitions on the preconditioner's own matrices and a random
block of right-hand sides, and the fastest one is written into the preconditioner. Every
candidate computes the same result bit for bit (each row block's accumulation is the same
instruction sequence whatever the shape), so tuning cannot change the solve. Results are
cached per (device, storage dtype, subdomain-size multiset).
"""

from __future__ import annotations

import hashlib

import numpy as np
import torch

from . import _lib

__all__ = ["autotune_mfma", "mfma_candidates"]

_CACHE: dict = {}


def mfma_candidates(n_items: int, n_cu: int):
    """(grid, waves, rpw) shapes: grid -1 one workgroup per subdomain; grid 0 one wavefront
    per work item; grid > 0 a persistent grid (rpw 4)."""
    out = [(-1, waves, rpw) for waves in (4, 8) for rpw in (2, 4, 8)]
    for waves in (1, 2, 4, 8):
        out.append((0, waves, 4))
        for per_cu in (1, 2, 4, 8):
            g = n_cu * per_cu
            if g * waves < n_items:
                out.append((g, waves, 4))
    return out


def _key(pc) -> tuple:
    sizes = np.diff(pc.sub_ptr.cpu().numpy())
    h = hashlib.sha1(np.sort(sizes).astype(np.int64).tobytes()).hexdigest()
    return (torch.cuda.get_device_name(pc.device), str(pc.local.mat_dtype), h)


def autotune_mfma(pc, reps: int = 5, verbose: bool = False) -> dict:
    """Pick (grid, waves, rpw) for `pc` (a Schwarz preconditioner set up with
    Inv(..., layout='mfma')); returns {'grid', 'waves', 'rpw', 'ms', 'timings'}."""
    mf = getattr(pc, "_mfma", None)
    if mf is None:
        raise ValueError("autotune_mfma needs a preconditioner set up with layout='mfma'")
    key = _key(pc)
    if key in _CACHE:
        best = _CACHE[key]
        mf["grid"], mf["waves"], mf["rpw"] = best["grid"], best["waves"], best["rpw"]
        return best
    dev = pc.device
    with torch.cuda.device(dev):
        n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
        gen = torch.Generator(device=dev).manual_seed(0)
        X = torch.rand((pc.n, _lib.NRHS), dtype=torch.float64, device=dev, generator=gen) - 0.5
        Y = mf["Ym"]
        pc._enqueue_local_multi(X, Y, None)  # pack once; the apply reads the pack
        timings = {}
        for shape in mfma_candidates(pc.local.n_items, n_cu):
            mf["grid"], mf["waves"], mf["rpw"] = shape
            pc._enqueue_mfma_apply(Y, None)  # warm
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(reps):
                pc._enqueue_mfma_apply(Y, None)
            t1.record()
            t1.synchronize()
            timings[shape] = t0.elapsed_time(t1) / reps
            if verbose:
                print(f"mfma grid={shape[0]} waves={shape[1]} rpw={shape[2]}: "
                      f"{timings[shape] * 1e3:.1f} us")
        (grid, waves, rpw), ms = min(timings.items(), key=lambda kv: kv[1])
    best = {"grid": int(grid), "waves": int(waves), "rpw": int(rpw), "ms": float(ms),
            "timings": timings}
    _CACHE[key] = best
    mf["grid"], mf["waves"], mf["rpw"] = best["grid"], best["waves"], best["rpw"]
    return best
