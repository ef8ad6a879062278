"""Time-to-solution benchmark of Schwarz-preconditioned CG on one MI355X.

Workload (default): 3D Poisson -div(k grad u) = f, Q1 hexahedra on a 128^3 grid
(n = 127^3 = 2,048,383 unknowns), k = exp(g) per element with g ~ N(0, 1), f ~ U(-1, 1) per
unknown (seed 0), Hybrid Schwarz with 4096 boxes of 8^3 elements and overlap 1, dense FP64
Cholesky coarse solve (n_c = 4096), FP64 PCG to rtol 1e-8, x0 = 0.

One step = one complete solve (time to solution), inputs already resident in HBM, setup
excluded (the reference's protocol). Prints one JSON line. `value` is the time to solution
of the mixed-precision form: local inverses stored in FP16 (scaled per subdomain) and
applied with FP32 accumulation, PCG recurrences, dots, residual check and coarse solve in
FP64. Beside it: the plain FP64 solve of the same system and the ratio, the FP32 and BF16
storage forms, the batched form (16 right-hand sides, FP16 MFMA local inverses), the SpMV;
the 64^3 system (FP64 and mixed, the BSR block-size sweep b = 2, 4, 8, sparse local
factors, batched forms), the uneven-subdomain batching workload (jittered 2D P1 512^2,
256 k-means subdomains, overlap 2, 16 right-hand sides), and the reference's own DG
precision-experiment runs and kernel-benchmark shapes beside its published A100 times.

    python bench.py [--steps K] [--warmup W] [--N 128] [--no-cpu] [--no-extra]
                    [--no-reference]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dd-solvers_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


_TNAME = {}


def _tnames():
    import torch
    _TNAME.update({torch.float64: "double", torch.float32: "float", torch.float16: "_Float16",
                   torch.bfloat16: "bf16_t"})


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(N: int, seed: int):
    from dd_solvers import problems as P

    t = time.perf_counter()
    pb = P.make_problem("q1_3d", N, box=8, overlap=1, seed=seed, lognormal=True)
    log(f"assembled Q1 {N}^3: n={pb.A.shape[0]} nnz={pb.A.nnz} subdomains="
        f"{pb.decomposition.n_subdomains} ({time.perf_counter() - t:.1f}s)")
    return pb


def make_solver(pb, storage, dev, bsr=None, scaled=False, use_graph=True, accumulate=None):
    import torch
    import dd_solvers as dds

    A = torch.sparse_csr_tensor(torch.as_tensor(pb.A.indptr), torch.as_tensor(pb.A.indices),
                                torch.as_tensor(pb.A.data), pb.A.shape).to(dev)
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(storage, scaled=scaled, accumulate=accumulate),
                                 dds.DenseCholesky()),
               rtol=1e-8, maxiter=2000, bsr_matmul=bsr is not None, check_every=8,
               use_graph=use_graph)
    kw = {"block_size": bsr} if bsr else {}
    t = time.perf_counter()
    s.setup(A, decomposition=pb.decomposition, **kw)
    torch.cuda.synchronize()
    log(f"setup storage={storage} accumulate={accumulate} bsr={bsr}: {time.perf_counter() - t:.1f}s")
    return s


def make_block_solver(pb, storage, dev, kind="hybrid", use_graph=True, autotune=True, layout=None):
    """CG over blocks of 16 right-hand sides on the matrix cores: the symmetric tiles
    (layout 'sym_mfma', half the bytes) where every subdomain has at most 1024 unknowns,
    else the full inverses as an MFMA batched GEMM (layout 'mfma', autotuned)."""
    import numpy as np
    import torch
    import dd_solvers as dds

    A = torch.sparse_csr_tensor(torch.as_tensor(pb.A.indptr), torch.as_tensor(pb.A.indices),
                                torch.as_tensor(pb.A.data), pb.A.shape).to(dev)
    cls = dds.HybridSchwarz if kind == "hybrid" else dds.AdditiveSchwarz
    if layout is None:
        layout = "sym_mfma" if int(np.diff(pb.decomposition.sub_ptr).max()) <= 1024 else "mfma"
    inv = dds.Inv(storage, scaled=True, layout="sym_mfma", accumulate=torch.float32) \
        if layout == "sym_mfma" else dds.Inv(storage, scaled=True, layout="mfma", autotune=autotune)
    s = dds.CG(cls(torch.float64, inv, dds.DenseCholesky()),
               rtol=1e-8, maxiter=2000, bsr_matmul=False, check_every=8, use_graph=use_graph)
    t = time.perf_counter()
    s.setup(A, decomposition=pb.decomposition)
    torch.cuda.synchronize()
    tn = getattr(s.preconditioner, "tuning", None)
    log(f"setup block storage={storage} layout={layout}: {time.perf_counter() - t:.1f}s"
        + (f", tuned grid={tn['grid']} waves={tn['waves']} rpw={tn['rpw']} ({tn['ms'] * 1e3:.1f} us)"
           if tn else ""))
    return s


def mfma_kernel_stats(M):
    """Event-timed average of one launch of the block apply alone (dds_mfma_apply, or
    dds_sym_mfma_apply16 on the symmetric tiles), with its algorithmic bytes (stored
    inverses + packed right-hand sides and scales + FP64 result) and MFMA flops (2 k^2 16 NT
    per subdomain: the symmetric tiles are used twice), and the pack beside it."""
    import numpy as np
    import torch
    from dd_solvers import _lib

    L = M.local
    sym = M._mfma is None
    mf = M._msym if sym else M._mfma
    sizes = M._host["sizes"].astype(np.int64)
    X = torch.rand((M.n, _lib.NRHS), dtype=torch.float64, device=M.device) - 0.5
    M._enqueue_local_multi(X, mf["Ym"], None)
    dt = _lib.dtype_code(L.mat_dtype)
    s = _lib.stream()
    nt = 2 if L.mat_dtype == torch.float16 else 3
    esz = L.mats.element_size()
    out = {}
    if sym:
        xp = M._xpack

        def launch():
            _lib.call("dds_sym_mfma_apply16", dt, M.n_solvers, L.max_k, mf["order"].data_ptr(),
                      M.sub_ptr.data_ptr(), xp["blk_ptr"].data_ptr(), L.mat_off.data_ptr(),
                      _lib.ptr(L.scale), mf["sigma"].data_ptr(), L.mats.data_ptr(),
                      mf["bp"].data_ptr(), mf["Ym"].data_ptr(), None, s)

        def pack():
            _lib.call("dds_sym_mfma_pack16", dt, xp["n_blk"], xp["blk_owner"].data_ptr(),
                      xp["blk_ptr"].data_ptr(), M.sub_ptr.data_ptr(), _lib.ptr(M.sub_idx),
                      X.data_ptr(), mf["sigma"].data_ptr(), mf["bp"].data_ptr(), None, s)

        kname = "sym16_apply_kernel"
        tp = kernel_time_ms(pack)
        pack_alg = int(sizes.sum() * 16 * 8 + mf["bp"].numel() * esz + mf["sigma"].numel() * 4
                       + (sizes.sum() * 4 if M.sub_idx is not None else 0))
        out["pack"] = {"kernel": f"sym16_pack_kernel<{'_Float16' if nt == 2 else 'bf16_t'}>",
                       "avg_launch_us": round(tp * 1e3, 2), "alg_bytes_per_launch": pack_alg,
                       "GBs": round(pack_alg / (tp * 1e-3) / 1e9, 1)}
    else:
        def launch():
            M._enqueue_mfma_apply(mf["Ym"], None)

        kname = "mfma_sub_kernel" if int(mf["grid"]) < 0 else "mfma_apply_kernel"
        out.update(grid=int(mf["grid"]), waves=int(mf["waves"]), rpw=int(mf["rpw"]))
    ms = kernel_time_ms(launch)
    alg = int(L.bytes_algorithmic + mf["bp"].numel() * esz + mf["sigma"].numel() * 4
              + sizes.sum() * 16 * 8)
    flops = float(2 * (sizes ** 2).sum() * 16 * nt)
    return {"kernel": f"{kname}<{'_Float16' if nt == 2 else 'bf16_t'}>", "layout": L.layout,
            "avg_launch_us": round(ms * 1e3, 2), "alg_bytes_per_launch": alg,
            "GBs": round(alg / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "mfma_TFLOPs": round(flops / (ms * 1e-3) / 1e12, 1), **out}


def batched_extra(pb, storage, dev, steps, warmup, seed, kind="hybrid", fp64_tts=None):
    """Solve 16 right-hand sides together (FP16/BF16 MFMA local inverses, FP64 Krylov);
    FP64-verified residual of every column."""
    import numpy as np
    import torch
    from dd_solvers import problems as P

    Bh = pb.b if (pb.b.ndim == 2 and pb.b.shape[1] == 16) else \
        P.uniform_rhs(pb.A.shape[0], seed + 100, 16)[pb.decomposition.perm]
    Bh = np.ascontiguousarray(Bh)
    B = torch.as_tensor(Bh, device=dev)
    s = make_block_solver(pb, storage, dev, kind=kind)
    X, meta, total, times = time_solves(s, B, steps, warmup)
    Xh = X.cpu().numpy()
    res = np.linalg.norm(Bh - pb.A @ Xh, axis=0) / np.linalg.norm(Bh, axis=0)
    tts = total / steps
    out = {"tts_s_16rhs": round(tts, 5), "per_rhs_s": round(tts / 16, 6),
           "iterations_min_max": [int(min(meta["iterations"])), int(max(meta["iterations"]))],
           "max_true_rel_residual": float(res.max()),
           "mfma": mfma_kernel_stats(s.preconditioner)}
    if fp64_tts:
        out["per_rhs_ratio_vs_fp64"] = round(tts / 16 / fp64_tts, 4)
    log(f"batched {storage} ({kind}): 16 RHS in {tts * 1e3:.2f} ms, iterations "
        f"{out['iterations_min_max']}, max residual {res.max():.1e}, mfma {out['mfma']}")
    del s
    torch.cuda.empty_cache()
    return out


def sparse_local_extra(pb, dev, steps, warmup, xh, fp64_tts, use_graph=True, precision=None):
    """The same FP64 Hybrid solve with sparse local solves (BatchSparseCholesky:
    nested-dissection block LDL^T of every subdomain, one forward and one backward sweep
    launch per tree height) in place of the dense inverses; the factor in FP64 (exact
    local solves) or stored in FP32 (`precision`, FP64 vectors and sums)."""
    import numpy as np
    import torch
    import dd_solvers as dds

    A = torch.sparse_csr_tensor(torch.as_tensor(pb.A.indptr), torch.as_tensor(pb.A.indices),
                                torch.as_tensor(pb.A.data), pb.A.shape).to(dev)
    b = torch.as_tensor(pb.b, device=dev)
    s = dds.CG(dds.HybridSchwarz(torch.float64,
                                 dds.BatchSparseCholesky(leaf_size=64, precision=precision),
                                 dds.DenseCholesky()),
               rtol=1e-8, maxiter=2000, bsr_matmul=False, check_every=8, use_graph=use_graph)
    t = time.perf_counter()
    s.setup(A, decomposition=pb.decomposition)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    x, meta, total, times = time_solves(s, b, steps, warmup)
    M = s.preconditioner
    plan = M.local_solver.plan
    xv = torch.randn(pb.A.shape[0], dtype=torch.float64, device=dev)
    Y = M.Y if M.Y is not None else torch.empty_like(xv)
    t_loc = kernel_time_ms(lambda: M._enqueue_local(xv, Y, None))
    # algorithmic bytes of one local solve: the stacked factor (forward) and the row-major
    # W (backward), the right-hand side gathered (value + index) and the result written
    nk = int(pb.decomposition.sub_ptr[-1])
    fac_bytes = (4 if precision == torch.float32 else 8) * \
        int(np.sum(plan.fnj * plan.fnj + 2 * plan.fns * plan.fnj))
    alg = fac_bytes + nk * (8 + 4) + nk * 8
    xs = x.cpu().numpy()
    out = {"tts_s": round(total / steps, 6), "min_tts_s": round(min(times), 6),
           "iterations": meta["iterations"],
           "rel_err_vs_fp64_inverses": float(np.linalg.norm(xs - xh) / np.linalg.norm(xh)),
           "true_rel_residual": float(np.linalg.norm(pb.b - pb.A @ xs) / np.linalg.norm(pb.b)),
           "ratio_vs_fp64_inverses": round(total / steps / fp64_tts, 3),
           "setup_s": round(t_setup, 2), "fronts": int(plan.n_fronts), "heights": int(plan.n_levels),
           "local_solve_us": round(t_loc * 1e3, 2), "local_alg_bytes": int(alg),
           "factor_bytes": int(fac_bytes),
           "local_GBs": round(alg / (t_loc * 1e-3) / 1e9, 1)}
    log(f"sparse local factor ({precision or 'fp64'}): TTS {out['tts_s'] * 1e3:.2f} ms, {out['iterations']} it, local "
        f"solve {out['local_solve_us']} us ({out['local_GBs']} GB/s), err vs inverses "
        f"{out['rel_err_vs_fp64_inverses']:.1e}")
    s.destroy()
    return out


def uneven_extra(dev, steps, warmup, seed):
    """Batching on uneven subdomains: jittered 2D P1 on 512^2, 256 k-means subdomains,
    overlap 2, 16 right-hand sides; beside the FP64 single-right-hand-side solve."""
    import numpy as np
    import torch
    from dd_solvers import problems as P

    t = time.perf_counter()
    pb = P.make_problem("p1_2d_kmeans", 512, overlap=2, seed=seed, nrhs=16, n_clusters=256, jitter=0.2)
    sizes = pb.decomposition.sizes
    log(f"built P1 512^2 k-means: n={pb.A.shape[0]} subdomains={sizes.shape[0]} "
        f"k in [{sizes.min()}, {sizes.max()}] ({time.perf_counter() - t:.1f}s)")
    s = make_solver(pb, torch.float64, dev)
    b0 = torch.as_tensor(np.ascontiguousarray(pb.b[:, 0]), device=dev)
    _, m0, tot0, _ = time_solves(s, b0, steps, warmup)
    fp64 = tot0 / steps
    log(f"uneven FP64 single RHS: {fp64 * 1e3:.2f} ms, {m0['iterations']} iterations")
    del s
    torch.cuda.empty_cache()
    out = {"workload": ("2D P1 512^2 (n=261121), nodes jittered U(-0.2h, 0.2h), 256 k-means "
                        "subdomains overlap 2, Hybrid Schwarz, 16 right-hand sides, rtol 1e-8"),
           "subdomain_sizes_min_max": [int(sizes.min()), int(sizes.max())],
           "fp64_single_rhs_tts_s": round(fp64, 5), "fp64_iterations": int(m0["iterations"])}
    for name, st in (("fp16", torch.float16), ("bf16", torch.bfloat16)):
        out[f"batched_{name}"] = batched_extra(pb, st, dev, steps, warmup, seed, fp64_tts=fp64)
    return out


def local_kernel_stats(M):
    """The local inverse apply alone (the dominant kernel of a solve): average launch time
    from HIP events on the launch stream, with its algorithmic bytes per launch. For the
    symmetric layout that is sym_apply_kernel, which reads the packed lower triangles and
    the gathered x (value + index) and writes the partial vectors; sym_combine_kernel reads
    those back and writes the subdomain-ordered result."""
    import numpy as np
    import torch

    L = M.local
    sizes = M._host["sizes"].astype(np.int64)
    xv = torch.randn(M.n, dtype=torch.float64, device=M.device)
    Y = M.Y if M.Y is not None else torch.empty_like(xv)
    pack = None
    if L.layout in ("sym", "sym_mfma"):
        from dd_solvers.solvers import _sym_written_rows
        if L.direct:  # one item per subdomain writes the result: no partial vectors
            t_comb, comb = 0.0, 0
            yp_rows = int(sizes.sum())
        else:
            t_comb = kernel_time_ms(lambda: M._enqueue_sym_combine(Y, None))
            yp_rows = _sym_written_rows(sizes, L.group)
            comb = int(yp_rows * 8 + sizes.sum() * 8)
        if L.layout == "sym":
            t_loc = kernel_time_ms(lambda: M._enqueue_sym_apply(xv, None, Y))
            alg = int(L.bytes_algorithmic + sizes.sum() * (8 + 4) + yp_rows * 8)
            kname = f"sym_apply_kernel<{_TNAME[L.mat_dtype]}, {_TNAME[M.pi]}, {_TNAME[L.acc_dtype]}>"
        else:
            # the two kernels apart: the pack of x (gathered x and indices in, terms and block
            # scales out), then the apply alone on that pack (lower triangles, terms and
            # scales in, the result out)
            xp = M._xpack
            nt = 2 if L.mat_dtype == torch.float16 else 3
            xbytes = xp["n_blk"] * (nt * 128 + 4)
            gx = int(sizes.sum() * (8 + (4 if M.sub_idx is not None else 0)))  # gathered x
            t_pack = kernel_time_ms(lambda: M._enqueue_sym_pack(xv, None))
            M._enqueue_sym_pack(xv, None)
            t_ap = kernel_time_ms(lambda: M._enqueue_sym_apply_packed(None, Y))
            ap_alg = int(L.bytes_algorithmic + xbytes + yp_rows * Y.element_size())
            pack = {"kernel": f"sym_pack_x_kernel<{_TNAME[L.mat_dtype]}, {_TNAME[M.pi]}>",
                    "avg_launch_us": round(t_pack * 1e3, 2), "alg_bytes_per_launch": gx + int(xbytes),
                    "GBs": round((gx + xbytes) / (t_pack * 1e-3) / 1e9, 1)}
            kname = f"sym_mfma_apply_kernel<{_TNAME[L.mat_dtype]}, {_TNAME[Y.dtype]}>"
            if L.mat_dtype == torch.float16 and M.fused_pack:
                # what the solve runs: one kernel packing x in its LDS (gathered x and the
                # lower triangles in, the result out); the pack and the apply from a packed
                # x, launched apart, are reported beside it
                pack["apply_packed_us"] = round(t_ap * 1e3, 2)
                pack["apply_packed_alg_bytes"] = ap_alg
                pack["apply_packed_GBs"] = round(ap_alg / (t_ap * 1e-3) / 1e9, 1)
                t_loc = kernel_time_ms(lambda: M._enqueue_sym_apply(xv, None, Y))
                alg = int(L.bytes_algorithmic + gx + yp_rows * Y.element_size())
                kname = (f"sym_mfma_apply_kernel<{_TNAME[L.mat_dtype]}, {_TNAME[Y.dtype]}, "
                         f"true, {_TNAME[M.pi]}>")
            else:
                t_loc, alg = t_ap, ap_alg
    else:
        t_loc = kernel_time_ms(lambda: M._enqueue_local(xv, Y, None))
        t_comb, comb = 0.0, 0
        alg = int(L.bytes_algorithmic + sizes.sum() * (8 + 4) + sizes.sum() * 8)
        kname = f"local_apply_kernel<{_TNAME[L.mat_dtype]}, {_TNAME[M.pi]}, {_TNAME[L.acc_dtype]}>"
    gbs = alg / (t_loc * 1e-3) / 1e9
    return {"kernel": kname, "avg_launch_us": round(t_loc * 1e3, 2), "alg_bytes_per_launch": alg,
            "stored_matrix_bytes": int(L.bytes_matrices), "GBs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "combine_kernel_us": round(t_comb * 1e3, 2),
            "combine_alg_bytes": comb, **({"pack": pack} if pack else {})}


def spmv_stats(op, n, dev):
    import torch

    xv = torch.randn(n, dtype=torch.float64, device=dev)
    q = torch.empty(op.shape[0], dtype=torch.float64, device=dev)
    if op.shape[0] > n:
        xv = torch.cat([xv, torch.zeros(op.shape[0] - n, dtype=torch.float64, device=dev)])
    t = kernel_time_ms(lambda: op._enqueue_matvec(xv, q, None, None))
    gbs = op.bytes_per_matvec / (t * 1e-3) / 1e9
    out = {"format": type(op).__name__, "us": round(t * 1e3, 2), "bytes": int(op.bytes_per_matvec),
           "GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    phys = getattr(op, "bytes_physical", None)
    if phys:
        out["physical_bytes"] = int(phys)
        out["physical_hbm_frac"] = round(phys / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def solve_record(pb, s, b, steps, warmup, xh=None):
    """Time `steps` solves (after `warmup`), then the FP64-verified residual and the error
    against the FP64 solution `xh`."""
    import numpy as np

    x, meta, total, times = time_solves(s, b, steps, warmup)
    xs = x.cpu().numpy()
    out = {"tts_s": round(total / steps, 6), "min_tts_s": round(min(times), 6),
           "iterations": int(meta["iterations"]),
           "true_rel_residual": float(np.linalg.norm(pb.b - pb.A @ xs) / np.linalg.norm(pb.b))}
    if xh is not None:
        out["rel_err_vs_fp64"] = float(np.linalg.norm(xs - xh) / np.linalg.norm(xh))
    return out, xs


# the precision forms of the mixed-precision workload: (name, storage, scaled, accumulate)
PRECISION_FORMS = [("fp16", "float16", True, "float32"), ("bf16", "bfloat16", False, "float32"),
                   ("fp32", "float32", False, "float32"), ("fp16_fp64acc", "float16", True, None)]


def precision_forms(pb, dev, steps, warmup, fp64, xh, forms, use_graph=True):
    """The same solve with low-precision local inverses: storage, accumulation, TTS,
    iterations (extra vs FP64), error vs the FP64 solve, local-apply kernel stats."""
    import torch

    out = {}
    for name, st, scaled, acc in forms:
        try:
            s = make_solver(pb, getattr(torch, st), dev, scaled=scaled, use_graph=use_graph,
                            accumulate=getattr(torch, acc) if acc else None)
            r, _ = solve_record(pb, s, torch.as_tensor(pb.b, device=dev), steps, warmup, xh)
            r.update({"storage": st, "scaled": scaled, "accumulate": acc or "float64",
                      "ratio_vs_fp64": round(r["tts_s"] / fp64["tts_s"], 4),
                      "extra_iterations_vs_fp64": r["iterations"] - fp64["iterations"],
                      "local_kernel": local_kernel_stats(s.preconditioner)})
            log(f"{name}: TTS {r['tts_s'] * 1e3:.2f} ms ({r['ratio_vs_fp64']}x FP64), "
                f"{r['iterations']} it, err vs FP64 {r['rel_err_vs_fp64']:.1e}, local "
                f"{r['local_kernel']['avg_launch_us']} us ({r['local_kernel']['hbm_frac'] * 100:.1f}%)")
            s.destroy()
            del s
        except Exception as e:  # report, keep the measured line
            r = {"error": str(e)[:200]}
            log(f"{name}: error {e}")
        torch.cuda.empty_cache()
        out[name] = r
    return out


def bsr_sweep(pb, dev, steps, warmup, xh, fp64_tts):
    """The FP64 solve with the BSR SpMV at block sizes 2, 4, 8 (FastBSR, the reference's
    CG(bsr_matmul=True)): TTS, stored vs useful bytes and SpMV bandwidth. When b does not
    divide n the operator is padded and the PCG vectors carry a zero tail."""
    import torch

    out = []
    for bs in (2, 4, 8):
        try:
            s = make_solver(pb, torch.float64, dev, bsr=bs)
            r, _ = solve_record(pb, s, torch.as_tensor(pb.b, device=dev), steps, warmup, xh)
            op = s.matrix
            st = spmv_stats(op, pb.A.shape[0], dev)
            csr_bytes = 12 * pb.A.nnz + 8 * (pb.A.shape[0] + 1) + 16 * pb.A.shape[0]
            st["useful_csr_bytes"] = int(csr_bytes)
            st["useful_GBs"] = round(csr_bytes / (st["us"] * 1e-6) / 1e9, 1)
            r.update({"b": bs, "n_padded": int(op.shape[0]), "nnzb": int(op.nnzb),
                      "ratio_vs_sell": round(r["tts_s"] / fp64_tts, 4), "spmv": st})
            log(f"BSR b={bs}: TTS {r['tts_s'] * 1e3:.2f} ms, {r['iterations']} it, SpMV {st['us']} us "
                f"({st['GBs']} GB/s stored, {st['useful_GBs']} GB/s useful)")
            s.destroy()
            del s
        except Exception as e:
            r = {"b": bs, "error": str(e)[:200]}
            log(f"BSR b={bs}: error {e}")
        torch.cuda.empty_cache()
        out.append(r)
    return out


def workload_64(dev, steps, warmup, seed, use_graph=True):
    """The 64^3 system (n = 250,047, 512 boxes): FP64 and mixed FP16 forms, the BSR
    block-size sweep, exact sparse local factors, the batched forms."""
    import torch

    pb = build_problem(64, seed)
    b = torch.as_tensor(pb.b, device=dev)
    s = make_solver(pb, torch.float64, dev, use_graph=use_graph)
    fp64, xh = solve_record(pb, s, b, steps, warmup)
    fp64["local_kernel"] = local_kernel_stats(s.preconditioner)
    fp64["spmv"] = spmv_stats(s.matrix, pb.A.shape[0], dev)
    log(f"64^3 FP64 TTS {fp64['tts_s'] * 1e3:.2f} ms, {fp64['iterations']} it, local "
        f"{fp64['local_kernel']['avg_launch_us']} us")
    s.destroy()
    del s
    torch.cuda.empty_cache()
    out = {"workload": ("3D Q1 Poisson 64^3 (n=250047), lognormal k, Hybrid Schwarz 512 boxes of "
                        "8^3 elements overlap 1, dense FP64 Cholesky coarse, rtol 1e-8"),
           "fp64": fp64}
    out.update(precision_forms(pb, dev, steps, warmup, fp64, xh, PRECISION_FORMS[:1], use_graph))
    out["bsr_block_sizes"] = bsr_sweep(pb, dev, steps, warmup, xh, fp64["tts_s"])
    for name, prec in (("sparse_local_fp64", None), ("sparse_local_fp32_factor", torch.float32)):
        try:
            out[name] = sparse_local_extra(pb, dev, steps, warmup, xh, fp64["tts_s"],
                                           use_graph=use_graph, precision=prec)
        except Exception as e:
            out[name] = {"error": str(e)[:200]}
        torch.cuda.empty_cache()
    for name, st in (("fp16", torch.float16), ("bf16", torch.bfloat16)):
        try:
            out[f"batched_{name}"] = batched_extra(pb, st, dev, steps, warmup, seed,
                                                   fp64_tts=fp64["tts_s"])
        except Exception as e:
            out[f"batched_{name}"] = {"error": str(e)[:200]}
        torch.cuda.empty_cache()
    return out


def stream_peaks(dev):
    """Achievable HBM bandwidth on this GPU: dds_stream read / copy / triad over 1 GiB
    arrays (beyond the 256 MB MALL), best grid of {1, 2, 4, 8} workgroups per CU, and the
    reference's own method (a tensor clone, experiments/memory_bandwidth.ipynb)."""
    import torch
    from dd_solvers import _lib

    n = 1 << 27  # doubles per array (1 GiB)
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.ones(n, dtype=torch.float64, device=dev)
    y = torch.empty(n, dtype=torch.float64, device=dev)
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    s = _lib.stream()
    out = {}
    for op, name, moved in ((0, "read", 8), (1, "copy", 16), (2, "triad", 24)):
        best = 0.0
        for per_cu in (1, 2, 4, 8):
            grid = min(n_cu * per_cu, _lib.MAX_PARTIALS)
            ms = kernel_time_ms(lambda: _lib.call("dds_stream", op, n, a.data_ptr(), b.data_ptr(),
                                                  0.5, y.data_ptr(), grid, s), reps=10)
            best = max(best, moved * n / (ms * 1e-3) / 1e9)
        out[f"{name}_GBs"] = round(best, 1)
    ms = kernel_time_ms(lambda: a.clone(), reps=10)
    out["torch_clone_GBs"] = round(16 * n / (ms * 1e-3) / 1e9, 1)
    del a, b, y
    torch.cuda.empty_cache()
    log(f"stream: {out}")
    return out


# The reference's own batched-GEMV benchmark (experiments/benchmark_batched_gemv.py: n =
# 2^27 / k^2 column-major k x k matrices, one vector each) on an A100 80GB: the best of its
# two kernels, min over repetitions, from results/benchmark_batched_gemv.csv (line numbers).
REF_GEMV_A100_MS = {
    (3, "float64", "float64"): (1.077, 35), (10, "float64", "float64"): (0.8179, 183),
    (20, "float64", "float64"): (0.7372, 393), (48, "float64", "float64"): (0.6623, 981),
    (128, "float64", "float64"): (0.6129, 2661),
    (3, "float16", "float64"): (0.6003, 38), (10, "float16", "float64"): (0.3248, 186),
    (20, "float16", "float64"): (0.2525, 396), (48, "float16", "float64"): (0.2118, 984),
    (128, "float16", "float64"): (0.1916, 2664),
    (3, "float32", "float32"): (0.5558, 23), (20, "float32", "float32"): (0.3751, 381),
    (128, "float32", "float32"): (0.3318, 2649),
    (3, "float16", "float32"): (0.4034, 26), (20, "float16", "float32"): (0.2212, 384),
    (128, "float16", "float32"): (0.1834, 2652),
}


def gemv_reference_shapes(dev):
    """dd_solvers.gemv_strided_batched (the drop-in op) on the reference benchmark's shapes,
    best of its two kernels (use_shared_memory False / True), beside the A100 times."""
    import torch
    import dd_solvers as dds

    out = []
    for (k, ad, bd), (ref_ms, line) in REF_GEMV_A100_MS.items():
        n = (1 << 27) // (k * k)
        A = torch.randn((n, k, k), device=dev, dtype=getattr(torch, ad)).transpose(1, 2).contiguous().transpose(1, 2)
        b = torch.randn((n, k), device=dev, dtype=getattr(torch, bd))
        ms = min(kernel_time_ms(lambda: dds.gemv_strided_batched(A, b, use_shared_memory=u), reps=20)
                 for u in (False, True))
        nbytes = n * k * k * A.element_size() + 2 * n * k * b.element_size()
        out.append({"k": k, "A": ad, "b": bd, "ms": round(ms, 4), "GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                    "reference_a100_ms": ref_ms, "speedup_vs_a100": round(ref_ms / ms, 2),
                    "reference_csv_line": line})
        del A, b
    torch.cuda.empty_cache()
    log("gemv vs reference A100: " + ", ".join(f"k={o['k']} {o['A'][:4]}/{o['b'][:4]} {o['ms']} ms "
                                                f"({o['speedup_vs_a100']}x)" for o in out))
    return out


# The reference's BSR SpMV benchmark (results/benchmark_bsrmv.csv, A100 80GB, FP64): 2D
# DG-SIPG matrices on the uniform mesh S_m (2 * 4^m triangles), block size b = DoFs per
# element, one block per triangle and per interior edge (each direction); its RPT-I kernel
# (the FastBSR default backend), min over repetitions (line numbers).
REF_BSR_A100_MS = {
    (11, 3): (1.9776, 3), (10, 6): (1.9016, 10), (10, 10): (4.6802, 17), (9, 15): (2.4548, 24),
    (8, 21): (1.1619, 31),
}


# The same benchmark's 3D rows: DG-SIPG on the tetrahedral S_m (6 * 8^m cells), one block
# per tetrahedron and per interior face (each direction); RPT-I, min over repetitions.
REF_BSR3D_A100_MS = {(7, 4): (6.7186, 38), (6, 10): (4.3284, 45), (5, 20): (1.9365, 52)}


def dg_tet_bsr(m: int, b: int, dev, seed: int = 0):
    """A BSR matrix with the block structure of DG on the tetrahedral mesh S_m
    (dg.UniformTetMesh: the 2^m grid of cubes, six Kuhn tetrahedra each), random FP64
    values; the block counts of the reference's 3D rows. Built on the device: the faces are
    matched by the sum of their vertices' grid coordinates (unique in a Kuhn mesh)."""
    import torch
    from dd_solvers.dg import _KUHN_ORDERS

    n = 1 << m
    cube = torch.arange(n ** 3, device=dev)
    base = torch.stack([cube % n, (cube // n) % n, cube // (n * n)], 1)
    eye = torch.eye(3, dtype=torch.int64, device=dev)
    path = torch.zeros((6, 4, 3), dtype=torch.int64, device=dev)
    for t, (a, b_, c) in enumerate(_KUHN_ORDERS.tolist()):
        path[t, 1] = eye[a]
        path[t, 2] = eye[a] + eye[b_]
        path[t, 3] = 1
    P = (base[:, None, None, :] + path[None]).reshape(-1, 4, 3)      # (T, 4, 3) grid coords
    T = P.shape[0]
    lf = torch.tensor([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]], device=dev)
    S = P[:, lf].sum(2).reshape(-1, 3)                                 # (4T, 3)
    key = (S[:, 2] * (3 * n + 1) + S[:, 1]) * (3 * n + 1) + S[:, 0]
    del P, S
    ks, order = torch.sort(key, stable=True)
    pair = ks[1:] == ks[:-1]
    i0 = torch.nonzero(pair).flatten()
    c1, c2 = order[i0] // 4, order[i0 + 1] // 4
    del key, ks, order, pair, i0
    rows = torch.cat([torch.arange(T, device=dev), c1, c2])
    cols = torch.cat([torch.arange(T, device=dev), c2, c1])
    del c1, c2
    o = torch.argsort(rows * T + cols)
    rows, cols = rows[o], cols[o]
    crow = torch.zeros(T + 1, dtype=torch.int64, device=dev)
    crow[1:] = torch.cumsum(torch.bincount(rows, minlength=T), 0)
    g = torch.Generator(device=dev).manual_seed(seed)
    vals = torch.rand((cols.shape[0], b, b), dtype=torch.float64, device=dev, generator=g)
    return torch.sparse_bsr_tensor(crow, cols.to(torch.int32), vals, (T * b, T * b))


def dg_triangle_bsr(m: int, b: int, dev, seed: int = 0):
    """A BSR matrix with the block structure of DG on the uniform triangle mesh S_m (N = 2^m
    squares per side, each cut into a lower and an upper triangle; a block per triangle and
    per interior edge), random FP64 values. Same block count as the reference's matrices."""
    import torch

    N = 1 << m
    ij = torch.arange(N * N, device=dev)
    i, j = ij // N, ij % N
    t0 = 2 * ij
    # lower triangle (bottom, left, diagonal neighbours), upper (diagonal, right, top)
    c0 = torch.stack([2 * ((i - 1) * N + j) + 1, 2 * (i * N + j - 1) + 1, t0, t0 + 1], 1)
    v0 = torch.stack([i > 0, j > 0, torch.ones_like(i, dtype=torch.bool), torch.ones_like(i, dtype=torch.bool)], 1)
    c1 = torch.stack([t0, t0 + 1, 2 * (i * N + j + 1), 2 * ((i + 1) * N + j)], 1)
    v1 = torch.stack([torch.ones_like(i, dtype=torch.bool), torch.ones_like(i, dtype=torch.bool), j < N - 1, i < N - 1], 1)
    cols = torch.stack([c0, c1], 1).reshape(-1, 4)   # triangle order t0, t1, t0, t1, ...
    valid = torch.stack([v0, v1], 1).reshape(-1, 4)
    counts = valid.sum(1)
    crow = torch.zeros(cols.shape[0] + 1, dtype=torch.int64, device=dev)
    crow[1:] = torch.cumsum(counts, 0)
    col = cols[valid].to(torch.int32)
    g = torch.Generator(device=dev).manual_seed(seed)
    vals = torch.rand((col.shape[0], b, b), dtype=torch.float64, device=dev, generator=g)
    n = cols.shape[0] * b
    return torch.sparse_bsr_tensor(crow, col, vals, (n, n))


def bsr_reference_shapes(dev):
    """FastBSR (the drop-in class) on the reference's DG benchmark shapes, 2D and 3D."""
    import torch
    import dd_solvers as dds

    out = []
    shapes = [(2, m, b, ref) for (m, b), ref in REF_BSR_A100_MS.items()] + \
             [(3, m, b, ref) for (m, b), ref in REF_BSR3D_A100_MS.items()]
    for dim, m, b, (ref_ms, line) in shapes:
        A = dg_triangle_bsr(m, b, dev) if dim == 2 else dg_tet_bsr(m, b, dev)
        op = dds.FastBSR(A)
        x = torch.rand(A.shape[1], dtype=torch.float64, device=dev)
        y = torch.empty(A.shape[0], dtype=torch.float64, device=dev)
        ms = kernel_time_ms(lambda: op._enqueue_matvec(x, y, None, None), reps=20)
        # the reference's GB/s: values + block columns + block-row pointers + x + y
        nbytes = op.bytes_per_matvec
        out.append({"dim": dim, "mesh": f"S{m}", "b": b, "dofs": int(A.shape[0]), "nnzb": op.nnzb,
                    "ms": round(ms, 4), "GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                    "reference_a100_ms": ref_ms, "speedup_vs_a100": round(ref_ms / ms, 2),
                    "reference_csv_line": line})
        del A, op, x, y
        torch.cuda.empty_cache()
    log("BSR vs reference A100: " + ", ".join(f"{o['dim']}D S{o['mesh'][1:]} b={o['b']} {o['ms']} ms "
                                               f"({o['GBs']} GB/s, {o['speedup_vs_a100']}x)" for o in out))
    return out


# The reference's precision experiment (results/experiment_precisions.csv, rtol 1e-9, A100)
# on its own DG systems: (p, fine mesh, subdomain mesh, coarse mesh).
REF_EXPERIMENTS = [(1, 10, "S8", "S8"), (1, 10, "S9", "S9"), (3, 9, "S8", "S8"), (3, 9, "S9", "S9"),
                   (3, 9, "S7", "S7")]
# run only when named in --reference-only: their host-side set-up (the S11 assembly, the
# nested-dissection plan of a 2-million-cell coarse space) takes minutes
REF_EXPERIMENTS_LARGE = [(1, 10, "S10", "S10"), (1, 11, "S9", "S9"), (1, 11, "S10", "S10"),
                         (1, 11, "S11", "S11"), (5, 8, "S6", "S6"), (5, 8, "S7", "S7"),
                         (5, 8, "S8", "S8")]
REF_SOLVERS = [("hybrid", "float64", None), ("hybrid", "float64", "float16"),
               ("additive", "float64", None), ("additive", "float32", "float16")]
# every configuration of the published precision experiment (--reference-solvers all):
# preconditioner precision x local-inverse storage, Additive and Hybrid
REF_SOLVERS_ALL = [(kind, pi, inv) for kind in ("hybrid", "additive")
                   for pi in ("float64", "float32")
                   for inv in ((None, "float32", "float16", "bfloat16") if pi == "float64"
                               else (None, "float16", "bfloat16"))]
# the reference's parallelism experiment (experiments/parallelism_experiments.py): p = 1 on
# S11 with subdomains and coarse cells from S_m and C_m, m = 8..11, two preconditioners;
# run when named "par_p1_S11_<solvers>_<coarse>" in --reference-only
REF_PARALLELISM = [(sm, cm) for kind in ("S", "C", "mixed") for c in range(8, 12)
                   for sm_ in range(c, 12)
                   for sm, cm in [(("C" if kind == "C" else "S") + str(sm_),
                                   ("S" if kind == "S" else "C") + str(c))]]
PAR_SOLVERS = [("additive", "float32", "float16"), ("hybrid", "float64", "float16")]


def reference_experiments(dev, steps, warmup, only=None, all_solvers=False):
    """The reference's DG experiment runs with this build in every slot (the BSR SpMV with
    block size = DoFs per cell, the local inverses in the reference's precisions, the
    sparse coarse solver for its CUDSS slot), graph-replayed iteration blocks, min over
    `steps` solves; beside the reference's published iteration count and best A100 time."""
    import numpy as np
    import torch
    import dd_solvers as dds
    from dd_solvers import dg

    with open(os.path.join(ROOT, "tests", "golden", "dg_published.json")) as f:
        pub = json.load(f)["solves"]
    T = {"float64": torch.float64, "float32": torch.float32, "float16": torch.float16,
         "bfloat16": torch.bfloat16}
    out = []
    meshes = {}
    fams = [(p, fine, sm, cm, "precisions", REF_SOLVERS_ALL if all_solvers else REF_SOLVERS,
             f"p{p}_S{fine}_{sm}_{cm}")
            for p, fine, sm, cm in REF_EXPERIMENTS + (REF_EXPERIMENTS_LARGE if only else [])]
    if only:
        fams += [(1, 11, sm, cm, "parallelism", PAR_SOLVERS, f"par_p1_S11_{sm}_{cm}")
                 for sm, cm in REF_PARALLELISM]
    for p, fine, sm, cm, experiment, solvers, name in fams:
        if only and name not in only:
            continue
        t = time.perf_counter()
        if (p, fine) not in meshes:
            meshes.clear()
            meshes[(p, fine)] = dg.assemble_sipg(dg.UniformTriangleMesh(fine), p)
        pr = meshes[(p, fine)]
        perm, A, bh, k, spc = dg.dg_experiment(pr, sm, cm)
        At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr.astype(np.int64)),
                                     torch.as_tensor(A.indices.astype(np.int64)),
                                     torch.as_tensor(A.data), A.shape).to(dev)
        b = torch.as_tensor(bh, device=dev)
        log(f"DG p={p} S{fine} solvers {sm} coarse {cm}: n={A.shape[0]} k={k} n_c={spc.shape[0]} "
            f"({time.perf_counter() - t:.1f}s)")
        for kind, pi, inv in solvers:
            ref = [r for r in pub if (r["experiment"], r["p"], r["fine"], r["solvers"], r["coarse"],
                                      r["kind"], r["precision"], r["inv"])
                   == (experiment, p, f"S{fine}", sm, cm, kind, pi, inv)]
            ref = ref[0] if ref else {}
            row = {"experiment": experiment, "p": p, "fine": f"S{fine}", "solvers": sm,
                   "coarse": cm, "kind": kind,
                   "precision": pi, "inv": inv, "dofs": int(A.shape[0]), "n_coarse": int(spc.shape[0]),
                   "reference_solver": ref.get("solver"),
                   "reference_iterations": ref.get("iterations"),
                   "reference_a100_ms": None if ref.get("best_ms") is None else round(ref["best_ms"], 2),
                   "reference_a100_setup_s": None if ref.get("setup_ms") is None
                   else round(ref["setup_ms"] / 1e3, 2)}
            try:
                cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
                P = cls(T[pi], dds.Inv(T[inv] if inv else None), dds.SparseCholesky())
                s = dds.CG(P, rtol=1e-9, maxiter=2000, use_graph=True)
                # the reference's protocol: solver.setup alone (experiments.py:235-241), its
                # component timings are not collected during it
                torch.cuda.synchronize()
                t = time.perf_counter()
                s.setup(At, dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
                        block_size=pr.block_size)
                torch.cuda.synchronize()
                setup_s = time.perf_counter() - t
                from dd_solvers.solvers import timeit
                x_c = torch.rand(P.n_coarse, device=dev, dtype=P.pi)
                x_f = torch.rand(P.n, device=dev, dtype=P.pi)
                tm = {"coarse solver solve time": timeit(lambda: P.coarse_solver.solve(x_c), 3, 10, 3),
                      "local solver solve time": timeit(lambda: P._local_only(x_f), 3, 10, 3)}
                for _ in range(warmup):
                    s.solve(b)
                best = None
                for _ in range(steps):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    x, md = s.solve(b)
                    e1.record()
                    torch.cuda.synchronize()
                    ms = e0.elapsed_time(e1)
                    best = ms if best is None else min(best, ms)
                xh = x.double().cpu().numpy()
                res = float(np.linalg.norm(A @ xh - bh) / np.linalg.norm(bh))
                row.update({"iterations": md["iterations"], "ms": round(best, 2),
                            "true_rel_residual": res, "setup_s": round(setup_s, 2),
                            "local_solve_ms": round(tm["local solver solve time"], 4),
                            "coarse_solve_ms": round(tm["coarse solver solve time"], 4),
                            "local_layout": P.local.layout if P.local is not None else None,
                            "coarse_levels": P.coarse_solver.plan.n_levels})
                if row["reference_a100_ms"]:
                    row["speedup_vs_a100"] = round(row["reference_a100_ms"] / best, 2)
                log(f"  {kind} {pi}+{inv}: {md['iterations']} it (reference {row['reference_iterations']}), "
                    f"{best:.1f} ms (A100 {row['reference_a100_ms']} ms), local {row['local_solve_ms']} ms, "
                    f"coarse {row['coarse_solve_ms']} ms, setup {setup_s:.2f} s (A100 "
                    f"{row['reference_a100_setup_s']} s)")
                s.destroy()
                del s, P
            except Exception as e:  # record and go on with the others
                row["error"] = str(e)[:200]
                if ref and ref.get("iterations") is None:
                    row["reference_raised"] = True  # the published run raised as well
                log(f"  {kind} {pi}+{inv}: error {e} (reference "
                    f"{'raised too' if row.get('reference_raised') else row['reference_iterations']})")
            torch.cuda.empty_cache()
            out.append(row)
        del At, A, b
    return out


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 PMC
    measurement (2 x FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            return int(json.load(f)[workload]["kernels"][kernel]["hbm_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def time_solves(solver, b, steps, warmup):
    import torch

    for _ in range(warmup):
        solver.solve(b)
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    meta = None
    for _ in range(steps):
        t = time.perf_counter()
        x, meta = solver.solve(b)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    return x, meta, total, times


def kernel_time_ms(fn, reps=20):
    """Average duration of `fn`'s launches, HIP events on the current stream."""
    import torch

    s = torch.cuda.current_stream()
    for _ in range(3):
        fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def cpu_baseline(pb, iterations, storage="float16", scaled=True, accumulate="float32",
                 sample_subdomains=128, seed=0):
    """The oracle (a numpy/scipy restatement of the reference's Hybrid Schwarz PCG) timed on
    this host on a bounded sample of the workload. Every piece of one PCG iteration is
    timed with the oracle's own functions: the SpMV (scipy CSR, single-threaded), R0A y,
    the coarse Cholesky solve, the prolongation and the fused-free vector updates on full
    vectors; the local inverse apply on `sample_subdomains` seeded random subdomains
    (their inverses formed as the oracle forms them: FP64 inverse, quantised to `storage`,
    scaled per subdomain, multiplied in `accumulate`), scaled up by the stored entries of
    all subdomains over those of the sample. Time to solution = start-up (T0, first
    residual) + `iterations` x per-iteration; setup is excluded, as on the GPU."""
    import numpy as np

    from oracle.schwarz import SchwarzOracle

    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:  # pragma: no cover
        threads = 1
    d = pb.decomposition
    A = pb.A
    n = A.shape[0]
    sizes = np.diff(d.sub_ptr).astype(np.int64)
    rng = np.random.default_rng(seed)
    sample = np.sort(rng.choice(sizes.shape[0], size=min(sample_subdomains, sizes.shape[0]),
                                replace=False))
    t = time.perf_counter()
    O = SchwarzOracle(A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind="hybrid",
                      storage=storage, scaled=scaled, accumulate=accumulate, local_subset=sample)
    t_setup = time.perf_counter() - t

    def best(fn, reps=3):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    v = np.asarray(pb.b, dtype=np.float64)
    w = v.copy()
    t_spmv = best(lambda: A @ v)
    t_local_sample = best(lambda: O.local(v))
    scale_up = float((sizes.astype(np.float64) ** 2).sum() / (sizes[sample].astype(np.float64) ** 2).sum())
    t_local = t_local_sample * scale_up
    zc = O.R0A @ v
    t_r0a = best(lambda: O.R0A @ v)
    t_coarse = best(lambda: O.coarse_solve(zc))
    yc = O.coarse_solve(zc)
    t_prolong = best(lambda: v - O.prolong(yc))

    def vec_ops():  # one PCG iteration's vector work (oracle/pcg.py): 2 dots, 3 updates, a norm
        rho = float(v @ w)
        p = w + 0.5 * v
        a = rho / float(p @ v)
        x = v + a * p
        r = w - a * p
        return float(np.sqrt(r @ r)) + x[0]
    t_vec = best(vec_ops)
    t_start = best(lambda: A @ O.T0(v)) + t_vec
    per_iter = t_spmv + t_local + t_r0a + t_coarse + t_prolong + t_vec
    tts = t_start + per_iter * iterations
    del O
    return {
        "value": round(tts, 4),
        "unit": "s",
        "cores": int(threads),
        "kind": "port",
        "sample": (f"oracle Hybrid Schwarz PCG (numpy/scipy, {threads} BLAS threads) on the full "
                   f"{n}-unknown system; per iteration: SpMV {t_spmv * 1e3:.1f} ms (scipy, 1 thread), "
                   f"local inverses ({storage} storage, scaled={scaled}, {accumulate} products) "
                   f"timed on {sample.shape[0]} of {sizes.shape[0]} subdomains "
                   f"({t_local_sample * 1e3:.1f} ms) and scaled by stored entries x{scale_up:.1f} "
                   f"-> {t_local * 1e3:.0f} ms, R0A {t_r0a * 1e3:.1f} ms, coarse Cholesky solve "
                   f"{t_coarse * 1e3:.1f} ms, prolongation {t_prolong * 1e3:.1f} ms, vectors "
                   f"{t_vec * 1e3:.1f} ms; time to solution = start-up {t_start * 1e3:.0f} ms + "
                   f"{iterations} iterations (the GPU solve's count) x {per_iter * 1e3:.0f} ms; "
                   f"setup excluded ({t_setup:.1f} s for the sample); cpu_count={os.cpu_count()}"),
    }


def profile_kernel_only(args, dev):
    """Counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): after setup, only the
    measured kernels, 20 launches each, then a 1 GiB device copy 20 times to calibrate the
    byte counters."""
    import numpy as np
    import torch
    import dd_solvers as dds
    from dd_solvers import _lib

    pb = build_problem(args.N, args.seed)
    st = getattr(torch, args.storage)
    if args.layout == "sparse":
        A = torch.sparse_csr_tensor(torch.as_tensor(pb.A.indptr), torch.as_tensor(pb.A.indices),
                                    torch.as_tensor(pb.A.data), pb.A.shape).to(dev)
        solver = dds.CG(dds.HybridSchwarz(torch.float64, dds.BatchSparseCholesky(leaf_size=64),
                                          dds.DenseCholesky()), rtol=1e-8, bsr_matmul=False)
        solver.setup(A, decomposition=pb.decomposition)
        M = solver.preconditioner
        xv = torch.randn(pb.A.shape[0], dtype=torch.float64, device=dev)
        Y = M.Y if M.Y is not None else torch.empty_like(xv)
        torch.cuda.synchronize()
        for _ in range(20):
            M._enqueue_local(xv, Y, None)
        torch.cuda.synchronize()
        plan = M.local_solver.plan
        out = {"profile_kernel": True, "layout": "sparse", "heights": int(plan.n_levels),
               "fronts": int(plan.n_fronts),
               "factor_bytes": 8 * int(np.sum(plan.fnj * plan.fnj + 2 * plan.fns * plan.fnj))}
        print(json.dumps(out))
        return
    if args.layout in ("mfma", "sym16"):
        solver = make_block_solver(pb, st, dev, use_graph=False, autotune=False,
                                   layout="mfma" if args.layout == "mfma" else "sym_mfma")
    else:
        acc = torch.float32 if st in (torch.float16, torch.bfloat16, torch.float32) else None
        solver = make_solver(pb, st, dev, scaled=st == torch.float16, accumulate=acc)
    M = solver.preconditioner
    L = M.local
    xv = torch.randn(pb.A.shape[0], dtype=torch.float64, device=dev)
    if args.layout in ("mfma", "sym16"):
        info = mfma_kernel_stats(M)  # 3 + 20 launches, event-timed
    else:
        info = local_kernel_stats(M)  # 3 + 20 launches of each kernel, event-timed
    q = torch.empty_like(xv)
    for _ in range(20):
        solver.matrix._enqueue_matvec(xv, q, None, None)
    torch.cuda.synchronize()
    src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(20):
        dst.copy_(src)
    torch.cuda.synchronize()
    out = {"profile_kernel": True, "N": args.N, "layout": L.layout, "storage": args.storage,
           "local": info, "spmv_alg_bytes": solver.matrix.bytes_per_matvec,
           "copy_bytes_each_way": src.numel() * 8}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline workload (FP64, the mixed forms, the batched form)")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the reference's DG precision-experiment runs")
    ap.add_argument("--reference-only", default=None, metavar="RUNS",
                    help="only the reference's DG runs named here (comma-separated, e.g. "
                         "p1_S10_S8_S8), printed as one JSON line")
    ap.add_argument("--reference-solvers", default="four", choices=["four", "all"],
                    help="with --reference-only: the four solver configurations of the default "
                         "run, or all 14 of the published precision experiment")
    ap.add_argument("--eager", action="store_true",
                    help="enqueue every iteration's kernels from Python instead of replaying a "
                         "captured hipGraph of each block of iterations")
    ap.add_argument("--profile-kernel", action="store_true",
                    help="after setup only launch the local apply, the SpMV and a 1 GiB device "
                         "copy, 20 times each (for rocprofv3 --pmc byte counters)")
    ap.add_argument("--storage", default="float16", choices=["float64", "float32", "float16", "bfloat16"])
    ap.add_argument("--layout", default="sym", choices=["sym", "mfma", "sym16", "sparse"],
                    help="--profile-kernel: the local-inverse layout to launch (mfma: 16 "
                         "right-hand sides on the full inverses; sym16: 16 on the symmetric "
                         "tiles; sparse: the sweeps of the sparse local factors, "
                         "BatchSparseCholesky)")
    ap.add_argument("--solve-only", default=None, choices=["fp64", "fp16"],
                    help="only set up and time this form of the headline workload (for "
                         "rocprofv3 --kernel-trace runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import dd_solvers as dds  # noqa: F401

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    _tnames()
    use_graph = not args.eager
    if args.reference_only:
        runs = reference_experiments(dev, args.steps, args.warmup, args.reference_only.split(","),
                                     all_solvers=args.reference_solvers == "all")
        print(json.dumps({"reference_experiments": runs}), flush=True)
        return
    if args.profile_kernel:
        profile_kernel_only(args, dev)
        return

    pb = build_problem(args.N, args.seed)
    b = torch.as_tensor(pb.b, device=dev)
    n = pb.A.shape[0]
    if args.solve_only:
        st = torch.float64 if args.solve_only == "fp64" else torch.float16
        s = make_solver(pb, st, dev, scaled=st == torch.float16, use_graph=use_graph,
                        accumulate=torch.float32 if st == torch.float16 else None)
        r, _ = solve_record(pb, s, b, args.steps, args.warmup)
        print(json.dumps({"solve_only": args.solve_only, "N": args.N, **r}), flush=True)
        return

    # ---- plain FP64 on the same system (the denominator of the ratio) ---------------------
    s = make_solver(pb, torch.float64, dev, use_graph=use_graph)
    fp64, xh = solve_record(pb, s, b, args.steps, args.warmup)
    fp64["local_kernel"] = local_kernel_stats(s.preconditioner)
    spmv = spmv_stats(s.matrix, n, dev)
    fp64_iter_bytes = s.preconditioner.bytes_per_apply + s.matrix.bytes_per_matvec + 9 * 8 * n
    log(f"FP64 TTS {fp64['tts_s'] * 1e3:.2f} ms, {fp64['iterations']} it, residual "
        f"{fp64['true_rel_residual']:.1e}; local {fp64['local_kernel']['avg_launch_us']} us; SpMV "
        f"{spmv['us']} us ({spmv['hbm_frac'] * 100:.1f}%)")
    s.destroy()
    del s
    torch.cuda.empty_cache()

    # ---- the mixed-precision form (the measured line): FP16 storage, FP32 accumulation ----
    s = make_solver(pb, torch.float16, dev, scaled=True, use_graph=use_graph, accumulate=torch.float32)
    mixed, _ = solve_record(pb, s, b, args.steps, args.warmup, xh)
    M = s.preconditioner
    kern = local_kernel_stats(M)
    iter_bytes = M.bytes_per_apply + s.matrix.bytes_per_matvec + 9 * 8 * n
    solve_bytes = iter_bytes * mixed["iterations"] + s.matrix.bytes_per_matvec
    mixed.update({"storage": "float16", "scaled": True, "accumulate": "float32",
                  "ratio_vs_fp64": round(mixed["tts_s"] / fp64["tts_s"], 4),
                  "extra_iterations_vs_fp64": mixed["iterations"] - fp64["iterations"]})
    log(f"mixed FP16 TTS {mixed['tts_s'] * 1e3:.2f} ms ({mixed['ratio_vs_fp64']}x FP64), "
        f"{mixed['iterations']} it, err vs FP64 {mixed['rel_err_vs_fp64']:.1e}; local "
        f"{kern['avg_launch_us']} us ({kern['hbm_frac'] * 100:.1f}%)")
    s.destroy()
    del s, M
    torch.cuda.empty_cache()
    forms = precision_forms(pb, dev, args.steps, args.warmup, fp64, xh, PRECISION_FORMS[1:], use_graph)
    try:
        batched = batched_extra(pb, torch.float16, dev, min(args.steps, 3), args.warmup, args.seed,
                                fp64_tts=fp64["tts_s"])
    except Exception as e:
        batched = {"error": str(e)[:200]}
    torch.cuda.empty_cache()

    extra = {}
    if not args.no_extra:
        for name, fn in (("q1_64", lambda: workload_64(dev, args.steps, args.warmup, args.seed, use_graph)),
                         ("batching_uneven", lambda: uneven_extra(dev, args.steps, args.warmup, args.seed)),
                         ("stream", lambda: stream_peaks(dev)),
                         ("gemv_reference_benchmark", lambda: gemv_reference_shapes(dev)),
                         ("bsr_reference_benchmark", lambda: bsr_reference_shapes(dev))):
            try:
                extra[name] = fn()
            except Exception as e:  # report, keep the measured line
                extra[name] = {"error": str(e)[:200]}
            torch.cuda.empty_cache()
        if not args.no_reference:
            try:
                extra["reference_experiments"] = reference_experiments(dev, 3, 1)
            except Exception as e:
                extra["reference_experiments"] = {"error": str(e)[:200]}

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(pb, mixed["iterations"], seed=args.seed)
        log(f"cpu baseline {cpu['value']} s ({time.perf_counter() - t:.1f}s to measure)")

    wl = f"q1_{args.N}_float16_sym"
    out = {
        "metric": "PCG time-to-solution @rtol 1e-8 (s), mixed/batched vs plain FP64, %HBM peak",
        "value": mixed["tts_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["tts_s"] * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 storage, f32 accumulation, f64 Krylov",
        "data": "synthetic",
        "config": {
            "workload": f"3D Q1 Poisson {args.N}^3 (n={n}), lognormal k=exp(N(0,1)), Hybrid Schwarz "
                        f"{pb.decomposition.n_subdomains} boxes of 8^3 elements overlap 1, local inverses "
                        "stored FP16 (per-subdomain scaled) applied with FP32 accumulation, FP64 PCG "
                        "vectors/dots/residual, dense FP64 Cholesky coarse, SELL-64 SpMV, rtol 1e-8, "
                        "x0 = 0, one right-hand side",
            "n": int(n),
            "nnz": int(pb.A.nnz),
            "subdomains": int(pb.decomposition.n_subdomains),
            "iterations": mixed["iterations"],
            "true_rel_residual": mixed["true_rel_residual"],
            "rel_err_vs_fp64": mixed["rel_err_vs_fp64"],
            "min_tts_s": mixed["min_tts_s"],
            "fp64_tts_s": fp64["tts_s"],
            "fp64_iterations": fp64["iterations"],
            "ratio_vs_fp64": mixed["ratio_vs_fp64"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kern["kernel"],
            "achieved": kern["GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": kern["hbm_frac"],
            "traffic": measured_traffic(wl, "sym_apply"),
            "alg_bytes_per_launch": kern["alg_bytes_per_launch"],
            "avg_launch_us": kern["avg_launch_us"],
            "stored_matrix_bytes": kern["stored_matrix_bytes"],
            "combine_kernel_us": kern["combine_kernel_us"],
            "combine_alg_bytes": kern["combine_alg_bytes"],
            "pack": kern.get("pack"),
            "pack_traffic": measured_traffic(wl, "sym_pack"),
        },
        "solve_roofline": {"alg_bytes": int(solve_bytes),
                           "achieved_GBs": round(solve_bytes / mixed["tts_s"] / 1e9, 1),
                           "frac": round(solve_bytes / mixed["tts_s"] / 1e9 / HBM_PEAK_GBS, 4)},
        "fp64": {**fp64, "solve_alg_bytes": int(fp64_iter_bytes * fp64["iterations"]),
                 "traffic": measured_traffic(f"q1_{args.N}_float64_sym", "sym_apply")},
        "mixed_fp16": mixed,
        **{f"mixed_{k}": v for k, v in forms.items()},
        "batched_fp16": batched,
        "spmv": spmv,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
