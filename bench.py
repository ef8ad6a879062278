"""Time-to-solution benchmark of Schwarz-preconditioned CG on one MI355X.

Workload (default): 3D Poisson -div(k grad u) = f, Q1 hexahedra on a 64^3 grid
(n = 63^3 = 250,047 unknowns), k = exp(g) per element with g ~ N(0, 1), f ~ U(-1, 1) per
unknown (seed 0), Hybrid Schwarz with 512 boxes of 8^3 elements and overlap 1, FP64 local
inverses, dense FP64 Cholesky coarse solve (n_c = 512), FP64 PCG to rtol 1e-8, x0 = 0.

One step = one complete solve (time to solution), inputs already resident in HBM, setup
excluded (the reference's protocol). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--N 64] [--no-cpu] [--no-extra]
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(N: int, seed: int):
    from dd_solvers import problems as P

    t = time.perf_counter()
    pb = P.make_problem("q1_3d", N, box=8, overlap=1, seed=seed, lognormal=True)
    log(f"assembled Q1 {N}^3: n={pb.A.shape[0]} nnz={pb.A.nnz} subdomains="
        f"{pb.decomposition.n_subdomains} ({time.perf_counter() - t:.1f}s)")
    return pb


def make_solver(pb, storage, dev, bsr=None, scaled=False, use_graph=False):
    import torch
    import dd_solvers as dds

    A = torch.sparse_csr_tensor(torch.as_tensor(pb.A.indptr), torch.as_tensor(pb.A.indices),
                                torch.as_tensor(pb.A.data), pb.A.shape).to(dev)
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(storage, scaled=scaled),
                                 dds.DenseCholesky()),
               rtol=1e-8, maxiter=2000, bsr_matmul=bsr is not None, check_every=8,
               use_graph=use_graph)
    kw = {"block_size": bsr} if bsr else {}
    t = time.perf_counter()
    s.setup(A, decomposition=pb.decomposition, **kw)
    torch.cuda.synchronize()
    log(f"setup storage={storage} bsr={bsr}: {time.perf_counter() - t:.1f}s")
    return s


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


def cpu_baseline(pb, iterations, budget_s=20.0):
    """The oracle's FP64 PCG components timed on this host on a bounded sample, scaled to
    one full solve: scipy CSR SpMV (whole matrix), numpy dense local inverse matvecs on a
    sample of subdomains (scaled by sum k_i^2), scipy cho_solve of the coarse system, and
    the PCG vector updates."""
    import numpy as np
    import scipy.linalg as sla

    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:  # pragma: no cover
        threads = 1
    A = pb.A
    d = pb.decomposition
    n = A.shape[0]
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n)

    def best(f, reps=3):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t)
        return min(ts)

    t_spmv = best(lambda: A @ x)
    sizes = d.sizes
    # Local inverse matvecs: a sample of subdomains, with as many distinct matrices as it
    # takes to exceed the host caches (>= 1 GiB), so the timing streams from DRAM as the
    # full set of inverses would. A matvec's time does not depend on the values, so the
    # sample's matrices are random with the sampled subdomains' sizes.
    sample = rng.choice(d.n_subdomains, size=min(d.n_subdomains, 256), replace=False)
    ks = sizes[sample].astype(np.int64)
    need = int(np.searchsorted(np.cumsum(ks ** 2 * 8), 1 << 30)) + 1
    ks = ks[:max(1, min(need, ks.shape[0]))]
    mats = [rng.standard_normal((int(k), int(k))) for k in ks]
    vecs = [rng.standard_normal(int(k)) for k in ks]
    t = time.perf_counter()
    np.linalg.inv(mats[0])
    t_inv = time.perf_counter() - t

    def local_pass():
        for Mi, vi in zip(mats, vecs):
            Mi @ vi

    t_loc = best(local_pass, reps=2)
    scale = float((sizes.astype(np.float64) ** 2).sum() / (ks.astype(np.float64) ** 2).sum())
    t_local = t_loc * scale
    del mats
    # coarse: dense SPD of size n_c
    nc = d.n_subdomains
    B = rng.standard_normal((nc, nc))
    AC = B @ B.T + nc * np.eye(nc)
    cf = sla.cho_factor(AC, lower=True)
    xc = rng.standard_normal(nc)
    t_coarse = best(lambda: sla.cho_solve(cf, xc))
    p = rng.standard_normal(n)
    r = rng.standard_normal(n)

    def vec():
        np.dot(r, p)
        p2 = p * 0.5 + r
        r2 = r - 0.1 * p2
        np.linalg.norm(r2)
        np.dot(p2, r2)
        return p2

    t_vec = best(vec)
    # Hybrid: SpMV (A p) + R0A matvec (~ n entries) + local + coarse + vectors per iteration
    t_iter = t_spmv + t_local + t_coarse + t_vec + t_spmv / 20
    tts = t_iter * iterations + t_spmv + t_coarse
    return {
        "value": round(tts, 4),
        "unit": "s",
        "cores": int(threads),
        "kind": "port",
        "sample": (f"oracle FP64 components on this host: full scipy CSR SpMV ({t_spmv * 1e3:.1f} ms), "
                   f"dense FP64 inverse matvecs over {len(ks)} matrices of the sampled "
                   f"subdomain sizes ({(ks ** 2).sum() * 8 / 2 ** 30:.2f} GiB, out of cache) "
                   f"scaled by sum k_i^2 ({t_local * 1e3:.1f} ms/iter), cho_solve n_c={nc} "
                   f"({t_coarse * 1e3:.2f} ms), vector ops ({t_vec * 1e3:.1f} ms); per-iteration "
                   f"sum x {iterations} iterations; setup excluded (one {int(ks[0])}^2 inverse "
                   f"takes {t_inv * 1e3:.0f} ms); cpu_count={os.cpu_count()}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay captured iteration blocks")
    ap.add_argument("--profile-kernel", action="store_true",
                    help="only launch the local-apply kernel 20 times, then a 1 GiB device "
                         "copy 20 times (calibration for rocprofv3 --pmc byte counters)")
    ap.add_argument("--storage", default="float64", choices=["float64", "float32", "float16", "bfloat16"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import dd_solvers as dds  # noqa: F401
    from dd_solvers import _lib

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    pb = build_problem(args.N, args.seed)
    b = torch.as_tensor(pb.b, device=dev)

    if args.profile_kernel:
        st = getattr(torch, args.storage)
        solver = make_solver(pb, st, dev, scaled=st == torch.float16)
        M = solver.preconditioner
        xv = torch.randn(pb.A.shape[0], dtype=torch.float64, device=dev)
        Y = M.Y if M.Y is not None else torch.empty_like(xv)
        for _ in range(20):
            M._enqueue_local(xv, Y, None)
        torch.cuda.synchronize()
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        L = M.local
        sizes = M._host["sizes"].astype(np.int64)
        log(f"profile-kernel: storage={args.storage} layout={L.layout} "
            f"algorithmic matrix bytes={L.bytes_algorithmic} stored={L.bytes_matrices} "
            f"vectors={int(sizes.sum() * 12 + Y.numel() * 8)}; copy 2 x {src.numel() * 8} bytes")
        print(json.dumps({"profile_kernel": True, "alg_matrix_bytes": L.bytes_algorithmic,
                          "stored_matrix_bytes": L.bytes_matrices}))
        return

    # ---- plain FP64 (the measured line) ------------------------------------------------
    solver = make_solver(pb, torch.float64, dev, use_graph=args.graph)
    x, meta, total, times = time_solves(solver, b, args.steps, args.warmup)
    iters = meta["iterations"]
    tts = total / args.steps
    xh = x.cpu().numpy()
    res = float(np.linalg.norm(pb.b - pb.A @ xh) / np.linalg.norm(pb.b))
    log(f"FP64 TTS {tts * 1e3:.2f} ms (min {min(times) * 1e3:.2f}), {iters} iterations, "
        f"true rel. residual {res:.2e}")

    # ---- dominant kernel: the local inverse apply ---------------------------------------
    M = solver.preconditioner
    L = M.local
    sizes = M._host["sizes"].astype(np.int64)
    xv = torch.randn(pb.A.shape[0], dtype=torch.float64, device=dev)
    Y = M.Y if M.Y is not None else torch.empty_like(xv)
    t_loc = kernel_time_ms(lambda: M._enqueue_local(xv, Y, None))
    # algorithmic bytes of one launch: the matrices (lower triangle for the symmetric
    # layout, whole matrices for "full"), the gathered x (value + index) and the result
    # (sym layout: plus the partial vectors, written by one launch and read by the next)
    yp = getattr(M, "Yp", None)
    alg_bytes = int(L.bytes_algorithmic + sizes.sum() * (8 + 4) + sizes.sum() * 8
                    + (2 * yp.numel() * 8 if yp is not None else 0))
    gbs = alg_bytes / (t_loc * 1e-3) / 1e9
    log(f"local apply: {t_loc * 1e3:.1f} us, {alg_bytes / 1e9:.3f} GB -> {gbs:.0f} GB/s "
        f"({gbs / HBM_PEAK_GBS * 100:.1f}% of HBM peak)")
    # SpMV alone (CSR), for the record
    op = solver.matrix
    q = torch.empty_like(xv)
    t_spmv = kernel_time_ms(lambda: op._enqueue_matvec(xv, q, None, None))
    spmv_bytes = op.bytes_per_matvec
    log(f"SpMV CSR: {t_spmv * 1e3:.1f} us, {spmv_bytes / 1e6:.1f} MB -> "
        f"{spmv_bytes / (t_spmv * 1e-3) / 1e9:.0f} GB/s")
    # whole-solve algorithmic bytes
    per_iter = M.bytes_per_apply + spmv_bytes + 9 * 8 * pb.A.shape[0]
    solve_bytes = per_iter * iters + spmv_bytes
    solve_gbs = solve_bytes / tts / 1e9

    extra = {}
    if not args.no_extra:
        # mixed precision: FP16 local inverses (per-subdomain scaled), FP64 everything else
        del solver
        torch.cuda.empty_cache()
        for name, st, scaled in (("fp32", torch.float32, False), ("fp16", torch.float16, True),
                                 ("bf16", torch.bfloat16, False)):
            try:
                s2 = make_solver(pb, st, dev, scaled=scaled, use_graph=args.graph)
                x2, m2, tot2, t2 = time_solves(s2, b, args.steps, args.warmup)
                e2 = float(np.linalg.norm(x2.cpu().numpy() - xh) / np.linalg.norm(xh))
                r2 = float(np.linalg.norm(pb.b - pb.A @ x2.cpu().numpy()) / np.linalg.norm(pb.b))
                extra[f"mixed_{name}"] = {"tts_s": round(tot2 / args.steps, 5),
                                          "iterations": m2["iterations"],
                                          "rel_err_vs_fp64": e2, "true_rel_residual": r2,
                                          "ratio_vs_fp64": round(tot2 / args.steps / tts, 3)}
                log(f"{name}: TTS {tot2 / args.steps * 1e3:.2f} ms, {m2['iterations']} it, "
                    f"err vs fp64 {e2:.1e}")
                del s2
                torch.cuda.empty_cache()
            except Exception as e:  # report, keep the measured line
                extra[f"mixed_{name}"] = {"error": str(e)[:200]}

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(pb, iters)
        log(f"cpu baseline {cpu['value']} s ({time.perf_counter() - t:.1f}s to measure)")

    out = {
        "metric": "PCG time-to-solution @rtol 1e-8 (s), mixed/batched vs plain FP64, %HBM peak",
        "value": round(tts, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tts * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Q1 Poisson {args.N}^3 (n={pb.A.shape[0]}), lognormal k=exp(N(0,1)), "
                        f"Hybrid Schwarz 8^3-element boxes overlap 1 ({pb.decomposition.n_subdomains} "
                        "subdomains), FP64 inverses, dense FP64 Cholesky coarse, CSR SpMV, rtol 1e-8",
            "n": int(pb.A.shape[0]),
            "nnz": int(pb.A.nnz),
            "iterations": int(iters),
            "true_rel_residual": res,
            "min_tts_s": round(min(times), 6),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("sym_apply_kernel<double,double,double> + sym_combine_kernel<double>"
                       if L.layout == "sym"
                       else "local_apply_kernel<double,double,double,4>"),
            "stored_matrix_bytes": int(L.bytes_matrices),
            "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
            "alg_bytes_per_launch": alg_bytes,
            "avg_launch_us": round(t_loc * 1e3, 2),
        },
        "solve_roofline": {"alg_bytes": int(solve_bytes), "achieved_GBs": round(solve_gbs, 1),
                           "frac": round(solve_gbs / HBM_PEAK_GBS, 4)},
        "spmv": {"us": round(t_spmv * 1e3, 2), "bytes": int(spmv_bytes),
                 "GBs": round(spmv_bytes / (t_spmv * 1e-3) / 1e9, 1)},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
