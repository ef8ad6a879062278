"""Extract the reference's published DG results into tests/golden/dg_published.json.

Reads two result files of the reference (text data, not code):
  results/experiment_amgx_pcg.csv    -> per (p, mesh level): DoFs, matrix nnz and ||b||
                                        (the first residual norm of CG from x0 = 0)
  results/experiment_precisions.csv  -> for the FP64 Schwarz runs (preconditioner and
                                        local-inverse precision float64): iteration
                                        count and residual-norm history per
                                        (p, fine, solvers, coarse, Additive/Hybrid);
                                        for every run ("solves"): the solver's
                                        precisions, its iteration count (null when it
                                        raised), its best solve time and its best
                                        solver set-up time on the A100
  results/experiment_parallelism.csv -> the same "solves" records for the p = 1 S11 runs
                                        over subdomain / coarse meshes S8-S11 and C8-C11
                                        ("experiment": "parallelism"; the precision runs
                                        carry "experiment": "precisions")
  results/benchmark_bsrmv.csv        -> per (dimension, mesh, degree) of the BSR SpMV
                                        benchmark: DoFs, DoFs per element, nnz, BSR nnz
                                        ("bsr_shapes")

Run from the repository root in a container that has the reference checked out at
/root/reference:  python tests/golden/make_dg_published.py
"""

import ast
import csv
import json
import os

REF = "/root/reference/results"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dg_published.json")


def main():
    csv.field_size_limit(1 << 30)
    levels = {}
    with open(os.path.join(REF, "experiment_amgx_pcg.csv")) as fh:
        for r in csv.DictReader(fh):
            if r["random rhs"] != "False" or not r["metadata"]:
                continue
            md = ast.literal_eval(r["metadata"])
            if "residual norms" not in md:
                continue
            key = f"p{r['p']}_{r['fine m']}"
            levels.setdefault(key, {"p": int(r["p"]), "level": int(r["fine m"][1:]),
                                    "dofs": int(r["DoFs"]), "nnz": int(r["matrix nnz"]),
                                    "b_norm": md["residual norms"][0]})
    runs = {}
    solves = []
    with open(os.path.join(REF, "experiment_precisions.csv")) as fh:
        columns = next(csv.reader(fh))

    def solve_record(r, experiment):
        s = r["solver"]
        pi = s.split("Schwarz(")[1].split(",")[0].replace("torch.", "")
        inv = s.split("Inv(")[1].split(")")[0].replace("torch.", "") or None
        times = ast.literal_eval(r["solve times"]) if r["solve times"] else []
        setup = ast.literal_eval(r["solver setup times"]) if r.get("solver setup times") else []
        md = ast.literal_eval(r["metadata"]) if r["metadata"] else {}
        return {"experiment": experiment, "p": int(r["p"]), "fine": r["fine m"],
                "solvers": r["solvers m"], "coarse": r["coarse m"],
                "kind": "additive" if "AdditiveSchwarz" in s else "hybrid",
                "precision": pi, "inv": inv,
                "iterations": None if r["exception"] else md.get("iterations"),
                "best_ms": min(times) if times and not r["exception"] else None,
                "setup_ms": min(setup) if setup else None,
                "solver": s}

    with open(os.path.join(REF, "experiment_parallelism.csv")) as fh:
        for r in csv.DictReader(fh):
            if r["solver"].startswith("CG(") and "Schwarz(" in r["solver"]:
                solves.append(solve_record(r, "parallelism"))
    with open(os.path.join(REF, "experiment_precisions.csv")) as fh:
        for r in csv.DictReader(fh):
            s = r["solver"]
            if s.startswith("CG(") and "Schwarz(" in s:
                solves.append(solve_record(r, "precisions"))
            if not s.startswith("CG(") or "torch.float64, Inv()" not in s or r["exception"]:
                continue
            kind = "additive" if "AdditiveSchwarz" in s else "hybrid"
            md = ast.literal_eval(r["metadata"])
            key = f"p{r['p']}_{r['fine m']}_{r['solvers m']}_{r['coarse m']}_{kind}"
            runs[key] = {"p": int(r["p"]), "fine": r["fine m"], "solvers": r["solvers m"],
                         "coarse": r["coarse m"], "kind": kind, "rtol": 1e-9,
                         "iterations": md["iterations"],
                         "residual_norms": md["residual norms"],
                         "solver": s}
    bsr_shapes = {}
    with open(os.path.join(REF, "benchmark_bsrmv.csv")) as fh:
        for r in csv.DictReader(fh):
            key = f"{r['dimension']}_{r['mesh']}_p{r['degree']}"
            bsr_shapes.setdefault(key, {"dim": int(r["dimension"][0]), "mesh": r["mesh"],
                                        "p": int(r["degree"]), "dofs": int(r["DOFs"]),
                                        "dofs_per_element": int(r["DOFs per element"]),
                                        "nnz": int(r["nnz"]), "bsr_nnz": int(r["bsr nnz"])})
    with open(OUT, "w") as fh:
        json.dump({"source": "mgr0dzicki/dd-solvers results/experiment_amgx_pcg.csv, "
                             "results/experiment_precisions.csv (rtol 1e-9, FP64 runs), "
                             "results/experiment_parallelism.csv and results/benchmark_bsrmv.csv",
                   "columns": columns, "levels": levels, "runs": runs, "solves": solves,
                   "bsr_shapes": bsr_shapes}, fh, indent=1)
    print(f"wrote {OUT}: {len(levels)} levels, {len(runs)} runs, {len(solves)} solves")


if __name__ == "__main__":
    main()
