"""Build the HIP kernel library `dd_solvers/libdd_solvers_hip.so` for gfx950 with hipcc.

Each csrc/*.hip translation unit is compiled to an object in parallel and the objects are
linked into one shared library exporting the C ABI of include/dd_solvers_hip.h. No
PyTorch headers are involved: the library sees only plain pointers and streams.

    python dd-solvers_amd/build.py [--jobs N] [--arch gfx950]
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "dd_solvers", "libdd_solvers_hip.so")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic (v_pk_*_f32).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


# Per-file extra flags. local_sym.hip: keep MFMA accumulators in VGPRs (the matrix-core
# symmetric apply reads and resets them between tiles; in AGPRs each would cost copies).
FILE_FLAGS = {"local_sym.hip": ["-mllvm", "--amdgpu-mfma-vgpr-form"]}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    # The host pass does not know the packed-fp32 feature and says so; drop that noise.
    noise = "is not a recognized feature for this target"
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if msg.strip():
        print(msg, file=sys.stderr)


def build(arch: str = "gfx950", jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    srcs = sources()
    objs = []
    cmds = []
    for src in srcs:
        obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
        objs.append(obj)
        cmds.append([cc, f"--offload-arch={arch}", *HIP_FLAGS,
                     *FILE_FLAGS.get(os.path.basename(src), []), "-I", INCLUDE, "-I", CSRC,
                     "-c", src, "-o", obj])
    jobs = jobs or min(8, len(cmds))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={arch}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.arch, a.jobs, a.verbose))
