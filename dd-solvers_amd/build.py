"""Build the HIP kernel library `dd_solvers/libdd_solvers_hip.so` for gfx950 with hipcc,
then the op module `dd_solvers/_C` on top of it.

Each csrc/*.hip translation unit is compiled to an object in parallel and the objects are
linked into one shared library exporting the C ABI of include/dd_solvers_hip.h. No
PyTorch headers are involved there: the library sees only plain pointers and streams.
csrc/torch_op.cpp (host code, PyTorch headers) registers the reference's native op
`torch.ops.dd_solvers_gemv.gemv_strided_batched` over that C ABI and is linked as the
Python extension module `dd_solvers._C`.

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


# host-only translation units of the kernel library (compiled as C++, no device code);
# torch_op.cpp is the op module, built separately against PyTorch
HOST_SOURCES = ("nd_plan.cpp",)


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f in HOST_SOURCES)


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
        base = os.path.basename(src)
        obj = os.path.join(BUILD, os.path.splitext(base)[0] + ".o")
        objs.append(obj)
        if base in HOST_SOURCES:
            cmds.append([cc, "-x", "c++", "-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread",
                         "-I", INCLUDE, "-c", src, "-o", obj])
            continue
        cmds.append([cc, f"--offload-arch={arch}", *HIP_FLAGS,
                     *FILE_FLAGS.get(base, []), "-I", INCLUDE, "-I", CSRC,
                     "-c", src, "-o", obj])
    jobs = jobs or min(8, len(cmds))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={arch}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_op_module()
    return LIB


def build_op_module() -> str:
    """dd_solvers/_C<EXT_SUFFIX>: csrc/torch_op.cpp against PyTorch and the kernel library,
    both found at run time through the rpath (the package directory, torch's lib)."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension

    out = os.path.join(HERE, "dd_solvers", "_C" + sysconfig.get_config_var("EXT_SUFFIX"))
    torch_lib = cpp_extension.library_paths()[0]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    inc = [f"-I{p}" for p in cpp_extension.include_paths()]
    rocm_inc = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "include")
    cmd = [hipcc(), "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-shared",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=_C",
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", *inc, f"-I{rocm_inc}",
           f"-I{sysconfig.get_paths()['include']}", "-I", INCLUDE,
           os.path.join(CSRC, "torch_op.cpp"), "-o", out + ".tmp",
           f"-L{torch_lib}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip",
           "-ltorch_python", f"-L{os.path.dirname(LIB)}", "-ldd_solvers_hip",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{torch_lib}"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.arch, a.jobs, a.verbose))
