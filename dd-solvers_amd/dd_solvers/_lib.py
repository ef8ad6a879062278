"""ctypes binding of the C ABI in include/dd_solvers_hip.h (libdd_solvers_hip.so).

The library is built in-tree by dd-solvers_amd/build.py and loaded after torch, so it
shares torch's HIP runtime (both resolve the same libamdhip64.so.7). Importing this module
fails with ImportError when the library is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library links against)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdd_solvers_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP kernels first "
        "(python dd-solvers_amd/build.py, or __graft_entry__.build())")

_lib = ctypes.CDLL(LIB_PATH)

# dtype codes (dds_dtype)
F64, F32, F16, BF16, I32, I64 = 0, 1, 2, 3, 4, 5
# finalize ops
FIN_CHECK, FIN_RHO, FIN_ALPHA, FIN_NEXT = 0, 1, 2, 3
MAX_PARTIALS = 16384
NRHS = 16  # right-hand sides per block in the multi-RHS kernels
ABI_VERSION = 4

_TORCH_TO_DDS = {
    torch.float64: F64,
    torch.float32: F32,
    torch.float16: F16,
    torch.bfloat16: BF16,
    torch.int32: I32,
    torch.int64: I64,
}


def dtype_code(dt: torch.dtype) -> int:
    try:
        return _TORCH_TO_DDS[dt]
    except KeyError:
        raise ValueError(f"unsupported dtype {dt}") from None


_p = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_ip = ctypes.POINTER(ctypes.c_int)

# name -> argtypes (restype int unless listed otherwise)
_SIGNATURES = {
    "dds_abi_version": [],
    "dds_gemv_strided_batched": [_i, _i, _p, _p, _p, _i64, _i64, _i, _p],
    "dds_bsrmv": [_i, _i, _i64, _i64, _p, _p, _p, _p, _p, _p, _ip, _p, _p],
    "dds_csrmv": [_i, _i, _i64, _p, _p, _p, _p, _p, _i, _p, _ip, _p, _p],
    "dds_sellmv": [_i, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _ip, _p, _p],
    "dds_local_vec_width": [_i],
    "dds_local_tile_rows": [_i],
    "dds_local_apply": [_i, _i, _i, _i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _i, _p, _p],
    "dds_local_apply_batched": [_i, _i, _p, _p, _p, _i64, _i64, _i, _p, _p, _p],
    "dds_local_apply_sym": [_i, _i, _i, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "dds_local_apply_sym_mfma": [_i, _i, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p,
                                 _i64, _p, _p, _p, _p, _p, _p, _p],
    "dds_local_apply_sym_mfma_fused": [_i, _i, _i, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p,
                                       _p, _p, _p, _p, _p, _p],
    "dds_sym_mfma_pack": [_i, _i, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "dds_sym_mfma_apply_packed": [_i, _i, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p,
                                  _p, _p, _p, _p],
    "dds_sym_combine": [_i, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p],
    "dds_local_apply_sym_mfma_noalloc": [_i, _i, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p,
                                         _p, _p, _i64, _p, _p, _p, _p, _p],
    "dds_overlap_sum": [_i, _i64, _p, _p, _p, _i, _p, _p, _d, _p, _p, _p, _ip, _p, _p],
    "dds_restrict": [_i, _i64, _p, _p, _p, _i, _p, _p],
    "dds_chol_factor": [_i, _i64, _p, _i64, _p, _p],
    "dds_chol_solve": [_i, _i64, _p, _i64, _p, _p, _p, _p],
    "dds_chol_inverse": [_i, _i64, _p, _i64, _p, _i64, _p],
    "dds_dense_symv": [_i, _i64, _p, _i64, _p, _p, _p, _p],
    "dds_nd_forward": [_i, _i, _i64, _p, _p, _p, _p, _p, _p, _p, _i64, _p, _p],
    "dds_nd_backward": [_i, _i, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _p, _p],
    "dds_nd_top_assemble": [_i, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p],
    "dds_nd_top_push": [_i, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p],
    "dds_nd_plan_build": [_i64, _p, _p, _i64, _i64, _p],
    "dds_nd_factor": [_i, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                      _p, _p, _p, _p, _p, _p, _p],
    "dds_pcg_residual": [_i64, _p, _p, _p, _p, _ip, _p, _p],
    "dds_pcg_update_p": [_i64, _p, _p, _p, _p],
    "dds_pcg_update_xr": [_i64, _p, _p, _p, _p, _p, _ip, _p, _p],
    "dds_dot_partials": [_i64, _p, _p, _p, _ip, _p, _p],
    "dds_pcg_finalize": [_i, _p, _i, _p, _p, _p, _p, _p],
    "dds_sum_partials": [_p, _i, _p, _p],
    "dds_cast": [_i, _i, _i64, _p, _p, _p, _p],
    "dds_q1_element_grad": [_i64, _p, _p, _p, _p, _p, _p],
    "dds_spd_inverse_batched": [_i, _i64, _p, _i64, _i64, _p, _p, _p],
    "dds_inverse_batched": [_i, _i64, _p, _i64, _i64, _p, _p, _p],
    "dds_segment_sum": [_i, _i, _i64, _p, _p, _p, _p],
    "dds_stream": [_i, _i64, _p, _p, _d, _p, _i64, _p],
    "dds_pcg_residual_m": [_i64, _p, _p, _p, _p, _ip, _p],
    "dds_pcg_update_p_m": [_i64, _p, _p, _p, _p],
    "dds_pcg_update_xr_m": [_i64, _p, _p, _p, _p, _p, _ip, _p, _p],
    "dds_dot_m": [_i64, _p, _p, _p, _ip, _p, _p],
    "dds_pcg_finalize_m": [_i, _p, _i, _p, _p, _p, _p, _p],
    "dds_sellmm": [_i64, _p, _p, _p, _p, _p, _p, _ip, _p, _p],
    "dds_csrmm_rows": [_i64, _p, _p, _p, _p, _p, _p, _p],
    "dds_dense_symm_m": [_i64, _p, _p, _p, _p, _p],
    "dds_restrict_m": [_i64, _p, _p, _p, _p, _p],
    "dds_overlap_sum_m": [_i64, _p, _p, _p, _p, _p, _d, _p, _p, _p, _ip, _p, _p],
    "dds_mfma_pack_x": [_i, _i64, _p, _p, _p, _p, _p, _p, _p, _p],
    "dds_mfma_apply": [_i, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p],
    "dds_mfma_apply_sub": [_i, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p],
    "dds_sym_mfma_pack16": [_i, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "dds_sym_mfma_apply16": [_i, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
}

_lib.dds_last_error.restype = ctypes.c_char_p
_lib.dds_last_error.argtypes = []
for _name, _args in _SIGNATURES.items():
    _fn = getattr(_lib, _name)
    _fn.argtypes = _args
    _fn.restype = ctypes.c_int
# the symbolic-plan accessors return a length and nothing
_lib.dds_nd_plan_field.argtypes = [_p, _i, _p]
_lib.dds_nd_plan_field.restype = ctypes.c_int64
_lib.dds_nd_plan_free.argtypes = [_p]
_lib.dds_nd_plan_free.restype = None

EXPORTED = ["dds_last_error", "dds_nd_plan_field", "dds_nd_plan_free", *_SIGNATURES]

if _lib.dds_abi_version() != ABI_VERSION:
    raise ImportError(f"{LIB_PATH}: ABI version {_lib.dds_abi_version()} != {ABI_VERSION}; rebuild")


class HipError(RuntimeError):
    """A dds_* entry point returned a non-zero status."""


def call(name: str, *args) -> None:
    status = getattr(_lib, name)(*args)
    if status != 0:
        msg = _lib.dds_last_error().decode(errors="replace")
        raise HipError(f"{name}: {msg} (status {status})")


def call_partials(name: str, *args) -> int:
    """Call an entry point whose second-to-last-but-two argument is `int* n_partials`."""
    n = ctypes.c_int(0)
    args = list(args)
    idx = args.index(PARTIALS_OUT)
    args[idx] = ctypes.byref(n)
    call(name, *args)
    return n.value


class _PartialsOut:
    """Placeholder marking where call_partials puts its `int* n_partials`."""


PARTIALS_OUT = _PartialsOut()


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream() -> int:
    """Handle of torch's current HIP stream."""
    return torch.cuda.current_stream().cuda_stream


def require_device(*tensors) -> None:
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError(
                "dd_solvers runs its kernels on the GPU: tensors must live on a HIP device "
                f"(got a tensor on {t.device})")
