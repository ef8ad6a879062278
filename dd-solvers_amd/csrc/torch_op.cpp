// dd_solvers._C: the native op dd_solvers_gemv::gemv_strided_batched of the reference
// (src/dd_solvers/cpp/gemv.cpp:9-57, registered at gemv.cpp:55-57; the importable module
// of cpp/pyinit.cpp:10-20) as a TORCH_LIBRARY unit over the C ABI dds_gemv_strided_batched
// of libdd_solvers_hip.so. Same schema (inferred from the function: Tensor, Tensor, bool ->
// Tensor), same checks and messages; the launch goes to torch's current stream of the
// tensors' device. Importing the module runs the registration.
#include <Python.h>

#include <ATen/ATen.h>
#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include "dd_solvers_hip.h"

namespace {

int dtype_code(at::ScalarType t) {
  switch (t) {
    case at::ScalarType::Double: return DDS_F64;
    case at::ScalarType::Float: return DDS_F32;
    case at::ScalarType::Half: return DDS_F16;
    case at::ScalarType::BFloat16: return DDS_BF16;
    default: return -1;
  }
}

at::Tensor gemv_strided_batched(const at::Tensor& mat, const at::Tensor& vec,
                                bool use_shared_memory) {
  TORCH_CHECK(mat.dim() == 3 && vec.dim() == 2 && mat.size(0) == vec.size(0) &&
                  mat.size(1) == mat.size(2) && mat.size(1) == vec.size(1),
              "Input tensors must have compatible dimensions: (n, k, k) and (n, k)");
  TORCH_CHECK(mat.stride(1) == 1 && mat.stride(2) == mat.size(1),
              "Matrices must be in column-major format");
  const int64_t n = mat.size(0), k = mat.size(1);
  TORCH_CHECK(n <= 1 || mat.stride(0) == k * k,
              "gemv_strided_batched: the matrices must follow each other (batch stride k*k)");
  TORCH_CHECK(mat.is_cuda() && vec.is_cuda() && mat.device() == vec.device(),
              "gemv_strided_batched: mat and vec must be on the same GPU (HIP kernels only)");
  const int mt = dtype_code(mat.scalar_type()), vt = dtype_code(vec.scalar_type());
  TORCH_CHECK(mt >= 0 && (vt == DDS_F64 || vt == DDS_F32),
              "gemv_strided_batched: unsupported dtypes: mat ", mat.scalar_type(), ", vec ",
              vec.scalar_type());
  TORCH_CHECK(!(mt == DDS_F64 && vt == DDS_F32), "Double-Float mixed precision is not supported.");
  const at::Tensor v = vec.contiguous();
  at::Tensor out = at::empty_like(v);
  if (n == 0 || k == 0) return out;
  const c10::DeviceGuard guard(vec.device());
  hipStream_t stream = c10::hip::getCurrentHIPStream(vec.device().index()).stream();
  const int rc = dds_gemv_strided_batched(mt, vt, mat.const_data_ptr(), v.const_data_ptr(),
                                          out.data_ptr(), n, k, use_shared_memory ? 1 : 0,
                                          stream);
  TORCH_CHECK(rc == 0, "gemv_strided_batched: ", dds_last_error());
  return out;
}

}  // namespace

TORCH_LIBRARY(dd_solvers_gemv, m) { m.def("gemv_strided_batched", &gemv_strided_batched); }

// A module with no methods: importing it loads this library, whose static initialisers
// above register the op.
extern "C" PyObject* PyInit__C(void) {
  static PyModuleDef module = {PyModuleDef_HEAD_INIT, "_C", nullptr, -1, nullptr};
  return PyModule_Create(&module);
}
