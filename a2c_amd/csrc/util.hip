// Plain device copy used to calibrate HBM byte counters (FETCH_SIZE and
// WRITE_SIZE of a copy of known size).
#include "fe_internal.hpp"

namespace a2c {
namespace {
__global__ void k_copy_calib(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace
}  // namespace a2c

extern "C" int a2c_dev_copy(a2c_fe_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  if (!dst || !src || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15))
    return a2c::set_err(c, A2C_EINVAL, "copy needs 16-byte aligned pointers and size");
  const size_t n = bytes / 16;
  if (n) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(a2c::k_copy_calib, dim3(grid), dim3(256), 0, c->stream, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n);
    A2C_HIP(c, hipGetLastError());
  }
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}
