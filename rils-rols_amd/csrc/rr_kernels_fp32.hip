// rr_kernels_fp32.hip -- FP32 precision form of the fitness path.
#include "rr_device.h"
#include "rr_kernels_fp32.h"

namespace rr {

namespace {
__global__ void k_to_fp32(const double* src, float* dst, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (float)src[i];
}
}  // namespace

hipError_t launch_to_fp32(const double* src, float* dst, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_fp32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_fp32(const Fp32Launch&, hipStream_t) { return hipErrorNotSupported; }

}  // namespace rr
