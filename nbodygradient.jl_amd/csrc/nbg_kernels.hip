// nbg_kernels.hip -- one-wave persistent launches of the AHL21 step with
// Jacobian propagation and transit search (device code: nbg_device.h), the
// copy kernel used to calibrate HBM byte counters, and their host launchers.
#include "nbg_device.h"

// ---------------------------------------------------------------------------
// Host launcher.  Body counts are served by launch classes: n = 2..9 share
// one kernel (its LDS, the largest of their layouts -- n = 6's, with the
// Jacobian in LDS up to n = 6 and in HBM from n = 7 on -- lets four
// workgroups share a CU, which is also what their register use allows);
// every larger n has its own.
// ---------------------------------------------------------------------------
extern "C" int nbg_class_of(int n) { return n <= 9 ? 9 : n; }

extern "C" hipError_t nbg_launch_group(int cls, nbg::GroupList gl, nbg::SysBufs bufs, int64_t chunk, int grid,
                                       unsigned long long* units, hipStream_t stream) {
  switch (cls) {
#define NBG_CASE(HI)                                                                                          \
  case HI:                                                                                                    \
    hipLaunchKernelGGL(nbg::nbg_run<HI>, dim3(grid), dim3(nbg::WAVE), 0, stream, gl, bufs, chunk, units);    \
    return hipGetLastError();
    NBG_CASE(9)
    NBG_CASE(10)
    NBG_CASE(11)
    NBG_CASE(12)
#undef NBG_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// Plain device copy (16-byte elements, grid-stride): the reference kernel
// for calibrating HBM byte counters and for measuring achieved bandwidth.
__global__ __launch_bounds__(256) void nbg_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                       size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" hipError_t nbg_launch_copy(const void* src, void* dst, size_t bytes, int cu_count, hipStream_t stream) {
  const size_t n = bytes / sizeof(double2);
  hipLaunchKernelGGL(nbg_copy_kernel, dim3(cu_count * 8), dim3(256), 0, stream, (const double2*)src, (double2*)dst, n);
  return hipGetLastError();
}

extern "C" size_t nbg_lds_bytes(int cls) {
  switch (cls) {
    case 9: return nbg::lds_bytes_upto<9>();
    case 10: return nbg::lds_bytes_upto<10>();
    case 11: return nbg::lds_bytes_upto<11>();
    case 12: return nbg::lds_bytes_upto<12>();
    default: return 0;
  }
}
