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

// Rewind of a batch's finished systems (step cursor past the last step) to
// their initial records -- state, Jacobians and counters -- so that the next
// round starts them again.  One workgroup per system; transit outputs are
// left as they are (a restarted run overwrites them as it counts again).
__global__ __launch_bounds__(256) void nbg_rewind_kernel(const nbg::SysDesc* __restrict__ descs, int nsys,
                                                         double* __restrict__ state, const double* __restrict__ state0,
                                                         double* __restrict__ jac, const double* __restrict__ jac0,
                                                         int64_t* __restrict__ ints) {
  for (int s = blockIdx.x; s < nsys; s += gridDim.x) {
    const nbg::SysDesc& D = descs[s];
    const bool fin = ints[D.off_int + D.n] > D.nsteps;
    __syncthreads();  // every thread has read the cursor before any clears it
    if (fin) {
      for (int64_t q = threadIdx.x; q < nbg::state_len(D.n); q += blockDim.x) state[D.off_state + q] = state0[D.off_state + q];
      for (int64_t q = threadIdx.x; q < nbg::jac_len(D.n); q += blockDim.x) jac[D.off_jac + q] = jac0[D.off_jac + q];
      for (int64_t q = threadIdx.x; q < nbg::int_len(D.n); q += blockDim.x) ints[D.off_int + q] = 0;
    }
    __syncthreads();
  }
}

extern "C" hipError_t nbg_launch_rewind(const nbg::SysDesc* descs, int nsys, double* state, const double* state0,
                                        double* jac, const double* jac0, int64_t* ints, int cu_count,
                                        hipStream_t stream) {
  if (nsys <= 0) return hipSuccess;
  const int grid = nsys < 8 * cu_count ? nsys : 8 * cu_count;
  hipLaunchKernelGGL(nbg_rewind_kernel, dim3(grid), dim3(256), 0, stream, descs, nsys, state, state0, jac, jac0, ints);
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
