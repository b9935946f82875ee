// nbg_coop.hip -- two-wave (cooperative) persistent launch: a system's
// transit searches run on a second wave while the first integrates (device
// code: nbg_device.h).  A translation unit of its own, so that the register
// allocation of the functions it shares with the one-wave kernels is made
// for each workgroup size separately.
#include "nbg_device.h"

// Cooperative launch (two waves per system) for class 9: `grid` workgroups.
extern "C" hipError_t nbg_launch_coop(int cls, nbg::GroupList gl, nbg::SysBufs bufs, nbg::CoopBufs coop, int grid,
                                      unsigned long long* units, hipStream_t stream) {
  if (cls != 9) return hipErrorInvalidValue;
  hipLaunchKernelGGL(nbg::nbg_run_coop<9>, dim3(grid), dim3(2 * nbg::WAVE), 0, stream, gl, bufs, coop, units);
  return hipGetLastError();
}
// HBM doubles per cooperative workgroup for its Jacobian snapshot slots.
extern "C" int64_t nbg_coop_stride(int cls) { return (int64_t)nbg::COOP_RQ * 2 * (6 * cls) * (7 * cls); }
extern "C" size_t nbg_lds_bytes_coop(int cls) {
  return cls == 9 ? 2 * nbg::lds_bytes_upto<9>() + sizeof(nbg::TransitRing<9>) + 16 : 0;
}
