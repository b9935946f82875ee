// nbg_coop.hip -- two-wave (cooperative) persistent launch: a system's
// transit searches run on a second wave while the first integrates (device
// code: nbg_device.h).  A translation unit of its own, so that the register
// allocation of the functions it shares with the one-wave kernels is made
// for each workgroup size separately.
#include "nbg_device.h"

// The cooperative launch (two waves per system) serves body counts up to
// NBG_COOP_MAX: LDS of two n <= 10 layouts plus the ring, ~80 KB, so that a CU
// holds one such workgroup beside two one-wave ones.
#define NBG_COOP_MAX 10
extern "C" int nbg_coop_max_bodies() { return NBG_COOP_MAX; }

// `grid` two-wave workgroups over the groups of `gl` (units: 14 entries).
extern "C" hipError_t nbg_launch_coop(nbg::GroupList gl, nbg::SysBufs bufs, nbg::CoopBufs coop, int64_t chunk, int grid,
                                      unsigned long long* units, hipStream_t stream) {
  hipLaunchKernelGGL(nbg::nbg_run_coop<NBG_COOP_MAX>, dim3(grid), dim3(2 * nbg::WAVE), 0, stream, gl, bufs, coop,
                     chunk, units);
  return hipGetLastError();
}
// HBM doubles per cooperative workgroup for its Jacobian snapshot slots.
extern "C" int64_t nbg_coop_stride() {
  return (int64_t)nbg::COOP_RQ * 2 * (6 * NBG_COOP_MAX) * (7 * NBG_COOP_MAX);
}
extern "C" size_t nbg_lds_bytes_coop() {
  return 2 * nbg::lds_bytes_upto<NBG_COOP_MAX>() + sizeof(nbg::TransitRing<NBG_COOP_MAX>) + 16;
}
