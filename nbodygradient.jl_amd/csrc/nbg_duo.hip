// nbg_duo.hip -- one-wave persistent launches with two systems per wave
// (nbg_run_duo; device code: nbg_device.h).  A translation unit of its own,
// so that the one-system kernels' register allocation does not depend on it.
#include "nbg_device.h"

extern "C" hipError_t nbg_launch_group_duo(int cls, nbg::GroupList gl, nbg::SysBufs bufs, int64_t chunk, int grid,
                                           unsigned long long* units, hipStream_t stream) {
  switch (cls) {
#define NBG_CASE(HI)                                                                                           \
  case HI:                                                                                                     \
    hipLaunchKernelGGL(nbg::nbg_run_duo<HI>, dim3(grid), dim3(nbg::WAVE), 0, stream, gl, bufs, chunk, units);  \
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

namespace {
template <int N, int HI> constexpr bool pairs_in() {
  if constexpr (N > HI) return false;
  else return nbg::duo_fits<N, HI>();
}
template <int HI> bool pairs_class(int n) {
  switch (n) {
    case 2: return pairs_in<2, HI>();
    case 3: return pairs_in<3, HI>();
    case 4: return pairs_in<4, HI>();
    case 5: return pairs_in<5, HI>();
    case 6: return pairs_in<6, HI>();
    case 7: return pairs_in<7, HI>();
    case 8: return pairs_in<8, HI>();
    case 9: return pairs_in<9, HI>();
    case 10: return pairs_in<10, HI>();
    case 11: return pairs_in<11, HI>();
    case 12: return pairs_in<12, HI>();
    default: return false;
  }
}
}  // namespace

// Whether systems of n bodies run two per wave in a launch of class cls.
extern "C" int nbg_duo_pairs(int cls, int n) {
  switch (cls) {
    case 9: return pairs_class<9>(n);
    case 10: return pairs_class<10>(n);
    case 11: return pairs_class<11>(n);
    case 12: return pairs_class<12>(n);
    default: return 0;
  }
}
