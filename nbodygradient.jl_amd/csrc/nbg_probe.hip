// nbg_probe.hip -- diagnostic only: the Jacobian row update of one Kepler
// pair (ahl21.jl:31-42: rows of bodies i and j of J, 12 x 7n, replaced by
// J + jac_ij * J with Kahan compensation, comp_sum_matrix!) timed in two
// forms on the same synthetic data, so that the choice between the VALU and
// the FP64 matrix cores rests on a measurement:
//   mode 0 (VALU): lane c owns column c and forms its 12 outputs with 14
//     fused multiply-adds each (168 v_fma_f64 per lane), as the persistent
//     kernel's phase C does;
//   mode 1 (MFMA): the 12 x 14 coefficient block (zero-padded to 16 x 16)
//     times the 14 x 64 operand (16 x 64) as 4 column tiles x 4 k-steps of
//     v_mfma_f64_16x16x4f64, then the same Kahan update per entry.
// The operand is n = 9's: 7n = 63 columns (a 64th, zero, pads the tile).
// Each workgroup (one wave) keeps its J, Jerr and two coefficient blocks in
// LDS and repeats the update `reps` times, alternating the blocks (as the
// kernel reads each pair's coefficients from its slot); no global memory is
// touched inside the loop.  The two forms round differently (the matrix core
// sums four products per instruction in its own order), so their results
// agree to rounding, not bit for bit.
//
// Also here: libm_kernel, the device build of nbg_libm.h evaluated on an
// array, so that a test can compare it bit for bit with the host build.
#include <hip/hip_runtime.h>

#include "nbg_libm.h"

namespace {

constexpr int ROWS = 12, COLS = 64, K = 14, N7 = 63;
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline void kahan_add(double& s, double& e, double add) {
  e += add;
  const double t = s + e;
  e += s - t;
  s = t;
}

// Synthetic data: coefficients of a few 1e-4 (the update then stays bounded
// over thousands of repetitions), J entries of order one.
__device__ inline double coef(int q, int k, int set) { return 2e-5 * (double)((7 * q + 3 * k + set) % 11 - 5); }
__device__ inline double jinit(int r, int c, int w) { return 1.0 + 0.01 * (double)((13 * r + 7 * c + w) % 17); }

template <int MODE>
__global__ __launch_bounds__(64) void probe_kernel(int reps, double* __restrict__ out) {
  __shared__ double J[ROWS * COLS], E[ROWS * COLS], C[2 * 16 * 16];
  const int l = threadIdx.x;
  for (int q = l; q < ROWS * COLS; q += 64) {
    const int r = q / COLS, c = q % COLS;
    J[q] = c < N7 ? jinit(r, c, blockIdx.x) : 0.0;
    E[q] = 0.0;
  }
  for (int q = l; q < 512; q += 64) {
    const int r = (q % 256) / 16, k = q % 16;
    C[q] = (r < ROWS && k < K) ? coef(r, k, q / 256) : 0.0;
  }
  __syncthreads();
  if constexpr (MODE == 0) {
    const int c = l;
    const double a6 = (c == 6) ? 1.0 : 0.0, a13 = (c == 13) ? 1.0 : 0.0;
#pragma unroll 1
    for (int it = 0; it < reps; ++it) {
      const double* cf = C + 256 * (it & 1);  // the same for every lane
      if (c < N7) {
        double a[ROWS], e[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          a[r] = J[r * COLS + c];
          e[r] = E[r * COLS + c];
        }
        // operand b = (a_i[0..5], a6, a_j[0..5], a13)
        double b[K];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          b[k] = a[k];
          b[7 + k] = a[6 + k];
        }
        b[6] = a6;
        b[13] = a13;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          double o = 0.0;
#pragma unroll
          for (int k = 0; k < K; ++k) o = __builtin_fma(cf[q * 16 + k], b[k], o);
          double s = a[q], se = e[q];
          kahan_add(s, se, o);
          J[q * COLS + c] = s;
          E[q * COLS + c] = se;
        }
      }
    }
  } else {
#pragma unroll 1
    for (int it = 0; it < reps; ++it) {
      // A fragments (coefficients): lane l holds A[l % 16][4 s + l / 16]
      const double* cf = C + 256 * (it & 1);
      double af[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) af[s] = cf[(l % 16) * 16 + 4 * s + l / 16];
      v4d acc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = 4 * s + l / 16;  // B fragment: lane l holds B[4 s + l / 16][16 t + l % 16]
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int col = 16 * t + l % 16;
          double bv;
          if (k < 6) bv = J[k * COLS + col];
          else if (k == 6) bv = (col == 6) ? 1.0 : 0.0;
          else if (k < 13) bv = J[(k - 1) * COLS + col];
          else if (k == 13) bv = (col == 13) ? 1.0 : 0.0;
          else bv = 0.0;
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], bv, acc[t], 0, 0, 0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);  // every lane has read its operands
      __builtin_amdgcn_wave_barrier();
      // D fragments: register r of lane l holds D[4 r + l / 16][16 t + l % 16]
      // (measured on gfx950 with one-hot operands)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 4 * r + l / 16, col = 16 * t + l % 16;
          if (row < ROWS && col < N7) {
            double s = J[row * COLS + col], se = E[row * COLS + col];
            kahan_add(s, se, acc[t][r]);
            J[row * COLS + col] = s;
            E[row * COLS + col] = se;
          }
        }
      __builtin_amdgcn_s_waitcnt(0xc07f);  // the next repetition reads other lanes' entries
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  if (blockIdx.x == 0)
    for (int q = l; q < ROWS * N7; q += 64) out[q] = J[(q / N7) * COLS + q % N7];
}

// out[i] = f(x[i]) for fn 0 sin, 1 cos, 2 exp, 3 sinh, 4 cosh, 5 cbrt (one
// element per lane, grid-stride).
__global__ __launch_bounds__(256) void libm_kernel(int fn, int64_t count, const double* __restrict__ x,
                                                   double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    double y;
    switch (fn) {
      case 0: y = nbg::sin(v); break;
      case 1: y = nbg::cos(v); break;
      case 2: y = nbg::exp(v); break;
      case 3: y = nbg::sinh(v); break;
      case 4: y = nbg::cosh(v); break;
      default: y = nbg::cbrt(v); break;
    }
    out[i] = y;
  }
}

}  // namespace

extern "C" hipError_t nbg_launch_libm(int fn, int64_t count, const double* x, double* out, hipStream_t stream) {
  const int64_t blocks = (count + 255) / 256;
  const int grid = (int)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096);
  hipLaunchKernelGGL(libm_kernel, dim3(grid), dim3(256), 0, stream, fn, count, x, out);
  return hipGetLastError();
}

// Runs the probe: `grid` one-wave workgroups, `reps` updates each; *ms is the
// kernel's time (HIP events), out (12 x 63, row-major) workgroup 0's final J.
extern "C" hipError_t nbg_launch_probe(int mode, int reps, int grid, double* out, hipStream_t stream) {
  if (mode == 0) hipLaunchKernelGGL(probe_kernel<0>, dim3(grid), dim3(64), 0, stream, reps, out);
  else hipLaunchKernelGGL(probe_kernel<1>, dim3(grid), dim3(64), 0, stream, reps, out);
  return hipGetLastError();
}
