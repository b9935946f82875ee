// nbg_device.h -- device code of the AHL21 step with Jacobian propagation
// and transit search on gfx950 (included by the translation units that
// instantiate its kernels: nbg_kernels.hip for the one-wave launches,
// nbg_coop.hip for the two-wave ones, so that neither's register allocation
// of the shared functions depends on the other's workgroup size).
//
// One wavefront (64 lanes) integrates one system; a persistent grid pulls
// systems from a work queue.  Inside a wavefront:
//  * the state (x, v, Kahan errors, masses, dq/dt) lives in LDS; the scalar
//    pair work (universal-variable Kepler solve and its analytic Jacobian)
//    runs one lane per pair of a level of disjoint pairs;
//  * the 7n x 7n Jacobian without its mass rows, which the AHL21 map never
//    changes (they stay unit rows): J and its compensation error are 6n x 7n,
//    row-major, in LDS up to n = 6 and in a per-system HBM work area from
//    n = 7 on; lane c owns column c for every update (pair maps, drifts, the
//    kick and force-gradient products, Kahan sums);
//  * the kick matrix K, the force-gradient matrix Phi and d(acceleration)/dq
//    have non-zeros only in velocity rows x (position, mass) columns; they are
//    stored as 3n x 4n blocks and built with one lane per column.
// Operation order follows NbodyGradient.jl src/integrator/ahl21/ahl21.jl:5-95
// (mul! as a k-ascending fused multiply-add accumulation), so results match
// the CPU restatement bit for bit.  Build with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include "nbg_pair.h"
#include "nbg_types.h"

namespace nbg {

constexpr int WAVE = 64;

template <int NB> struct Shape {
  static constexpr int N3 = 3 * NB, N4 = 4 * NB, N6 = 6 * NB, N7 = 7 * NB;
};

// One pair of a sweep chunk: inputs and results of the universal-variable
// Kepler drift (phase A) and its analytic 6x8 Jacobian (phase B).
struct PairSlot {
  double x0[3], v0[3];
  KParams P;
  double delxv[6];
  double mi, mj, mijinv, m_i, m_j;
  int valid, i, j, pad;
  double kep[48], jm[6];
  double mc[24];  // jac_ij's mass columns 6 (rows 0-5, 7-12) and 13 (same rows)
};

template <int NB, bool G = (NB >= JGLOBAL_MIN)> struct Smem {
  static constexpr int N3 = 3 * NB, N4 = 4 * NB, N6 = 6 * NB, N7 = 7 * NB;
  static constexpr int JOFF = N6 * N7;
  // jac_step without mass rows (row 6b+k <-> Julia row 7b+k), then jac_error
  // in the same layout at JOFF.  2*N6*N7 >= N7*N7, so jac_init fits here too.
  // From JGLOBAL_MIN bodies on, these live in an HBM work area (jg) and the
  // LDS keeps everything else: four such systems fit a CU instead of two.
  // Every access to them is made by the lane that owns the column.
  static constexpr bool JG = G;
  double JJ[JG ? 1 : 2 * N6 * N7];
  double* jg;
  static constexpr int NP = NB * (NB - 1) / 2;  // pairs per sweep
  // Kepler pairs (i, j) of the forward / backward sweep grouped by level:
  // a pair's level is one more than the latest earlier pair sharing a body,
  // so the pairs of a level are disjoint; lvw holds the level widths.
  uint8_t plist[2][2 * (NP > 0 ? NP : 1)];
  uint8_t lvw[2][NP > 0 ? NP : 1];
  int nlv[2];
  // Source of jac_ij for a pair whose masses sum to zero (the reference leaves
  // jac_ij from the pair before it, ahl21.jl:712-713): the last pair with mass
  // when it lies in an earlier chunk, or zero at the start of a step.
  PairSlot jlast;
  int seq;  // the system has such a pair: sweeps keep the sequential order
  int rowupd;  // SysDesc::rowupd: the form of the Jacobian row update (pair_cols)
  double dqij[14];       // dqdt_ij
  double x[N3], v[N3], xe[N3], ve[N3], m[NB];
  double px[N3], pv[N3], pxe[N3], pve[N3];  // s_prior
  double a[N3];
  double dqdt[N7], pdqdt[N7], dqdt_phi[N7], dqdt_kick[N7];
  double gsave[NB];
  double t, pt;
  int64_t count[NB];
  uint8_t pair[NB * NB];
};

// The work area of a step's stages (LDS).  Lifetimes are disjoint: jac_phi
// and dadq live in the force-gradient stage (dadq also holds jac_kick around
// the kicks), the pair slots in the sweeps; each stage builds what it reads.
// One system's layout (Solo) keeps it beside the system's record; a pair of
// systems run by one wave (Duo) shares one, the slots split between them.
template <int NB> struct Swork {
  static constexpr int N3 = 3 * NB, N4 = 4 * NB;
  static constexpr int NP = NB * (NB - 1) / 2;
  static constexpr int PQ = 13;  // per-pair factors of phisalpha!, see phisalpha_cols
  // pair slots share the space of the force-gradient arrays
  static constexpr int CH_FIT = (int)(((2 * N3 * N4 + (NP > 0 ? NP : 1) * PQ) * sizeof(double)) / sizeof(PairSlot));
  static constexpr int CH_MIN = CH_FIT < 4 ? 4 : CH_FIT;
  static constexpr int CH = NP < CH_MIN ? NP : (CH_MIN > 64 ? 64 : CH_MIN);  // pairs per sweep chunk
  union {
    struct {
      double phi[N3 * N4];   // jac_phi block: row 3b+k (velocity k of body b), col 3d+p (x_p of d) or N3+d (m_d)
      double dadq[N3 * N4];  // d a / d q: row 3b+k, same columns
      double pq[(NP > 0 ? NP : 1) * PQ];  // per pair (lexicographic index)
    } f;
    PairSlot slot[CH];
  };
};

// LDS of one system run by one wave: its record and its work area.
template <int NB> struct Solo {
  Smem<NB> s;
  Swork<NB> w;
};

// The Jacobian store of a system: an LDS array, or (from JGLOBAL_MIN bodies
// on) the HBM work area, addressed as global memory from a wave-uniform
// base so that accesses are global loads/stores with scalar bases rather
// than flat ones.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ gdouble* uniform_global(const double* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return (gdouble*)((((uint64_t)hi) << 32) | lo);
}
template <int NB, bool G> __device__ __forceinline__ auto jj(Smem<NB, G>& S) {
  if constexpr (G) return uniform_global(S.jg);
  else return static_cast<double*>(S.JJ);
}
template <int NB, bool G> __device__ __forceinline__ auto jj(const Smem<NB, G>& S) {
  if constexpr (G) return (const gdouble*)uniform_global(S.jg);
  else return static_cast<const double*>(S.JJ);
}

// Synchronisation of the lanes of one wave (a system's lanes are one wave):
// LDS (and scalar) accesses complete, no reordering across.  A one-wave
// workgroup's __syncthreads() lowers to exactly this; in a cooperative
// workgroup the two waves run different code and sync only with themselves.
__device__ inline void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0), vmcnt/expcnt unchanged (gfx9 encoding)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class RS, class RE> __device__ inline void kahan(RS& s, RE& e, double add) {
  // comp_sum: utils.jl:16-23
  e += add;
  const double t = s + e;
  e = (s - t) + e;
  s = t;
}

// jac_ij (ahl21.jl:734-758) is a function of the pair's slot: rows 0-5 and
// 7-12 (body i, body j), column k.  jac_ij starts at zero and accumulates
// +-m*kep (ahl21.jl:735-742); the mass columns 6 and 13 are set from jac_mass,
// delxv and kep's k-derivative column; rows 6 and 13 stay zero.
__device__ inline double jij_coef(const PairSlot& sl, bool rowi, int rk, int k) {
  const double mi = sl.mi, mj = sl.mj, mijinv = sl.mijinv;
  if (k == 6)
    return rowi ? sl.jm[rk] * sl.m_j : (((-mj) * sl.delxv[rk]) * mijinv) - ((NBG_GNEWT * mi) * sl.kep[rk + 36]);
  if (k == 13)
    return rowi ? ((mi * sl.delxv[rk]) * mijinv) + ((NBG_GNEWT * mj) * sl.kep[rk + 36]) : (-sl.jm[rk]) * sl.m_i;
  const int ck = k < 7 ? k : k - 7;
  const double kv = sl.kep[rk + 6 * ck];
  if (rowi == (k < 7)) return 0.0 + (rowi ? mj : mi) * kv;
  return 0.0 - (rowi ? mj : mi) * kv;
}

// ---------------------------------------------------------------------------
// The pair Kepler drift (jac_delxv_gamma!, ahl21.jl:766-890) and its analytic
// Jacobian (compute_jacobian_gamma!, ahl21.jl:896-1139) as single out-of-line
// copies shared by every body count's kernel: inlined, each would be
// duplicated per kernel and per call site, and the hot loop of a step would
// no longer fit the instruction cache.  Inputs and outputs pass through the
// pair's slot in LDS; the arithmetic runs in registers.
// ---------------------------------------------------------------------------
static __device__ __attribute__((noinline)) void pair_drift(PairSlot* sl, double gm, double hh, bool drift_first) {
  drift_first = __builtin_amdgcn_readfirstlane((int)drift_first) != 0;  // the same for the whole sweep
  double x0[3], v0[3], delxv[6];
  for (int k = 0; k < 3; ++k) {
    x0[k] = sl->x0[k];
    v0[k] = sl->v0[k];
  }
  const KParams P = jac_delxv_gamma(x0, v0, delxv, gm, hh, drift_first);
  sl->P = P;
  for (int k = 0; k < 6; ++k) sl->delxv[k] = delxv[k];
}

static __device__ __attribute__((noinline)) void pair_jacobian(PairSlot* sl, bool drift_first) {
  drift_first = __builtin_amdgcn_readfirstlane((int)drift_first) != 0;  // the same for the whole sweep
  double x0[3], v0[3], kep[48], jm[6];
  for (int k = 0; k < 3; ++k) {
    x0[k] = sl->x0[k];
    v0[k] = sl->v0[k];
  }
  for (int q = 0; q < 48; ++q) kep[q] = 0.0;
  for (int q = 0; q < 6; ++q) jm[q] = 0.0;
  const KParams P = sl->P;
  const double mi = sl->mi, mj = sl->mj, mijinv = sl->mijinv, m_i = sl->m_i, m_j = sl->m_j;
  double dl[6];
  for (int q = 0; q < 6; ++q) dl[q] = sl->delxv[q];
  compute_jacobian_gamma(P, x0, v0, kep, jm, drift_first);
  // jac_ij's mass columns 6 and 13 (the k == 6, 13 cases of jij_coef),
  // formed from registers: reading them back from the slot after the
  // stores below would cost an LDS round trip each
  double mc[24];
  for (int rk = 0; rk < 6; ++rk) {
    mc[rk] = jm[rk] * m_j;
    mc[6 + rk] = (((-mj) * dl[rk]) * mijinv) - ((NBG_GNEWT * mi) * kep[rk + 36]);
    mc[12 + rk] = ((mi * dl[rk]) * mijinv) + ((NBG_GNEWT * mj) * kep[rk + 36]);
    mc[18 + rk] = (-jm[rk]) * m_i;
  }
  for (int q = 0; q < 48; ++q) sl->kep[q] = kep[q];
  for (int q = 0; q < 6; ++q) sl->jm[q] = jm[q];
  for (int q = 0; q < 24; ++q) sl->mc[q] = mc[q];
}

// ---------------------------------------------------------------------------
// Lane <-> column helpers for the 3n x 4n blocks: column cc is (d, p) with
// p in 0..2 for positions and p = 3 for the mass.
// ---------------------------------------------------------------------------
template <int NB> __device__ inline void col_dp(int cc, int& d, int& p) {
  constexpr int N3 = 3 * NB;
  if (cc < N3) { d = cc / 3; p = cc - 3 * d; }
  else { d = cc - N3; p = 3; }
}

// ---------------------------------------------------------------------------
// kickfast! with Jacobian (ahl21.jl:337-386): impulses for kicked pairs.
// v (Kahan) and dqdt_kick are uniform (lane 0 stores); jac_kick is built with
// one lane per column into W.f.dadq.
// ---------------------------------------------------------------------------
template <int NB, bool G> __device__ void kick_build(Smem<NB, G>& S, Swork<NB>& W, double h, bool needQ, int lane) {
  constexpr int N3 = 3 * NB, N4 = 4 * NB;
  double* K = W.f.dadq;
  if (lane < N4)
    for (int r = 0; r < N3; ++r) K[r * N4 + lane] = 0.0;
  int myd = -1, myp = -1;
  if (lane < N4) col_dp<NB>(lane, myd, myp);
  for (int i = 0; i < NB - 1; ++i) {
    for (int j = i + 1; j < NB; ++j) {
      if (!S.pair[i + NB * j]) continue;
      double rij[3];
      for (int k = 0; k < 3; ++k) rij[k] = S.x[3 * i + k] - S.x[3 * j + k];
      const double r2inv = 1.0 / dot3(rij);
      const double r3inv = r2inv * __builtin_sqrt(r2inv);
      const double fac2 = (h * NBG_GNEWT) * r3inv;
      const double mi = S.m[i], mj = S.m[j];
      for (int k = 0; k < 3; ++k) {
        const double fac = fac2 * rij[k];
        if (lane == 0) {
          kahan(S.v[3 * i + k], S.ve[3 * i + k], (-mj) * fac);
          kahan(S.v[3 * j + k], S.ve[3 * j + k], mi * fac);
          if (needQ) {
            S.dqdt_kick[7 * i + 3 + k] -= (mj * fac) / h;
            S.dqdt_kick[7 * j + 3 + k] += (mi * fac) / h;
          }
        }
        if (lane < N4) {
          const int ri = 3 * i + k, rj = 3 * j + k;
          if (myp == 3 && myd == j) K[ri * N4 + lane] -= fac;
          if (myp == 3 && myd == i) K[rj * N4 + lane] += fac;
          const double facx = fac * (3.0 * r2inv);
          if (myp < 3 && myd == i) {
            K[ri * N4 + lane] += (facx * mj) * rij[myp];
            K[rj * N4 + lane] -= (facx * mi) * rij[myp];
          }
          if (myp < 3 && myd == j) {
            K[ri * N4 + lane] -= (facx * mj) * rij[myp];
            K[rj * N4 + lane] += (facx * mi) * rij[myp];
          }
          if (myp == k && myd == i) {
            K[ri * N4 + lane] -= fac2 * mj;
            K[rj * N4 + lane] += fac2 * mi;
          }
          if (myp == k && myd == j) {
            K[ri * N4 + lane] += fac2 * mj;
            K[rj * N4 + lane] -= fac2 * mi;
          }
        }
      }
      wsync();
    }
  }
}

// phisalpha! (ahl21.jl:558-700, alpha = 2) for one system, on the lanes
// whose column index within the system is col = 0, 1, ..., its pair-factor
// loops striding by STEP lanes: one system on a whole wave (STEP = WAVE,
// col = lane) or two on its halves (STEP = 32, phi_build2).  The lane of
// body d's coordinate p writes the acceleration a[p, d]; col 0 writes the
// other uniform values (v, dqdt_phi).  FACT: jac_phi in the factored form.
template <int NB, bool G, int STEP, bool FACT = false>
__device__ __attribute__((always_inline)) inline void phisalpha_cols(Smem<NB, G>& S, Swork<NB>& W, double h, bool anyKick,
                                                                     bool needQ, int col) {
  constexpr int N3 = 3 * NB, N4 = 4 * NB;
  const bool own = col < N4;
  int myd = -1, myp = -1;
  if (own) col_dp<NB>(col, myd, myp);
  double* PHI = W.f.phi;
  double* D = W.f.dadq;
  // phisalpha! (ahl21.jl:558-700, alpha = 2) for the Kepler pairs.
  // First one lane per pair evaluates the pair's factors (the divisions and
  // square roots) into pq; then each lane forms its dadq column (loop 1),
  // and, one row block (the three rows of body b) at a time, keeps the
  // block's rows of its jac_phi column in registers and accumulates them
  // from pq (loop 2; the factored form goes pair by pair instead, see
  // there).  Body b's rows receive
  // contributions only from the pairs that contain b, and in the reference's
  // lexicographic pair order those come partner by partner in increasing
  // order; looping over the partner c in increasing order therefore repeats
  // every accumulation in the reference's order.  a and dadq start at zero
  // (ahl21.jl:559-563); jac_phi keeps phic!'s terms.
  constexpr int NP = Swork<NB>::NP, PQ = Swork<NB>::PQ;
  double* pq = W.f.pq;
  auto pidx = [](int i, int j) { return i * (2 * NB - i - 1) / 2 + (j - i - 1); };
  const double coeff = (((2.0 * (h * h * h)) / 96.0) * 2.0) * NBG_GNEWT;
  // loop 1 factors (ahl21.jl:569-575): fac2 = G/r^3, fac_k, f3_k = fac_k * 3/r^2
  for (int p = col; p < NP; p += STEP) {
    int i = 0, q = p;
    while (q >= NB - 1 - i) { q -= NB - 1 - i; ++i; }
    const int j = i + 1 + q;
    double rij[3];
    for (int k = 0; k < 3; ++k) rij[k] = S.x[3 * i + k] - S.x[3 * j + k];
    const double r2 = dot3(rij);
    const double r3 = r2 * __builtin_sqrt(r2);
    const double fac2 = NBG_GNEWT / r3;
    double* o = pq + p * PQ;
    o[0] = fac2;
    for (int k = 0; k < 3; ++k) {
      const double fac = fac2 * rij[k];
      o[1 + k] = fac;
      o[4 + k] = fac * (3.0 / r2);
    }
  }
  wsync();
  // loop 1: accelerations and dadq (ahl21.jl:567-602), column by column.
  // For the lane's column (body myd, coordinate myp; myp = 3 for the mass)
  // the reference's conditional updates of pair (i, j) reduce, per role of
  // myd in the pair, to at most two additions in a fixed order, t1 then t2
  // (-0.0, which leaves any sum unchanged, when one does not apply).  Row
  // block b != myd receives them from one pair only, (b, myd) -- myd is b's
  // partner there -- so its entry is (0 + t1) + t2.  Row block myd receives
  // them from every partner c in increasing order, the order of the
  // reference's pair loop, summed in registers; so does the acceleration
  // a[myp, myd] on the lane whose coordinate is myp.  Every entry therefore
  // gets exactly the reference's additions.
  const bool mpos = myp >= 0 && myp < 3, mmass = myp == 3;
  const int ry = mpos ? myp : 0;
  if (own) {
#pragma unroll 1
    for (int b = 0; b < NB; ++b) {
      if (b == myd) continue;
      const int i = b < myd ? b : myd, j = b < myd ? myd : b;
      const bool bi = (b == i);
      double t[3] = {0.0, 0.0, 0.0};
      if (!(anyKick && S.pair[i + NB * j])) {
        const double* o = pq + pidx(i, j) * PQ;
        const double m = S.m[myd];  // b's partner's mass
        const double f2m = o[0] * m;
        const double rmy = S.x[3 * i + ry] - S.x[3 * j + ry];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double fac = o[1 + k];
          const double x3 = (o[4 + k] * m) * rmy;
          const double t1 = mmass ? (bi ? -fac : fac) : -x3;
          const double t2 = (mpos && myp == k) ? f2m : -0.0;
          t[k] = (0.0 + t1) + t2;
        }
      }
      for (int k = 0; k < 3; ++k) D[(3 * b + k) * N4 + col] = t[k];
    }
    double db[3] = {0.0, 0.0, 0.0};
    double ab = 0.0;
#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
      if (c == myd) continue;
      const int i = c < myd ? c : myd, j = c < myd ? myd : c;
      if (anyKick && S.pair[i + NB * j]) continue;
      const double* o = pq + pidx(i, j) * PQ;
      const bool bi = (myd == i);
      const double m = S.m[c];  // the partner's mass
      const double f2m = o[0] * m;
      const double rmy = S.x[3 * i + ry] - S.x[3 * j + ry];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double x3 = (o[4 + k] * m) * rmy;
        const double t1 = mpos ? x3 : -0.0;
        const double t2 = (mpos && myp == k) ? -f2m : -0.0;
        db[k] = (db[k] + t1) + t2;
      }
      const double mf = m * o[1 + ry];
      ab += bi ? -mf : mf;
    }
    for (int k = 0; k < 3; ++k) D[(3 * myd + k) * N4 + col] = db[k];
    if (mpos) S.a[3 * myd + myp] = ab;
  }
  wsync();
  // loop 2 factors (ahl21.jl:609-640), needing the accelerations above.
  // Products of a pair scalar with a component (f5 = fac * (5 / r2) and
  // ff = (3 * fac1) * rij[k]) are stored as the scalar and formed where they
  // are used, with the same operations.
  for (int p = col; p < NP; p += STEP) {
    int i = 0, q = p;
    while (q >= NB - 1 - i) { q -= NB - 1 - i; ++i; }
    const int j = i + 1 + q;
    double rij[3], aij[3];
    for (int k = 0; k < 3; ++k) {
      aij[k] = S.a[3 * i + k] - S.a[3 * j + k];
      rij[k] = S.x[3 * i + k] - S.x[3 * j + k];
    }
    const double r2 = dot3(rij);
    const double r1 = __builtin_sqrt(r2);
    const double ardot = dot3(aij, rij);
    const double fac1 = coeff / ((r2 * r2) * r1);
    const double mi = S.m[i], mj = S.m[j];
    const double fac2 = ((2.0 * NBG_GNEWT) * (mi + mj)) / r1 + 3.0 * ardot;
    double* o = pq + p * PQ;
    for (int k = 0; k < 3; ++k) {
      o[k] = fac1 * (rij[k] * fac2 - r2 * aij[k]);                                             // fac
      o[4 + k] = (((2.0 * NBG_GNEWT) * fac1) * rij[k]) / r1;                                   // fm
      o[7 + k] = ((-2.0) * fac1) * ((((rij[k] * NBG_GNEWT) * (mi + mj)) / (r2 * r1)) + aij[k]);  // fr
    }
    o[3] = 5.0 / r2;       // f5 = fac * o[3]
    // (fac1 * 3) * rij[k] multiplies aij[p] in fac3 and is ff = (3 * fac1) * rij[k]
    // as well (IEEE multiplication commutes)
    o[10] = fac1 * 3.0;
    o[11] = fac1 * fac2;   // fd
    o[12] = (-fac1) * r2;  // fa
  }
  wsync();
  // loop 2: velocity corrections, dqdt_phi and jac_phi (ahl21.jl:607-698),
  // with the same role reduction as loop 1: per k, at most three
  // role-dependent additions (t1, t2, t3), then the dadq terms.
  const double q3h = 3.0 / h;
  if constexpr (FACT) {
    // Factored (rounds differently from the reference's order): for pair
    // (i, j) the rows of body i receive mj * X and those of body j -mi * X
    // of the same column term X, plus the first mass derivative on the
    // partner's mass column, so X is formed once per pair.  Pairs run in the
    // reference's lexicographic order, so body i's rows (and its uniform
    // velocity and dqdt_phi, summed exactly as in the reference) stay in
    // registers across the partners j > i; a body's rows as a partner j are
    // read and written back in LDS.  jac_phi starts at phic!'s terms.
    if (own && !anyKick)
      for (int r = 0; r < N3; ++r) PHI[r * N4 + col] = 0.0;
    int p = 0;
#pragma unroll 1
    for (int i = 0; i < NB - 1; ++i) {
      double vi[3], vie[3], qi[3], pi[3], xi[3], Di[3];
      for (int k = 0; k < 3; ++k) {
        vi[k] = S.v[3 * i + k];
        vie[k] = S.ve[3 * i + k];
        qi[k] = needQ ? S.dqdt_phi[7 * i + 3 + k] : 0.0;
        pi[k] = own ? PHI[(3 * i + k) * N4 + col] : 0.0;
        xi[k] = S.x[3 * i + k];
        Di[k] = own ? D[(3 * i + k) * N4 + col] : 0.0;
      }
      const double mi = S.m[i], ai = S.a[3 * i + ry];
      const bool di = myd == i;
#pragma unroll 1
      for (int j = i + 1; j < NB; ++j, ++p) {
        if (anyKick && S.pair[i + NB * j]) continue;
        const double* o = pq + p * PQ;
        const double mj = S.m[j];
        double rij[3], Dj[3], vj[3], vje[3], qj[3], pj[3];
        for (int k = 0; k < 3; ++k) {
          rij[k] = xi[k] - S.x[3 * j + k];
          Dj[k] = own ? D[(3 * j + k) * N4 + col] : 0.0;
          vj[k] = S.v[3 * j + k];
          vje[k] = S.ve[3 * j + k];
          qj[k] = needQ ? S.dqdt_phi[7 * j + 3 + k] : 0.0;
          pj[k] = own ? PHI[(3 * j + k) * N4 + col] : 0.0;
        }
        double dot = 0.0;
        for (int k = 0; k < 3; ++k) dot += rij[k] * (Di[k] - Dj[k]);
        const double rmy = ry == 0 ? rij[0] : (ry == 1 ? rij[1] : rij[2]);
        const double amy = ai - S.a[3 * j + ry];
        const bool dj = myd == j;
        const double fd = o[11], fa = o[12];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double fac = o[k];
          kahan(vi[k], vie[k], mj * fac);
          kahan(vj[k], vje[k], -(mi * fac));
          if (needQ) {
            qi[k] += (q3h * mj) * fac;
            qj[k] += -((q3h * mi) * fac);
          }
          const double ff = o[10] * rij[k];
          const double base = (((myp == k) ? fd : 0.0) - (fac * o[3]) * rmy) + (o[7 + k] * rmy + ff * amy);
          const double role = mpos ? (di ? base : (dj ? -base : 0.0)) : ((mmass && (di || dj)) ? o[4 + k] : 0.0);
          const double X = (role + fa * (Di[k] - Dj[k])) + ff * dot;
          pi[k] += mj * X + ((mmass && dj) ? fac : 0.0);
          pj[k] += (-mi) * X + ((mmass && di) ? -fac : 0.0);
        }
        for (int k = 0; k < 3; ++k) {
          if (own) PHI[(3 * j + k) * N4 + col] = pj[k];
          if (col == 0) {
            S.v[3 * j + k] = vj[k];
            S.ve[3 * j + k] = vje[k];
            if (needQ) S.dqdt_phi[7 * j + 3 + k] = qj[k];
          }
        }
      }
      for (int k = 0; k < 3; ++k) {
        if (own) PHI[(3 * i + k) * N4 + col] = pi[k];
        if (col == 0) {
          S.v[3 * i + k] = vi[k];
          S.ve[3 * i + k] = vie[k];
          if (needQ) S.dqdt_phi[7 * i + 3 + k] = qi[k];
        }
      }
    }
  } else {
#pragma unroll 1
  for (int b = 0; b < NB; ++b) {
    double vb[3], veb[3], qb[3], pb[3];
    for (int k = 0; k < 3; ++k) {
      vb[k] = S.v[3 * b + k];
      veb[k] = S.ve[3 * b + k];
      qb[k] = needQ ? S.dqdt_phi[7 * b + 3 + k] : 0.0;
      pb[k] = (anyKick && own) ? PHI[(3 * b + k) * N4 + col] : 0.0;
    }
#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
      if (c == b) continue;
      const int i = c < b ? c : b, j = c < b ? b : c;
      if (anyKick && S.pair[i + NB * j]) continue;
      const double* o = pq + pidx(i, j) * PQ;
      double rij[3], Di[3], Dj[3];
      for (int k = 0; k < 3; ++k) {
        rij[k] = S.x[3 * i + k] - S.x[3 * j + k];
        Di[k] = own ? D[(3 * i + k) * N4 + col] : 0.0;
        Dj[k] = own ? D[(3 * j + k) * N4 + col] : 0.0;
      }
      double dot = 0.0;
      if (own)
        for (int k = 0; k < 3; ++k) dot += rij[k] * (Di[k] - Dj[k]);
      const double rmy = ry == 0 ? rij[0] : (ry == 1 ? rij[1] : rij[2]);
      const double amy = S.a[3 * i + ry] - S.a[3 * j + ry];
      const bool bi = (b == i);
      const double m = bi ? S.m[j] : S.m[i];  // the partner's mass
      const bool self_pos = mpos && myd == b, other_pos = mpos && myd == c;
      const bool self_mass = mmass && myd == b, other_mass = mmass && myd == c;
      const double fd = o[11], fa = o[12];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double fac = o[k], f5 = fac * o[3], fm = o[4 + k], fr = o[7 + k], ff = o[10] * rij[k];
        const double mf = m * fac;
        kahan(vb[k], veb[k], bi ? mf : -mf);
        if (needQ) {
          const double qm = (q3h * m) * fac;
          qb[k] += bi ? qm : -qm;
        }
        const double diff = Di[k] - Dj[k];
        const double x5 = (f5 * m) * rmy;
        const double y = fm * m;
        const double sy = bi ? y : -y, sfac = bi ? fac : -fac;
        const double fdm = fd * m;
        const double fac3 = fr * rmy + ff * amy;
        const double mf3 = m * fac3;
        const bool pk = myp == k;
        const double t1 = self_pos ? -x5 : (other_pos ? x5 : (self_mass ? sy : (other_mass ? sfac : -0.0)));
        const double t2 = (pk && self_pos) ? fdm : ((pk && other_pos) ? -fdm : (other_mass ? sy : -0.0));
        const double t3 = self_pos ? mf3 : (other_pos ? -mf3 : -0.0);
        const double td = (fa * m) * diff, tf = (ff * m) * dot;
        pb[k] = (((((pb[k] + t1) + t2) + t3) + (bi ? td : -td)) + (bi ? tf : -tf));
      }
    }
    if (own)
      for (int k = 0; k < 3; ++k) PHI[(3 * b + k) * N4 + col] = pb[k];
    if (col == 0)
      for (int k = 0; k < 3; ++k) {
        S.v[3 * b + k] = vb[k];
        S.ve[3 * b + k] = veb[k];
        if (needQ) S.dqdt_phi[7 * b + 3 + k] = qb[k];
      }
  }
  }
}

// ---------------------------------------------------------------------------
// phic! (ahl21.jl:392-552, kicked pairs) and phisalpha! (ahl21.jl:558-700,
// Kepler pairs, alpha = 2) with Jacobian.  Phi and dadq columns are owned by
// lanes; v and dqdt_phi are uniform (lane 0 stores), and phisalpha!'s
// accelerations are stored by the lane of each body and coordinate.
// ---------------------------------------------------------------------------
template <int NB, bool G, bool FACT = false>
__device__ void phi_build(Smem<NB, G>& S, Swork<NB>& W, double h, bool anyKick, bool needQ, int lane) {
  constexpr int N3 = 3 * NB, N4 = 4 * NB;
  const bool own = lane < N4;
  int myd = -1, myp = -1;
  if (own) col_dp<NB>(lane, myd, myp);
  double* PHI = W.f.phi;
  double* D = W.f.dadq;
  if (anyKick) {
    // phic!: a, dadq, jac_phi set to zero (ahl21.jl:393-399)
    if (own)
      for (int r = 0; r < N3; ++r) { PHI[r * N4 + lane] = 0.0; D[r * N4 + lane] = 0.0; }
    if (lane == 0)
      for (int r = 0; r < N3; ++r) S.a[r] = 0.0;
    wsync();
  }
  if (anyKick) {
    const double coeff = ((h * h * h) / 36.0) * NBG_GNEWT;
    // loop 1: kicks 2h/3, accelerations, dadq (ahl21.jl:403-466)
    for (int i = 0; i < NB - 1; ++i)
      for (int j = i + 1; j < NB; ++j) {
        if (!S.pair[i + NB * j]) continue;
        double rij[3];
        for (int k = 0; k < 3; ++k) rij[k] = S.x[3 * i + k] - S.x[3 * j + k];
        const double r2inv = 1.0 / dot3(rij);
        const double r3inv = r2inv * __builtin_sqrt(r2inv);
        const double mi = S.m[i], mj = S.m[j];
        for (int k = 0; k < 3; ++k) {
          double fac = (NBG_GNEWT * rij[k]) * r3inv;
          double facv = ((fac * 2.0) * h) / 3.0;
          const int ri = 3 * i + k, rj = 3 * j + k;
          if (lane == 0) {
            kahan(S.v[ri], S.ve[ri], (-mj) * facv);
            kahan(S.v[rj], S.ve[rj], mi * facv);
            if (needQ) {
              S.dqdt_phi[7 * i + 3 + k] -= ((1.0 / h) * mj) * facv;
              S.dqdt_phi[7 * j + 3 + k] += ((1.0 / h) * mi) * facv;
            }
            S.a[ri] -= mj * fac;
            S.a[rj] += mi * fac;
          }
          if (own) {
            if (myp == 3 && myd == j) PHI[ri * N4 + lane] -= facv;
            if (myp == 3 && myd == i) PHI[rj * N4 + lane] += facv;
            const double fv3 = facv * (3.0 * r2inv);
            if (myp < 3 && myd == i) {
              PHI[ri * N4 + lane] += (fv3 * mj) * rij[myp];
              PHI[rj * N4 + lane] -= (fv3 * mi) * rij[myp];
            }
            if (myp < 3 && myd == j) {
              PHI[ri * N4 + lane] -= (fv3 * mj) * rij[myp];
              PHI[rj * N4 + lane] += (fv3 * mi) * rij[myp];
            }
            const double facd = (((2.0 * h) / 3.0) * NBG_GNEWT) * r3inv;
            if (myp == k && myd == i) {
              PHI[ri * N4 + lane] -= facd * mj;
              PHI[rj * N4 + lane] += facd * mi;
            }
            if (myp == k && myd == j) {
              PHI[ri * N4 + lane] += facd * mj;
              PHI[rj * N4 + lane] -= facd * mi;
            }
            if (myp == 3 && myd == j) D[ri * N4 + lane] -= fac;
            if (myp == 3 && myd == i) D[rj * N4 + lane] += fac;
            const double f3 = fac * (3.0 * r2inv);
            if (myp < 3 && myd == i) {
              D[ri * N4 + lane] += (f3 * mj) * rij[myp];
              D[rj * N4 + lane] -= (f3 * mi) * rij[myp];
            }
            if (myp < 3 && myd == j) {
              D[ri * N4 + lane] -= (f3 * mj) * rij[myp];
              D[rj * N4 + lane] += (f3 * mi) * rij[myp];
            }
            const double fg = NBG_GNEWT * r3inv;
            if (myp == k && myd == i) {
              D[ri * N4 + lane] -= fg * mj;
              D[rj * N4 + lane] += fg * mi;
            }
            if (myp == k && myd == j) {
              D[ri * N4 + lane] += fg * mj;
              D[rj * N4 + lane] -= fg * mi;
            }
          }
        }
        wsync();
      }
    // loop 2: fourth-order correction for kicked pairs (ahl21.jl:471-550)
    for (int i = 0; i < NB - 1; ++i)
      for (int j = i + 1; j < NB; ++j) {
        if (!S.pair[i + NB * j]) continue;
        double rij[3], aij[3];
        for (int k = 0; k < 3; ++k) {
          aij[k] = S.a[3 * i + k] - S.a[3 * j + k];
          rij[k] = S.x[3 * i + k] - S.x[3 * j + k];
        }
        double dot = 0.0;
        if (own)
          for (int k = 0; k < 3; ++k) dot += rij[k] * (D[(3 * i + k) * N4 + lane] - D[(3 * j + k) * N4 + lane]);
        const double r2 = dot3(rij);
        const double r1 = __builtin_sqrt(r2);
        const double ardot = dot3(aij, rij);
        const double fac1 = coeff / ((r2 * r2) * r1);
        const double fac2 = 3.0 * ardot;
        const double mi = S.m[i], mj = S.m[j];
        for (int k = 0; k < 3; ++k) {
          const double fac = fac1 * (rij[k] * fac2 - r2 * aij[k]);
          const int ri = 3 * i + k, rj = 3 * j + k;
          if (lane == 0) {
            kahan(S.v[ri], S.ve[ri], mj * fac);
            kahan(S.v[rj], S.ve[rj], (-mi) * fac);
            if (needQ) {
              S.dqdt_phi[7 * i + 3 + k] += ((3.0 / h) * mj) * fac;
              S.dqdt_phi[7 * j + 3 + k] -= ((3.0 / h) * mi) * fac;
            }
          }
          if (own) {
            double* pi_ = &PHI[ri * N4 + lane];
            double* pj_ = &PHI[rj * N4 + lane];
            if (myp == 3 && myd == j) *pi_ += fac;
            if (myp == 3 && myd == i) *pj_ -= fac;
            const double f5 = fac * (5.0 / r2);
            if (myp < 3 && myd == i) { *pi_ -= (f5 * mj) * rij[myp]; *pj_ += (f5 * mi) * rij[myp]; }
            if (myp < 3 && myd == j) { *pi_ += (f5 * mj) * rij[myp]; *pj_ -= (f5 * mi) * rij[myp]; }
            const double fd = fac1 * fac2;
            if (myp == k && myd == i) { *pi_ += fd * mj; *pj_ -= fd * mi; }
            if (myp == k && myd == j) { *pi_ -= fd * mj; *pj_ += fd * mi; }
            const double fr = ((-2.0) * fac1) * aij[k];
            if (myp < 3 && (myd == i || myd == j)) {
              const double fac3 = fr * rij[myp] + (((fac1 * 3.0) * rij[k]) * aij[myp]);
              if (myd == i) { *pi_ += mj * fac3; *pj_ -= mi * fac3; }
              else { *pi_ -= mj * fac3; *pj_ += mi * fac3; }
            }
            const double fa = (-fac1) * r2;
            const double diff = D[ri * N4 + lane] - D[rj * N4 + lane];
            *pi_ += (fa * mj) * diff;
            *pj_ -= (fa * mi) * diff;
            const double ff = (3.0 * fac1) * rij[k];
            *pi_ += (ff * mj) * dot;
            *pj_ -= (ff * mi) * dot;
          }
        }
        wsync();
      }
  }
  phisalpha_cols<NB, G, WAVE, FACT>(S, W, h, anyKick, needQ, lane);
}

// ---------------------------------------------------------------------------
// Jacobian column kernels (lane owns column c).
// ---------------------------------------------------------------------------
// Position rows of column c (pre-update), the operand of K*J and Phi*J.
template <int NB, bool G> __device__ inline void load_pos(const Smem<NB, G>& S, int c, double* P) {
  const auto J_ = jj<NB>(S);
  constexpr int N7 = 7 * NB;
#pragma unroll
  for (int d = 0; d < NB; ++d)
#pragma unroll
    for (int p = 0; p < 3; ++p) P[3 * d + p] = J_[(unsigned)((6 * d + p) * N7 + c)];
}

// J[velocity rows, c] += (B * J)[velocity rows, c] with Kahan compensation,
// B = K or Phi (3n x 4n block); then the zero-addend Kahan pass over the
// position rows that comp_sum_matrix! applies to the rest of the matrix.
template <int NB, bool G> __device__ inline void vel_product_add(Smem<NB, G>& S, const double* B, int c, const double* P) {
  const auto J_ = jj<NB>(S);
  constexpr int N3 = 3 * NB, N4 = 4 * NB, N7 = 7 * NB;
#pragma unroll 1
  for (int rv = 0; rv < N3; ++rv) {
    double s = 0.0;
    const double* Br = B + rv * N4;
#pragma unroll
    for (int d = 0; d < NB; ++d) {
      s = fmad(Br[3 * d + 0], P[3 * d + 0], s);
      s = fmad(Br[3 * d + 1], P[3 * d + 1], s);
      s = fmad(Br[3 * d + 2], P[3 * d + 2], s);
      s = fmad(Br[N3 + d], (c == 7 * d + 6) ? 1.0 : 0.0, s);
    }
    const int b = rv / 3, k = rv - 3 * b;
    const int r = 6 * b + 3 + k;
    kahan(J_[(unsigned)(r * N7 + c)], J_[(unsigned)(Smem<NB>::JOFF + r * N7 + c)], s);
  }
}

template <int NB, bool G> __device__ inline void zero_add_rows(Smem<NB, G>& S, int c, bool pos, bool vel) {
  const auto J_ = jj<NB>(S);
  constexpr int N7 = 7 * NB;
#pragma unroll 1
  for (int b = 0; b < NB; ++b)
    for (int k = 0; k < 6; ++k) {
      if ((k < 3 && !pos) || (k >= 3 && !vel)) continue;
      const int r = 6 * b + k;
      auto& J = J_[(unsigned)(r * N7 + c)];
      auto& E = J_[(unsigned)(Smem<NB>::JOFF + r * N7 + c)];
      E += 0.0;
      const double t = J + E;
      E += J - t;
      J = t;
    }
}

// drift_grad! Jacobian part (ahl21.jl:326-328)
template <int NB, bool G> __device__ inline void drift_cols(Smem<NB, G>& S, int c, double hh) {
  const auto J_ = jj<NB>(S);
  constexpr int N7 = 7 * NB;
#pragma unroll 1
  for (int b = 0; b < NB; ++b)
    for (int k = 0; k < 3; ++k) {
      const int r = 6 * b + k;
      kahan(J_[(unsigned)(r * N7 + c)], J_[(unsigned)(Smem<NB>::JOFF + r * N7 + c)], hh * J_[(unsigned)((r + 3) * N7 + c)]);
    }
}

// Pair map on rows (i, j): copy_submatrix!, mul!, comp_sum_matrix!,
// ypoc_submatrix! (ahl21.jl:31-42, utils.jl:51-100) for column c.  Row q of
// jac_ij * a is summed over k = 0..13 in ascending order with one fused
// multiply-add per term, as jij_coef's coefficients would give it: the
// Kepler-block coefficient +-(0 + m*kep) enters as m*kep times +-a[k] (the
// running sum starts at +0 and can never become -0, so the sign of a zero
// coefficient cannot change it), and the mass columns 6 and 13 take their
// coefficients from the slot, where pair_jacobian stored them.
template <int NB, bool G, bool FACT = false>
__device__ inline void pair_cols(Smem<NB, G>& S, const PairSlot& sl, int i, int j, int c) {
  const auto J_ = jj<NB>(S);
  constexpr int N7 = 7 * NB;
  const int ri = 6 * i, rj = 6 * j;
  double a[12], e[12];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    a[k] = J_[(unsigned)((ri + k) * N7 + c)];
    a[6 + k] = J_[(unsigned)((rj + k) * N7 + c)];
    e[k] = J_[(unsigned)(Smem<NB>::JOFF + (ri + k) * N7 + c)];
    e[6 + k] = J_[(unsigned)(Smem<NB>::JOFF + (rj + k) * N7 + c)];
  }
  const double a6 = (c == 7 * i + 6) ? 1.0 : 0.0;
  const double a13 = (c == 7 * j + 6) ? 1.0 : 0.0;
  const double mi = sl.mi, mj = sl.mj;
  double o[12];
  if constexpr (FACT) {
    // Factored (rounds differently from the reference's mul!): the Kepler
    // block of jac_ij gives the rows of i +mj*kep and those of j -mi*kep
    // times the same difference a_i - a_j, so kep*(a_i - a_j) is formed once.
    double d[6], sk[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = a[k] - a[6 + k];
#pragma unroll
    for (int rk = 0; rk < 6; ++rk) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) t = fmad(sl.kep[rk + 6 * k], d[k], t);
      sk[rk] = t;
    }
#pragma unroll
    for (int rk = 0; rk < 6; ++rk) {
      o[rk] = fmad(sl.mc[12 + rk], a13, fmad(sl.mc[rk], a6, mj * sk[rk]));
      o[6 + rk] = fmad(sl.mc[18 + rk], a13, fmad(sl.mc[6 + rk], a6, (-mi) * sk[rk]));
    }
  } else {
#pragma unroll
  for (int q = 0; q < 12; ++q) o[q] = 0.0;
  // k = 0..5: rows of i get +mj*kep * a_i[k], rows of j -mi*kep * a_i[k]
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int rk = 0; rk < 6; ++rk) {
      const double kv = sl.kep[rk + 6 * k];
      o[rk] = fmad(mj * kv, a[k], o[rk]);
      o[6 + rk] = fmad(mi * kv, -a[k], o[6 + rk]);
    }
#pragma unroll
  for (int q = 0; q < 12; ++q) o[q] = fmad(sl.mc[q], a6, o[q]);
  // k = 7..12: rows of i get -mj*kep * a_j[k], rows of j +mi*kep * a_j[k]
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int rk = 0; rk < 6; ++rk) {
      const double kv = sl.kep[rk + 6 * k];
      o[rk] = fmad(mj * kv, -a[6 + k], o[rk]);
      o[6 + rk] = fmad(mi * kv, a[6 + k], o[6 + rk]);
    }
#pragma unroll
  for (int q = 0; q < 12; ++q) o[q] = fmad(sl.mc[12 + q], a13, o[q]);
  }
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    double sv = a[q], se = e[q];
    se += o[q];
    const double t = sv + se;
    se += sv - t;
    sv = t;
    const int r = q < 6 ? ri + q : rj + (q - 6);
    J_[(unsigned)(r * N7 + c)] = sv;
    J_[(unsigned)(Smem<NB>::JOFF + r * N7 + c)] = se;
  }
}

// ---------------------------------------------------------------------------
// One pair sweep of ahl21! (forward: ahl21.jl:25-45, drift_first = true;
// backward: :56-78, drift_first = false) in chunks of pairs:
//  A. sequentially per pair: the universal-variable Kepler drift
//     (jac_delxv_gamma!, ahl21.jl:766-890) and the Kahan state update of
//     kepler_driftij_gamma! (:706-733), all lanes alike, lane 0 stores;
//  B. the analytic Jacobians (compute_jacobian_gamma!, :896-1139) of all the
//     chunk's pairs at once, one lane per pair;
//  C. sequentially per pair: jac_ij (:734-758), the Jacobian row update
//     (copy_submatrix!, mul!, comp_sum_matrix!, ypoc_submatrix!) with one lane
//     per column, and the dq/dt update.
// The state never reads the Jacobian and the Jacobian updates never read the
// state, so running all state updates of a chunk before its Jacobian updates
// gives the reference's results exactly.
// ---------------------------------------------------------------------------
// Phase A for one lane: the Kepler drift of pair (i, j) of system S into
// slot sl and the Kahan update of the two bodies' state.
template <int NB, bool G>
__device__ __attribute__((always_inline)) inline void pair_level_lane(Smem<NB, G>& S, PairSlot& sl, int i, int j,
                                                                      double hh, bool drift_first) {
  // every LDS operand is read into registers before anything is written
  // back: the compiler cannot tell the slot from the state arrays, so
  // interleaved reads and writes would each wait for the previous write (a
  // dozen LDS round trips in a row per level)
  {
    double xi[3], xj[3], vi[3], vj[3];
    for (int k = 0; k < 3; ++k) {
      xi[k] = S.x[3 * i + k];
      xj[k] = S.x[3 * j + k];
      vi[k] = S.v[3 * i + k];
      vj[k] = S.v[3 * j + k];
    }
    for (int k = 0; k < 3; ++k) {
      sl.x0[k] = xi[k] - xj[k];
      sl.v0[k] = vi[k] - vj[k];
    }
  }
  const double m_i = S.m[i], m_j = S.m[j];
  const double gm = NBG_GNEWT * (m_i + m_j);
  if (gm != 0.0) {
    pair_drift(&sl, gm, hh, drift_first);
    const double mijinv = 1.0 / (m_i + m_j);
    const double mi = m_i * mijinv;
    const double mj = m_j * mijinv;
    double d[6], xi[3], xj[3], vi[3], vj[3], exi[3], exj[3], evi[3], evj[3];
    for (int k = 0; k < 6; ++k) d[k] = sl.delxv[k];
    for (int k = 0; k < 3; ++k) {
      xi[k] = S.x[3 * i + k];
      exi[k] = S.xe[3 * i + k];
      xj[k] = S.x[3 * j + k];
      exj[k] = S.xe[3 * j + k];
      vi[k] = S.v[3 * i + k];
      evi[k] = S.ve[3 * i + k];
      vj[k] = S.v[3 * j + k];
      evj[k] = S.ve[3 * j + k];
    }
    for (int k = 0; k < 3; ++k) {
      kahan(xi[k], exi[k], mj * d[k]);
      kahan(xj[k], exj[k], (-mi) * d[k]);
    }
    for (int k = 0; k < 3; ++k) {
      kahan(vi[k], evi[k], mj * d[3 + k]);
      kahan(vj[k], evj[k], (-mi) * d[3 + k]);
    }
    for (int k = 0; k < 3; ++k) {
      S.x[3 * i + k] = xi[k];
      S.xe[3 * i + k] = exi[k];
      S.x[3 * j + k] = xj[k];
      S.xe[3 * j + k] = exj[k];
      S.v[3 * i + k] = vi[k];
      S.ve[3 * i + k] = evi[k];
      S.v[3 * j + k] = vj[k];
      S.ve[3 * j + k] = evj[k];
    }
    sl.mi = mi; sl.mj = mj; sl.mijinv = mijinv; sl.m_i = m_i; sl.m_j = m_j;
  }
  sl.valid = (gm != 0.0); sl.i = i; sl.j = j;
}

// Phase C of one chunk of `cnt` pairs (slots[0..cnt)) of system S, with the
// row update in the reference's order or (FACT) factored.
template <int NB, bool G, int STEP = WAVE, bool FACT = false>
__device__ __attribute__((always_inline)) inline void sweep_rows(Smem<NB, G>& S, const PairSlot* slots, int cnt,
                                                                 bool needJ, bool needQ, int col) {
  constexpr int N7 = 7 * NB;
  int lastq = -1;  // last slot of this chunk whose pair has mass
#pragma unroll 1
  for (int q = 0; q < cnt; ++q) {
    const PairSlot& sl = slots[q];
    const int i = sl.i, j = sl.j;
    if (sl.valid) {
      lastq = q;
      if (needQ && col < 6) {  // dqdt_ij (ahl21.jl:743-748)
        S.dqij[col] = sl.mj * sl.kep[col + 42];
        S.dqij[7 + col] = (-sl.mi) * sl.kep[col + 42];
      }
    }
    // (gm == 0: the reference returns before touching jac_ij / dqdt_ij and
    // the caller reuses the previous values: the last pair with mass, or
    // jlast when that lies in an earlier chunk; dqij keeps its values.)
    const PairSlot& src = lastq >= 0 ? slots[lastq] : S.jlast;
    if (needJ)
      for (int c = col; c < N7; c += STEP) pair_cols<NB, G, FACT>(S, src, i, j, c);
    if (needQ) {
      double q1[14];
      for (int k = 0; k < 7; ++k) {
        q1[k] = S.dqdt[7 * i + k];
        q1[7 + k] = S.dqdt[7 * j + k];
      }
      if (col < 14) {
        double s2 = 0.0;
        if (col != 6 && col != 13) {
          const bool rowi = col < 7;
          const int rk = rowi ? col : col - 7;
          for (int k = 0; k < 14; ++k) s2 = fmad(jij_coef(src, rowi, rk, k), q1[k], s2);
        } else {
          for (int k = 0; k < 14; ++k) s2 = fmad(0.0, q1[k], s2);
        }
        const double nd = (S.dqij[col] * 0.5) + (q1[col] + s2);
        S.dqij[col] = nd;
        S.dqdt[col < 7 ? 7 * i + col : 7 * j + col - 7] = nd;
      }
    }
    wsync();
  }
  // keep the last pair with mass for a massless pair at the start of a later chunk
  if (S.seq && lastq >= 0) {
    const double* from = reinterpret_cast<const double*>(&slots[lastq]);
    double* to = reinterpret_cast<double*>(&S.jlast);
    for (int w = col; w < (int)(sizeof(PairSlot) / sizeof(double)); w += STEP) to[w] = from[w];
    wsync();
  }
}

template <int NB, bool G, bool FACT = false>
__device__ __attribute__((always_inline)) inline void sweep(Smem<NB, G>& S, Swork<NB>& W, int dir, double hh, bool needJ,
                                                            bool needQ, int lane) {
  constexpr int CH = Swork<NB>::CH;
  const bool drift_first = (dir == 0);
  const int nlev = S.nlv[dir];
  const uint8_t* pl = S.plist[dir];
  const uint8_t* lw = S.lvw[dir];
  int lev = 0, c0 = 0;
  while (lev < nlev) {
    // a chunk is a run of whole levels holding at most CH pairs
    int cnt = 0, lev_end = lev;
    while (lev_end < nlev && cnt + lw[lev_end] <= CH) cnt += lw[lev_end++];
    // ---- A: Kepler drifts and state updates, one level at a time: the
    // pairs of a level share no body, so lane l solves pair l of the level
    int q0 = 0;
#pragma unroll 1
    for (int L = lev; L < lev_end; ++L) {
      const int w = lw[L];
      if (lane < w) {
        const int q = q0 + lane;
        pair_level_lane<NB>(S, W.slot[q], pl[2 * (c0 + q)], pl[2 * (c0 + q) + 1], hh, drift_first);
      }
      q0 += w;
      wsync();
    }
    lev = lev_end;
    c0 += cnt;
    if (!(needJ || needQ)) continue;
    // ---- B: analytic Kepler-drift Jacobians, one lane per pair
    if (lane < cnt && W.slot[lane].valid) pair_jacobian(&W.slot[lane], drift_first);
    wsync();
    // ---- C: jac_ij, Jacobian rows and dq/dt, in sweep order
    sweep_rows<NB, G, WAVE, FACT>(S, W.slot, cnt, needJ, needQ, lane);
  }
}

// The same sweep for two systems of one wave (Duo) with the same level
// schedule (same body count and pair kinds, no massless pair): phases A and
// B serve both systems' pairs at once -- lanes [0, w) the first system's
// pairs of a level, [w, 2w) the second's; phase C runs for each in turn.
// Each system's slots are its own (slots[0, ch) and slots[ch, 2 ch)), and
// every operation of a system is the one-system sweep's, so each system's
// results are exactly its one-system results.
template <int NB, bool G, bool HALF = false, bool FACT = false>
__device__ __attribute__((always_inline)) inline void sweep2(Smem<NB, G>& S0, Smem<NB, G>& S1, PairSlot* slots, int ch,
                                                             int dir, double hh0, double hh1, bool nJ0, bool nQ0,
                                                             bool nJ1, bool nQ1, int lane) {
  const bool drift_first = (dir == 0);
  const int nlev = S0.nlv[dir];
  const uint8_t* pl = S0.plist[dir];
  const uint8_t* lw = S0.lvw[dir];
  const bool pj0 = nJ0 || nQ0, pj1 = nJ1 || nQ1;
  int lev = 0, c0 = 0;
  while (lev < nlev) {
    int cnt = 0, lev_end = lev;
    while (lev_end < nlev && cnt + lw[lev_end] <= ch) cnt += lw[lev_end++];
    int q0 = 0;
#pragma unroll 1
    for (int L = lev; L < lev_end; ++L) {
      const int w = lw[L];
      if (lane < 2 * w) {
        const int sys = lane >= w ? 1 : 0;
        const int q = q0 + lane - sys * w;
        Smem<NB, G>& S = sys ? S1 : S0;
        pair_level_lane<NB>(S, slots[sys * ch + q], pl[2 * (c0 + q)], pl[2 * (c0 + q) + 1], sys ? hh1 : hh0,
                            drift_first);
      }
      q0 += w;
      wsync();
    }
    lev = lev_end;
    c0 += cnt;
    if (!(pj0 || pj1)) continue;
    if (lane < 2 * cnt) {
      const int sys = lane >= cnt ? 1 : 0;
      PairSlot* sl = slots + sys * ch + (lane - sys * cnt);
      if ((sys ? pj1 : pj0) && sl->valid) pair_jacobian(sl, drift_first);
    }
    wsync();
    if constexpr (HALF) {  // both systems' rows at once, one per half of the wave
      const bool sys = lane >= 32;
      sweep_rows<NB, G, 32, FACT>(sys ? S1 : S0, slots + (sys ? ch : 0), cnt, sys ? nJ1 : nJ0, sys ? nQ1 : nQ0,
                                  lane & 31);
    } else {
      if (pj0) sweep_rows<NB, G, WAVE, FACT>(S0, slots, cnt, nJ0, nQ0, lane);
      if (pj1) sweep_rows<NB, G, WAVE, FACT>(S1, slots + ch, cnt, nJ1, nQ1, lane);
    }
  }
}

// ---------------------------------------------------------------------------
// One AHL21 step: ahl21!(s, d, h) (ahl21.jl:5-95).  needJ: propagate the
// Jacobian; needQ: compute dq/dt.  The state arithmetic is identical for
// every flag combination (so grad and no-grad runs agree bit for bit, as the
// reference's no-grad map ahl21_no_grad.jl:6-18 does).
// ---------------------------------------------------------------------------
// The stages of one step between the pair sweeps: before the forward sweep
// (kickfast! and the first drift), between the sweeps (phic!/phisalpha!),
// after the backward sweep (the second drift and kickfast!).  The
// zero-addend compensated passes of comp_sum_matrix! over the rows an
// operation leaves unchanged (zero_add_rows) re-normalise pairs that their
// last update already normalised; the factored form (FACT) leaves them out.
template <int NB, bool G, int STEP = WAVE, bool FACT = false>
__device__ __attribute__((always_inline)) inline void step_pre(Smem<NB, G>& S, Swork<NB>& W, double h, bool needJ, bool needQ, bool anyKick, int col) {
  constexpr int N3 = 3 * NB, N4 = 4 * NB, N7 = 7 * NB;
  const double h2 = 0.5 * h;
  const double h6 = h / 6.0;
  // zero_out!(d); fill!(s.dqdt, 0)
  if (col < 14) S.dqij[col] = 0.0;
  if (S.seq)  // jac_ij = 0 at the start of the step (zero_out!, ahl21.jl:8)
    for (int w = col; w < (int)(sizeof(PairSlot) / sizeof(double)); w += STEP)
      reinterpret_cast<double*>(&S.jlast)[w] = 0.0;
  if (needQ)
    for (int r = col; r < N7; r += STEP) {
      S.dqdt[r] = 0.0;
      S.dqdt_phi[r] = 0.0;
      S.dqdt_kick[r] = 0.0;
    }
  wsync();
  // ---- kickfast!(s, d, h/6)
  if constexpr (STEP == WAVE) {  // kicked pairs: whole-wave form only
    if (anyKick) kick_build<NB>(S, W, h6, needQ, col);
  }
  wsync();
  if (needQ) {
    // dqdt_kick ./= 6; tmp7n = K*dqdt (dqdt = 0 here); dqdt .+= dqdt_kick .+ tmp7n
    for (int r = col; r < N7; r += STEP) {
      double tmp = 0.0;
      if (anyKick) {
        const int b = r / 7, k = r - 7 * b;
        if (k >= 3 && k < 6) {
          const double* Kr = W.f.dadq + (3 * b + k - 3) * N4;
          for (int d = 0; d < NB; ++d) {
            tmp = fmad(Kr[3 * d + 0], S.dqdt[7 * d + 0], tmp);
            tmp = fmad(Kr[3 * d + 1], S.dqdt[7 * d + 1], tmp);
            tmp = fmad(Kr[3 * d + 2], S.dqdt[7 * d + 2], tmp);
            tmp = fmad(Kr[N3 + d], S.dqdt[7 * d + 6], tmp);
          }
        }
      }
      const double dk = S.dqdt_kick[r] / 6.0;
      S.dqdt[r] = S.dqdt[r] + (dk + tmp);
    }
  }
  // ---- Jacobian: jac_copy = K*J (pre-drift), drift_grad!, J += jac_copy
  if (needJ)
    for (int c = col; c < N7; c += STEP) {
      double P[N3];  // the pre-drift position rows, the operand of K*J
      if (anyKick) load_pos<NB>(S, c, P);
      drift_cols<NB>(S, c, h2);
      if constexpr (!FACT) zero_add_rows<NB>(S, c, true, !anyKick);
      if (anyKick) vel_product_add<NB>(S, W.f.dadq, c, P);
    }
  // ---- drift of the state (drift_grad! / drift!): entries are independent
  for (int q = col; q < N3; q += STEP) {
    double xq = S.x[q], eq = S.xe[q];
    kahan(xq, eq, h2 * S.v[q]);
    S.x[q] = xq;
    S.xe[q] = eq;
  }
  wsync();
  if (needQ) {
    for (int q = col; q < N3; q += STEP) {
      const int b = q / 3, k = q - 3 * b;
      S.dqdt[7 * b + k] = 0.5 * S.v[q] + h2 * S.dqdt[7 * b + 3 + k];
    }
  }
  wsync();
}

template <int NB, bool G, int STEP = WAVE, bool FACT = false>
__device__ __attribute__((always_inline)) inline void step_mid_rest(Smem<NB, G>& S, Swork<NB>& W, bool needJ, bool needQ,
                                                                    int col);

template <int NB, bool G, bool FACT = false>
__device__ __attribute__((always_inline)) inline void step_mid(Smem<NB, G>& S, Swork<NB>& W, double h, bool needJ, bool needQ, bool anyKick, int lane) {
  // ---- phic!, phisalpha!
  phi_build<NB, G, FACT>(S, W, h, anyKick, needQ, lane);
  wsync();
  step_mid_rest<NB, G, WAVE, FACT>(S, W, needJ, needQ, lane);
}

// The rest of the stage between the sweeps, once jac_phi is built:
// J_vel += jac_phi * J_pos, and dq/dt.
template <int NB, bool G, int STEP, bool FACT>
__device__ __attribute__((always_inline)) inline void step_mid_rest(Smem<NB, G>& S, Swork<NB>& W, bool needJ, bool needQ,
                                                                    int col) {
  constexpr int N3 = 3 * NB, N4 = 4 * NB, N7 = 7 * NB;
  if (needJ)
    for (int c = col; c < N7; c += STEP) {
      double P[N3];
      load_pos<NB>(S, c, P);
      vel_product_add<NB>(S, W.f.phi, c, P);
      if constexpr (!FACT) zero_add_rows<NB>(S, c, true, false);
    }
  if (needQ) {
    // tmp7n = jac_phi*dqdt; dqdt .+= dqdt_phi .+ tmp7n
    double nv[2];
    int nr = 0;
    for (int r = col; r < N7; r += STEP, ++nr) {
      double tmp = 0.0;
      const int b = r / 7, k = r - 7 * b;
      if (k >= 3 && k < 6) {
        const double* Pr = W.f.phi + (3 * b + k - 3) * N4;
        for (int d = 0; d < NB; ++d) {
          tmp = fmad(Pr[3 * d + 0], S.dqdt[7 * d + 0], tmp);
          tmp = fmad(Pr[3 * d + 1], S.dqdt[7 * d + 1], tmp);
          tmp = fmad(Pr[3 * d + 2], S.dqdt[7 * d + 2], tmp);
          tmp = fmad(Pr[N3 + d], S.dqdt[7 * d + 6], tmp);
        }
      }
      nv[nr] = S.dqdt[r] + (S.dqdt_phi[r] + tmp);
    }
    wsync();
    nr = 0;
    for (int r = col; r < N7; r += STEP, ++nr) S.dqdt[r] = nv[nr];
  }
  wsync();
}

template <int NB, bool G, int STEP = WAVE, bool FACT = false>
__device__ __attribute__((always_inline)) inline void step_post(Smem<NB, G>& S, Swork<NB>& W, double h, bool needJ, bool needQ, bool anyKick, int col) {
  constexpr int N3 = 3 * NB, N4 = 4 * NB, N7 = 7 * NB;
  const double h2 = 0.5 * h;
  const double h6 = h / 6.0;
  // ---- drift
  if (needJ)
    for (int c = col; c < N7; c += STEP) drift_cols<NB>(S, c, h2);
  for (int q = col; q < N3; q += STEP) {
    double xq = S.x[q], eq = S.xe[q];
    kahan(xq, eq, h2 * S.v[q]);
    S.x[q] = xq;
    S.xe[q] = eq;
  }
  wsync();
  if (needQ) {
    for (int q = col; q < N3; q += STEP) {
      const int b = q / 3, k = q - 3 * b;
      S.dqdt[7 * b + k] += 0.5 * S.v[q] + h2 * S.dqdt[7 * b + 3 + k];
    }
    for (int r = col; r < N7; r += STEP) S.dqdt_kick[r] = 0.0;
  }
  wsync();
  // ---- kickfast!(s, d, h/6)
  if constexpr (STEP == WAVE) {  // kicked pairs: whole-wave form only
    if (anyKick) kick_build<NB>(S, W, h6, needQ, col);
  }
  wsync();
  if (needQ) {
    double nv[2];
    int nr = 0;
    for (int r = col; r < N7; r += STEP, ++nr) {
      double tmp = 0.0;
      if (anyKick) {
        const int b = r / 7, k = r - 7 * b;
        if (k >= 3 && k < 6) {
          const double* Kr = W.f.dadq + (3 * b + k - 3) * N4;
          for (int d = 0; d < NB; ++d) {
            tmp = fmad(Kr[3 * d + 0], S.dqdt[7 * d + 0], tmp);
            tmp = fmad(Kr[3 * d + 1], S.dqdt[7 * d + 1], tmp);
            tmp = fmad(Kr[3 * d + 2], S.dqdt[7 * d + 2], tmp);
            tmp = fmad(Kr[N3 + d], S.dqdt[7 * d + 6], tmp);
          }
        }
      }
      const double dk = S.dqdt_kick[r] / 6.0;
      nv[nr] = S.dqdt[r] + (dk + tmp);
    }
    wsync();
    nr = 0;
    for (int r = col; r < N7; r += STEP, ++nr) S.dqdt[r] = nv[nr];
  }
  if (needJ)
    for (int c = col; c < N7; c += STEP) {
      if (anyKick) {
        double P[N3];
        load_pos<NB>(S, c, P);
        vel_product_add<NB>(S, W.f.dadq, c, P);
        if constexpr (!FACT) zero_add_rows<NB>(S, c, true, false);
      } else {
        if constexpr (!FACT) zero_add_rows<NB>(S, c, true, true);
      }
    }
  wsync();
}

template <int NB, bool G, bool FACT>
__device__ void ahl21_step_form(Smem<NB, G>& S, Swork<NB>& W, double h, bool needJ, bool needQ, bool anyKick,
                                int lane) {
  step_pre<NB, G, WAVE, FACT>(S, W, h, needJ, needQ, anyKick, lane);
  sweep<NB, G, FACT>(S, W, 0, 0.5 * h, needJ, needQ, lane);  // forward pair sweep (drift first)
  step_mid<NB, G, FACT>(S, W, h, needJ, needQ, anyKick, lane);
  sweep<NB, G, FACT>(S, W, 1, 0.5 * h, needJ, needQ, lane);  // backward pair sweep (drift after)
  step_post<NB, G, WAVE, FACT>(S, W, h, needJ, needQ, anyKick, lane);
}

// The step in the record's form of the Jacobian row update (SysDesc::rowupd):
// two instantiations, so that neither form's code is in the other's loops.
template <int NB, bool G>
__device__ void ahl21_step(Smem<NB, G>& S, Swork<NB>& W, double h, bool needJ, bool needQ, bool anyKick, int lane) {
  if (__builtin_amdgcn_readfirstlane(S.rowupd))
    ahl21_step_form<NB, G, true>(S, W, h, needJ, needQ, anyKick, lane);
  else
    ahl21_step_form<NB, G, false>(S, W, h, needJ, needQ, anyKick, lane);
}

// One step of each of the two systems of a Duo, in lockstep: the stages
// outside the sweeps run for one system, then the other (they share the
// work area), the sweeps for both at once (sweep2).  Each stage is written
// out for each record (fixed LDS addresses): a loop over the two records
// with a selected base took 1.5 % longer, its register pressure higher.
// phisalpha! of both systems of a Duo at once, where the 4n columns and the
// pairs of one system fit half a wave: lanes [0, 32) serve the first
// system's columns and pair factors, [32, 64) the second's, each lane with
// exactly the operations phi_build gives it; each system builds into its own
// work area.  Systems without kicked pairs only.
template <int NB, bool G, bool FACT = false>
__device__ __attribute__((always_inline)) inline void phi_build2(Smem<NB, G>& S0, Smem<NB, G>& S1, Swork<NB>& W0,
                                                                 Swork<NB>& W1, double h0, double h1, bool nQ0, bool nQ1,
                                                                 int lane) {
  static_assert(4 * NB <= 32 && Swork<NB>::NP <= 32, "a system's columns and pairs must fit half a wave");
  const bool sys = lane >= 32;
  phisalpha_cols<NB, G, 32, FACT>(sys ? S1 : S0, sys ? W1 : W0, sys ? h1 : h0, false, sys ? nQ1 : nQ0, lane & 31);
}

template <int NB, bool G, bool PHI2, bool FACT>
__device__ void ahl21_step2(Smem<NB, G>& S0, Smem<NB, G>& S1, Swork<NB>& W, int ch, double h0, double h1,
                            bool nJ0, bool nQ0, bool nJ1, bool nQ1, bool k0, bool k1, int lane) {
  step_pre<NB, G, WAVE, FACT>(S0, W, h0, nJ0, nQ0, k0, lane);
  step_pre<NB, G, WAVE, FACT>(S1, W, h1, nJ1, nQ1, k1, lane);
  sweep2<NB, G, false, FACT>(S0, S1, W.slot, ch, 0, 0.5 * h0, 0.5 * h1, nJ0, nQ0, nJ1, nQ1, lane);
  bool shared = false;
  if constexpr (PHI2) {
    if (!k0 && !k1) {  // the second work area follows the first
      Swork<NB>& W1 = *reinterpret_cast<Swork<NB>*>(reinterpret_cast<unsigned char*>(&W) + sizeof(Swork<NB>));
      phi_build2<NB, G, FACT>(S0, S1, W, W1, h0, h1, nQ0, nQ1, lane);
      wsync();
      step_mid_rest<NB, G, WAVE, FACT>(S0, W, nJ0, nQ0, lane);
      step_mid_rest<NB, G, WAVE, FACT>(S1, W1, nJ1, nQ1, lane);
      shared = true;
    }
  }
  if (!shared) {
    step_mid<NB, G, FACT>(S0, W, h0, nJ0, nQ0, k0, lane);
    step_mid<NB, G, FACT>(S1, W, h1, nJ1, nQ1, k1, lane);
  }
  sweep2<NB, G, false, FACT>(S0, S1, W.slot, ch, 1, 0.5 * h0, 0.5 * h1, nJ0, nQ0, nJ1, nQ1, lane);
  step_post<NB, G, WAVE, FACT>(S0, W, h0, nJ0, nQ0, k0, lane);
  step_post<NB, G, WAVE, FACT>(S1, W, h1, nJ1, nQ1, k1, lane);
}

// One step of both systems of a Duo entirely at once, for body counts whose
// 7n columns fit half a wave (records with their Jacobians in LDS): lanes
// [0, 32) do the first system's work of every stage, [32, 64) the second's,
// each lane with the operations the one-system step gives it; the sweeps'
// Kepler levels and pair Jacobians are shared as in sweep2.  No kicked pairs.
template <int NB, bool FACT>
__device__ void ahl21_step2h(Smem<NB, false>& S0, Smem<NB, false>& S1, Swork<NB>& W0, Swork<NB>& W1, PairSlot* slots,
                             int ch, double h0, double h1, bool nJ0, bool nQ0, bool nJ1, bool nQ1, int lane) {
  static_assert(7 * NB <= 32, "a system's columns must fit half a wave");
  const bool sys = lane >= 32;
  const int col = lane & 31;
  Smem<NB, false>& S = sys ? S1 : S0;
  Swork<NB>& W = sys ? W1 : W0;
  const double h = sys ? h1 : h0;
  const bool nJ = sys ? nJ1 : nJ0, nQ = sys ? nQ1 : nQ0;
  step_pre<NB, false, 32, FACT>(S, W, h, nJ, nQ, false, col);
  sweep2<NB, false, true, FACT>(S0, S1, slots, ch, 0, 0.5 * h0, 0.5 * h1, nJ0, nQ0, nJ1, nQ1, lane);
  phisalpha_cols<NB, false, 32, FACT>(S, W, h, false, nQ, col);
  wsync();
  step_mid_rest<NB, false, 32, FACT>(S, W, nJ, nQ, col);
  sweep2<NB, false, true, FACT>(S0, S1, slots, ch, 1, 0.5 * h0, 0.5 * h1, nJ0, nQ0, nJ1, nQ1, lane);
  step_post<NB, false, 32, FACT>(S, W, h, nJ, nQ, false, col);
}

// ---------------------------------------------------------------------------
// Transit search (timing.jl:3-110) and element gradients (timing.jl:112-138)
// ---------------------------------------------------------------------------
template <int NB, bool G> __device__ inline double gsky(const Smem<NB, G>& S, int i, int j) {
  return (S.x[3 * j] - S.x[3 * i]) * (S.v[3 * j] - S.v[3 * i]) +
         (S.x[3 * j + 1] - S.x[3 * i + 1]) * (S.v[3 * j + 1] - S.v[3 * i + 1]);
}
template <int NB, bool G> __device__ inline double gdot(const Smem<NB, G>& S, int i, int j) {
  return (((S.x[3 * j] - S.x[3 * i]) * (S.dqdt[7 * j + 3] - S.dqdt[7 * i + 3]) +
           (S.x[3 * j + 1] - S.x[3 * i + 1]) * (S.dqdt[7 * j + 4] - S.dqdt[7 * i + 4])) +
          (S.v[3 * j] - S.v[3 * i]) * (S.dqdt[7 * j + 0] - S.dqdt[7 * i + 0])) +
         (S.v[3 * j + 1] - S.v[3 * i + 1]) * (S.dqdt[7 * j + 1] - S.dqdt[7 * i + 1]);
}

template <int NB, bool G> __device__ inline void save_prior(Smem<NB, G>& S, int lane) {
  constexpr int N3 = 3 * NB, N7 = 7 * NB;
  for (int q = lane; q < N3; q += WAVE) {
    S.px[q] = S.x[q]; S.pv[q] = S.v[q]; S.pxe[q] = S.xe[q]; S.pve[q] = S.ve[q];
  }
  for (int q = lane; q < N7; q += WAVE) S.pdqdt[q] = S.dqdt[q];
  if (lane == 0) S.pt = S.t;
  wsync();
}
template <int NB, bool G> __device__ inline void restore_prior(Smem<NB, G>& S, int lane) {
  constexpr int N3 = 3 * NB, N7 = 7 * NB;
  for (int q = lane; q < N3; q += WAVE) {
    S.x[q] = S.px[q]; S.v[q] = S.pv[q]; S.xe[q] = S.pxe[q]; S.ve[q] = S.pve[q];
  }
  for (int q = lane; q < N7; q += WAVE) S.dqdt[q] = S.pdqdt[q];
  if (lane == 0) S.t = S.pt;
  wsync();
}

template <int NB, bool G>
__device__ void findtransit(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, int ti, int jb, double dt0, int lane,
                            double* ws) {
  const auto J_ = jj<NB>(S);
  constexpr int N6 = 6 * NB, N7 = 7 * NB;
  const int n = NB;
  double stmp = 0.0;
  double tt1 = dt0 + 1.0;
  double tt2 = dt0 + 2.0;
  int iter = 0;
  while (true) {
    tt2 = tt1;
    tt1 = dt0;
    restore_prior<NB>(S, lane);
    ahl21_step<NB>(S, W, dt0, false, true, D.any_kick != 0, lane);
    const double gs = gsky<NB>(S, ti, jb);
    const double gd = gdot<NB>(S, ti, jb);
    const double dt = -gs / gd;
    kahan(dt0, stmp, dt);
    iter += 1;
    if (iter >= 20 || dt0 == tt1 || dt0 == tt2) break;
  }
  const int64_t c = S.count[jb] - 1;
  const int nb = D.ntbv;
  double* tt = B.tt + D.off_tt;
  if (D.grad) {
    restore_prior<NB>(S, lane);
    // keep the prior Jacobian: the step below updates J in place
    for (int col = lane; col < N7; col += WAVE)
      for (int r = 0; r < N6; ++r) {
        ws[r * N7 + col] = J_[(unsigned)(r * N7 + col)];
        ws[N6 * N7 + r * N7 + col] = J_[(unsigned)(Smem<NB>::JOFF + r * N7 + col)];
      }
    ahl21_step<NB>(S, W, dt0, true, true, D.any_kick != 0, lane);
    const double gd = gdot<NB>(S, ti, jb);
    const double dx = S.x[3 * jb] - S.x[3 * ti], dy = S.x[3 * jb + 1] - S.x[3 * ti + 1];
    const double dvx = S.v[3 * jb] - S.v[3 * ti], dvy = S.v[3 * jb + 1] - S.v[3 * ti + 1];
    double* dq = B.dq + D.off_dq;
    const int64_t ntt = D.ntt;
    for (int col = lane; col < N7; col += WAVE) {
      const int rj = 6 * jb, ri = 6 * ti;
      const double d0 = -((((J_[(unsigned)(rj * N7 + col)] - J_[(unsigned)(ri * N7 + col)]) * dvx +
                            (J_[(unsigned)((rj + 1) * N7 + col)] - J_[(unsigned)((ri + 1) * N7 + col)]) * dvy) +
                           (J_[(unsigned)((rj + 3) * N7 + col)] - J_[(unsigned)((ri + 3) * N7 + col)]) * dx) +
                          (J_[(unsigned)((rj + 4) * N7 + col)] - J_[(unsigned)((ri + 4) * N7 + col)]) * dy) /
                        gd;
      const int64_t base = (int64_t)jb + n * (c + ntt * (int64_t)col);
      dq[0 + nb * base] = d0;
      if (nb == 3) {
        const double vsky = __builtin_sqrt(dvx * dvx + dvy * dvy);
        const double dvdt = (dvx * (S.dqdt[7 * jb + 3] - S.dqdt[7 * ti + 3]) +
                             dvy * (S.dqdt[7 * jb + 4] - S.dqdt[7 * ti + 4])) / vsky;
        dq[1 + nb * base] = ((J_[(unsigned)((rj + 3) * N7 + col)] - J_[(unsigned)((ri + 3) * N7 + col)]) * dvx +
                             (J_[(unsigned)((rj + 4) * N7 + col)] - J_[(unsigned)((ri + 4) * N7 + col)]) * dvy) / vsky +
                            dvdt * d0;
        dq[2 + nb * base] = 2.0 * ((J_[(unsigned)(rj * N7 + col)] - J_[(unsigned)(ri * N7 + col)]) * dx +
                                   (J_[(unsigned)((rj + 1) * N7 + col)] - J_[(unsigned)((ri + 1) * N7 + col)]) * dy);
      }
    }
    if (lane == 0) {
      tt[0 + nb * (jb + n * c)] = S.t + dt0;
      if (nb == 3) {
        tt[1 + nb * (jb + n * c)] = __builtin_sqrt(dvx * dvx + dvy * dvy);
        tt[2 + nb * (jb + n * c)] = dx * dx + dy * dy;
      }
    }
    for (int col = lane; col < N7; col += WAVE)
      for (int r = 0; r < N6; ++r) {
        J_[(unsigned)(r * N7 + col)] = ws[r * N7 + col];
        J_[(unsigned)(Smem<NB>::JOFF + r * N7 + col)] = ws[N6 * N7 + r * N7 + col];
      }
  } else if (lane == 0) {
    tt[0 + nb * (jb + n * c)] = S.t + dt0;
    if (nb == 3) {
      const double dx = S.x[3 * jb] - S.x[3 * ti], dy = S.x[3 * jb + 1] - S.x[3 * ti + 1];
      const double dvx = S.v[3 * jb] - S.v[3 * ti], dvy = S.v[3 * jb + 1] - S.v[3 * ti + 1];
      tt[1 + nb * (jb + n * c)] = __builtin_sqrt(dvx * dvx + dvy * dvy);
      tt[2 + nb * (jb + n * c)] = dx * dx + dy * dy;
    }
  }
  if (lane == 0) {
    int64_t* ints = B.ints + D.off_int;
    ints[NB + 1] += 1;
    ints[NB + 2] += iter;
  }
  wsync();
}

template <int NB, bool G>
__device__ void detect_transits(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, int lane, double* ws) {
  // set_state!(s_prior, s) (timing.jl:6) is taken lazily, before the first
  // search of this step; a search leaves the state modified and the prior
  // state is restored right after it (timing.jl:19-20), so bodies without a
  // transit see the unmodified state either way.
  bool saved = false;
  const int ti = D.ti;
  for (int i = 0; i < NB; ++i) {
    if (i == ti) continue;
    const double gi = gsky<NB>(S, i, ti);
    const double ri = __builtin_sqrt((S.x[3 * i] * S.x[3 * i] + S.x[3 * i + 1] * S.x[3 * i + 1]) + S.x[3 * i + 2] * S.x[3 * i + 2]);
    if (gi > 0.0 && S.gsave[i] < 0.0 && -S.x[3 * i + 2] > 0.25 * ri && ri < 1e12) {
      const int64_t cnt = S.count[i] + 1;
      wsync();
      if (lane == 0) S.count[i] = cnt;
      wsync();
      if (cnt >= 1 && cnt <= D.ntt) {  // count starts at 0; the lower bound keeps tt/dq indexing in range
        if (!saved) {
          save_prior<NB>(S, lane);
          saved = true;
        }
        const double dt0 = ((-gi) * D.hstep) / (gi - S.gsave[i]);
        findtransit<NB>(S, W, D, B, ti, i, dt0, lane, ws);
        restore_prior<NB>(S, lane);
      }
    }
    wsync();
    if (lane == 0) S.gsave[i] = gi;
    wsync();
  }
}

// calc_dtdelements! (timing.jl:112-138) for one system; jac_init is staged
// transposed in the Jacobian's LDS space (the Jacobian is already stored).
template <int NB, bool G> __device__ void dtdelements(Smem<NB, G>& S, const SysDesc& D, const SysBufs& B, int lane) {
  const auto J_ = jj<NB>(S);
  constexpr int N7 = 7 * NB;
  const int n = NB;
  const auto jt = J_;  // N7*N7 <= 2*N6*N7 doubles
  const double* jinit = B.jac + D.off_jac + 2 * (int64_t)N7 * N7;
  for (int col = lane; col < N7; col += WAVE)
    for (int cp = 0; cp < N7; ++cp) jt[col + N7 * cp] = jinit[cp + (int64_t)N7 * col];
  wsync();
  const double* dq = B.dq + D.off_dq;
  double* del = B.del + D.off_del;
  const int nb = D.ntbv;
  const int64_t ntt = D.ntt;
  for (int b = 0; b < nb; ++b)
    for (int i = 0; i < n; ++i) {
      const int64_t cnt = S.count[i] < ntt ? S.count[i] : ntt;
      for (int64_t jj = 0; jj < cnt; ++jj) {
        for (int col = lane; col < N7; col += WAVE) {
          double acc = 0.0;
          for (int cp = 0; cp < N7; ++cp)
            acc = acc + dq[b + nb * ((int64_t)i + n * (jj + ntt * cp))] * jt[col + N7 * cp];
          del[b + nb * ((int64_t)i + n * (jj + ntt * col))] = acc;
        }
      }
    }
  wsync();
}

// ---------------------------------------------------------------------------
// Load / store of a system between its HBM record and LDS.
// ---------------------------------------------------------------------------
template <int NB, bool G> __device__ void load_system(Smem<NB, G>& S, const SysDesc& D, const SysBufs& B, int lane,
                                              bool loadJ = true) {
  double* J_ = S.JJ;
  if constexpr (G) {
    J_ = B.ws + D.off_ws + 2 * (int64_t)(6 * NB) * (7 * NB);  // after the transit-step copy
    if (lane == 0) S.jg = J_;
  }
  constexpr int N3 = 3 * NB, N6 = 6 * NB, N7 = 7 * NB;
  const double* st = B.state + D.off_state;
  for (int q = lane; q < N3; q += WAVE) {
    S.x[q] = st[off_x(NB) + q];
    S.v[q] = st[off_v(NB) + q];
    S.xe[q] = st[off_xe(NB) + q];
    S.ve[q] = st[off_ve(NB) + q];
  }
  for (int q = lane; q < NB; q += WAVE) {
    S.m[q] = st[off_m(NB) + q];
    S.gsave[q] = st[off_gsave(NB) + q];
    S.count[q] = B.ints[D.off_int + q];
  }
  for (int q = lane; q < N7; q += WAVE) S.dqdt[q] = st[off_dqdt(NB) + q];
  for (int q = lane; q < NB * NB; q += WAVE) S.pair[q] = B.pair[D.off_pair + q];
  if (lane == 0) {
    S.t = st[off_t(NB)];
    S.rowupd = D.rowupd;
  }
  wsync();
  if (lane == 0) {
    constexpr int NPS = Smem<NB>::NP > 0 ? Smem<NB>::NP : 1;
    // a massless pair (gm == 0) leaves jac_ij from the pair before it in
    // place (ahl21.jl:712-713), so systems with one keep the sequential order
    bool seq = false;
    for (int i = 0; i < NB - 1; ++i)
      for (int j = i + 1; j < NB; ++j)
        if (!S.pair[i + NB * j] && NBG_GNEWT * (S.m[i] + S.m[j]) == 0.0) seq = true;
    S.seq = seq ? 1 : 0;
    for (int dir = 0; dir < 2; ++dir) {
      uint8_t pi_[NPS], pj_[NPS], lv[NPS];
      int np = 0;
      if (dir == 0) {
        for (int i = 0; i < NB - 1; ++i)
          for (int j = i + 1; j < NB; ++j)
            if (!S.pair[i + NB * j]) { pi_[np] = (uint8_t)i; pj_[np] = (uint8_t)j; ++np; }
      } else {
        for (int i = NB - 2; i >= 0; --i)
          for (int j = NB - 1; j >= i + 1; --j)
            if (!S.pair[i + NB * j]) { pi_[np] = (uint8_t)i; pj_[np] = (uint8_t)j; ++np; }
      }
      int last[NB];
      for (int b = 0; b < NB; ++b) last[b] = 0;
      int nl = 0;
      for (int p = 0; p < np; ++p) {
        const int a = last[pi_[p]], b = last[pj_[p]];
        const int l = seq ? p + 1 : (a > b ? a : b) + 1;
        lv[p] = (uint8_t)l;
        last[pi_[p]] = l;
        last[pj_[p]] = l;
        nl = l > nl ? l : nl;
      }
      int w = 0;
      for (int l = 1; l <= nl; ++l) {
        int cntl = 0;
        for (int p = 0; p < np; ++p)
          if (lv[p] == l) {
            S.plist[dir][2 * w] = pi_[p];
            S.plist[dir][2 * w + 1] = pj_[p];
            ++w;
            ++cntl;
          }
        S.lvw[dir][l - 1] = (uint8_t)cntl;
      }
      S.nlv[dir] = nl;
    }
  }
  if (D.grad && loadJ) {
    const double* J = B.jac + D.off_jac;
    const double* JE = J + (int64_t)N7 * N7;
    for (int c = lane; c < N7; c += WAVE)
      for (int b = 0; b < NB; ++b)
        for (int k = 0; k < 6; ++k) {
          J_[(unsigned)((6 * b + k) * N7 + c)] = J[(7 * b + k) + (int64_t)N7 * c];
          J_[(unsigned)(Smem<NB>::JOFF + (6 * b + k) * N7 + c)] = JE[(7 * b + k) + (int64_t)N7 * c];
        }
  }
  (void)N6;
  wsync();
}

template <int NB, bool G> __device__ void store_system(Smem<NB, G>& S, const SysDesc& D, const SysBufs& B, int lane) {
  const auto J_ = jj<NB>(S);
  constexpr int N3 = 3 * NB, N7 = 7 * NB;
  wsync();
  double* st = B.state + D.off_state;
  for (int q = lane; q < N3; q += WAVE) {
    st[off_x(NB) + q] = S.x[q];
    st[off_v(NB) + q] = S.v[q];
    st[off_xe(NB) + q] = S.xe[q];
    st[off_ve(NB) + q] = S.ve[q];
  }
  for (int q = lane; q < NB; q += WAVE) {
    st[off_gsave(NB) + q] = S.gsave[q];
    B.ints[D.off_int + q] = S.count[q];
  }
  for (int q = lane; q < N7; q += WAVE) st[off_dqdt(NB) + q] = S.dqdt[q];
  if (lane == 0) st[off_t(NB)] = S.t;
  if (D.grad) {
    double* J = B.jac + D.off_jac;
    double* JE = J + (int64_t)N7 * N7;
    for (int c = lane; c < N7; c += WAVE)
      for (int b = 0; b < NB; ++b)
        for (int k = 0; k < 6; ++k) {
          J[(7 * b + k) + (int64_t)N7 * c] = J_[(unsigned)((6 * b + k) * N7 + c)];
          JE[(7 * b + k) + (int64_t)N7 * c] = J_[(unsigned)(Smem<NB>::JOFF + (6 * b + k) * N7 + c)];
        }
  }
  wsync();
}

// The state before step k of a MODE_CARTESIAN run into its trajectory record
// (o.states[k+1] = deepcopy(s), Outputs.jl:37-40): t, x, v, and by flag
// xerror/verror, jac_step, jac_error (the mass rows, which no step changes:
// unit rows of jac_step, zero rows of jac_error).  Out of line, so that the
// step loop's register allocation is the same as without it.
template <int NB, bool G>
__device__ __attribute__((noinline)) void write_snapshot(const Smem<NB, G>& S, const SysDesc& D, const SysBufs& B,
                                                         int64_t k, int lane) {
  constexpr int N3 = 3 * NB, N7 = 7 * NB;
  const int64_t ns = D.nsteps;
  const int f = D.traj;
  double* base = B.traj + D.off_traj;
  wsync();
  if (lane == 0) base[k] = (k == 0) ? D.t0 : D.t0 + (D.h * (double)k);  // s.t[1] = t0 + (h * i)
  for (int q = lane; q < N3; q += WAVE) {
    base[traj_off_x(NB, ns) + N3 * k + q] = S.x[q];
    base[traj_off_v(NB, ns) + N3 * k + q] = S.v[q];
    if (f & TRAJ_ERR) {
      base[traj_off_xe(NB, ns) + N3 * k + q] = S.xe[q];
      base[traj_off_ve(NB, ns) + N3 * k + q] = S.ve[q];
    }
  }
  if (f & (TRAJ_JAC | TRAJ_JACERR)) {
    const auto J_ = jj<NB>(S);
    const int64_t NN = (int64_t)N7 * N7;
    for (int r = lane; r < N7; r += WAVE) {
      const int b = r / 7, kk = r - 7 * b;
      const unsigned row = (unsigned)((6 * b + (kk < 6 ? kk : 0)) * N7);  // mass rows read nothing
      if (f & TRAJ_JAC) {
        double* o = base + traj_off_jac(NB, ns, f) + NN * k + r;
        for (int c = 0; c < N7; ++c) o[(int64_t)N7 * c] = kk == 6 ? (c == r ? 1.0 : 0.0) : J_[row + c];
      }
      if (f & TRAJ_JACERR) {
        double* o = base + traj_off_jacerr(NB, ns, f) + NN * k + r;
        for (int c = 0; c < N7; ++c) o[(int64_t)N7 * c] = kk == 6 ? 0.0 : J_[Smem<NB>::JOFF + row + c];
      }
    }
  }
}

// check_step (Integrator.jl:249-259): the direction of an integration from
// t0 towards tmax.
__device__ inline double check_step_dev(double t0, double tmax) {
  if (__builtin_fabs(tmax) > __builtin_fabs(t0)) return nbg::sign(tmax);
  if (nbg::sign(tmax) != nbg::sign(t0)) return nbg::sign(tmax);
  return -1.0 * nbg::sign(tmax);
}

// (intr)(s, ts::TransitSnapshot; grad) (Transits.jl:186-211): integrate to
// each time in turn as (intr)(s, time) does (Integrator.jl:159-197: full
// steps of h, then one step onto the time), then b^2 and v_sky of every body
// relative to body 1 (calc_bv) and, with grad, their derivatives as
// calc_dbvdq! forms them (snapshot.jl:17-41, its transit-time terms
// included), one lane per Jacobian column; at the end the element gradients
// of calc_dbvdelements! (snapshot.jl:44-54), which the reference runs with
// or without grad.  Out of line, so the step loop's register allocation
// stays as it is.
template <int NB, bool G>
__device__ __attribute__((noinline)) void run_snapshot(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, int lane) {
  constexpr int N7 = 7 * NB;
  const int n = NB;
  const int64_t nt = D.nsteps;
  double* base = B.traj + D.off_traj;
  double* vsky = base + snap_off_vsky(n, nt);
  double* bsky2 = base + snap_off_bsky2(n, nt);
  double* dq = base + snap_off_dq(n, nt);
  double* del = base + snap_off_del(n, nt);
  const bool grad = D.grad != 0, anyKick = D.any_kick != 0;
  const double hstep = D.hstep;
#pragma unroll 1
  for (int64_t j = 0; j < nt; ++j) {
    const double time = base[j];
    const double t0 = S.t;
    const int64_t ns = (int64_t)__builtin_fabs(__builtin_rint((time - t0) / hstep));  // abs(round(Int64, ...))
    const double hs = hstep * check_step_dev(t0, time);
    const double tmax = t0 + (hs * (double)ns);
    const bool land = (tmax != time);
#pragma unroll 1
    for (int64_t k = 0; k < ns; ++k) ahl21_step<NB>(S, W, hs, grad, grad && k + 1 == ns && !land, anyKick, lane);
    if (land) ahl21_step<NB>(S, W, time - tmax, grad, grad, anyKick, lane);
    if (lane == 0) S.t = time;
    wsync();
    const auto J_ = jj<NB>(S);
    for (int b = 1; b < n; ++b) {
      const double dvx = S.v[3 * b] - S.v[0], dvy = S.v[3 * b + 1] - S.v[1];
      const double dx = S.x[3 * b] - S.x[0], dy = S.x[3 * b + 1] - S.x[1];
      const double vs = __builtin_sqrt(dvx * dvx + dvy * dvy);
      if (lane == 0) {
        vsky[b + n * j] = vs;
        bsky2[b + n * j] = dx * dx + dy * dy;
      }
      if (!grad) continue;
      const double* q = S.dqdt;
      const double gs = dx * dvx + dy * dvy;
      const double gd = ((dx * (q[7 * b + 3] - q[3]) + dy * (q[7 * b + 4] - q[4])) + dvx * (q[7 * b] - q[0])) +
                        dvy * (q[7 * b + 1] - q[1]);
      const double dvdt = (dvx * (q[7 * b + 3] - q[3]) + dvy * (q[7 * b + 4] - q[4])) / vs;
      const double dbdt = 2.0 * gs;
      // rows 7b + k of jac_step (k < 6) are row 6b + k of the compressed layout
      for (int c = lane; c < N7; c += WAVE) {
        auto Jr = [&](int body, int k) { return J_[(unsigned)((6 * body + k) * N7 + c)]; };
        const double Jx = Jr(b, 0) - Jr(0, 0), Jy = Jr(b, 1) - Jr(0, 1);
        const double Jvx = Jr(b, 3) - Jr(0, 3), Jvy = Jr(b, 4) - Jr(0, 4);
        const double dtdq = -(((Jx * dvx + Jy * dvy) + Jvx * dx) + Jvy * dy) / gd;
        const int64_t at = 2 * ((int64_t)b + n * (j + nt * c));
        dq[at] = (Jvx * dvx + Jvy * dvy) / vs + dvdt * dtdq;
        dq[at + 1] = 2.0 * (Jx * dx + Jy * dy) + dbdt * dtdq;
      }
    }
  }
  store_system<NB>(S, D, B, lane);
  if (lane == 0) B.ints[D.off_int + NB] = D.nsteps + 1;
  // calc_dbvdelements!: the sum over p, then q, of dbvdq0[ibv, i, j, q, p] x
  // jac_init[7p + q, 7k + l], one lane per output column 7k + l; jac_init
  // transposed into the Jacobian's store (no longer needed) first
  const auto jt = jj<NB>(S);
  const double* jinit = B.jac + D.off_jac + 2 * (int64_t)N7 * N7;
  for (int col = lane; col < N7; col += WAVE)
    for (int cp = 0; cp < N7; ++cp) jt[col + N7 * cp] = jinit[cp + (int64_t)N7 * col];
  wsync();
  for (int ibv = 0; ibv < 2; ++ibv)
    for (int i = 0; i < n; ++i)
      for (int64_t j = 0; j < nt; ++j)
        for (int col = lane; col < N7; col += WAVE) {
          double acc = 0.0;
          for (int cp = 0; cp < N7; ++cp) acc = acc + dq[ibv + 2 * ((int64_t)i + n * (j + nt * cp))] * jt[col + N7 * cp];
          del[ibv + 2 * ((int64_t)i + n * (j + nt * col))] = acc;
        }
  wsync();
}

// Run (or resume) one system for at most `chunk` steps.  Returns true when
// the system has finished.
template <int NB, bool G>
__device__ void run_system(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, int64_t chunk, int lane) {
  load_system<NB>(S, D, B, lane);
  int64_t* ints = B.ints + D.off_int;
  const int64_t istep0 = ints[NB];
  if (istep0 > D.nsteps) return;  // finished earlier
  if (D.mode == MODE_SNAPSHOT) {
    run_snapshot<NB>(S, W, D, B, lane);
    return;
  }
  const int64_t iend = (chunk > 0 && istep0 + chunk < D.nsteps) ? istep0 + chunk : D.nsteps;
  const bool anyKick = D.any_kick != 0;
  const bool grad = D.grad != 0;
  double* ws = B.ws + D.off_ws;
  if (D.mode == MODE_TRANSITS && istep0 == 0) {
    // gsave[i] = g!(i, ti, x, v) before the loop (Transits.jl:146-149)
    if (lane == 0)
      for (int i = 0; i < NB; ++i)
        if (i != D.ti) S.gsave[i] = gsky<NB>(S, i, D.ti);
    wsync();
  }
#pragma unroll 1
  for (int64_t is = istep0; is < iend; ++is) {
    if (D.mode == MODE_CARTESIAN) write_snapshot<NB>(S, D, B, is, lane);
    // dq/dt of the run's last step (every earlier one is overwritten); only
    // a gradient step forms it: ahl21!(s, h) leaves s.dqdt as it was
    // (ahl21_no_grad.jl:6-18), and the transit search restores it
    // (set_state!, timing.jl:25-26)
    const bool last = (is + 1 == D.nsteps);
    const bool needQ = grad && last && (D.mode == MODE_TRANSITS || D.hlast == 0.0);
    ahl21_step<NB>(S, W, D.h, grad, needQ, anyKick, lane);
    if (D.mode == MODE_TRANSITS) {
      if (lane == 0) S.t = D.t0 + ((double)(is + 1) * D.h);
      wsync();
      detect_transits<NB>(S, W, D, B, lane, ws);
    }
  }
  const bool done = (iend == D.nsteps);
  if (done && D.mode == MODE_STEPS && D.hlast != 0.0) ahl21_step<NB>(S, W, D.hlast, grad, grad, anyKick, lane);
  store_system<NB>(S, D, B, lane);
  if (lane == 0) ints[NB] = done ? D.nsteps + 1 : iend;
  if (done && D.mode == MODE_TRANSITS && grad) dtdelements<NB>(S, D, B, lane);
}

// Run system D with the LDS layout of its body count N (a compile-time
// chain of scalar branches on the uniform D.n).
template <int N, int HI>
__device__ __forceinline__ void dispatch_n(unsigned char* raw, const SysDesc& D, const SysBufs& B, int64_t chunk,
                                           int lane) {
  if constexpr (N <= HI) {
    if (D.n == N) {
      Solo<N>& L = *reinterpret_cast<Solo<N>*>(raw);
      run_system<N>(L.s, L.w, D, B, chunk, lane);
      return;
    }
    dispatch_n<N + 1, HI>(raw, D, B, chunk, lane);
  }
}


// LDS of a launch that may run any body count up to HI.
template <int N> constexpr size_t lds_bytes_upto() {
  if constexpr (N <= 2) return sizeof(Solo<2>);
  else return sizeof(Solo<N>) > lds_bytes_upto<N - 1>() ? sizeof(Solo<N>) : lds_bytes_upto<N - 1>();
}

// ---------------------------------------------------------------------------
// Two systems per wave (a Duo), for transit runs of equal body count.  A
// wave's lanes are mostly idle while it solves the Kepler pairs of a level
// (at most n/2 of them), so one wave steps two systems in lockstep: their
// Kepler drifts and pair Jacobians share the lanes (sweep2), the rest of each
// step runs for one system, then the other.  The two records keep their
// Jacobians in HBM (at every n) and share one work area, which takes the rest
// of the launch's LDS and whose pair slots are split in halves.  A body count
// is paired only when its Duo fits in the LDS of the launch's one-system
// layouts (duo_fits), so the launch keeps its occupancy; other systems, and
// pairs that cannot share a schedule, run one after the other as in nbg_run.
// ---------------------------------------------------------------------------
// LDS of a two-systems-per-wave launch of class HI: the CU's 160 KB over as
// many workgroups as the one-system layouts allow, so occupancy is theirs.
template <int HI> constexpr size_t duo_lds_bytes() {
  constexpr size_t CU = 163840, k = CU / lds_bytes_upto<HI>();
  return (CU / k) / 256 * 256;
}
template <int N, int HI, bool G = true> struct Duo {
  static constexpr size_t REC = sizeof(Smem<N, G>);
  static constexpr size_t FREE = duo_lds_bytes<HI>() > 2 * REC ? duo_lds_bytes<HI>() - 2 * REC : 0;
  // the work area takes what the launch's LDS leaves (at least one Swork<N>)
  static constexpr size_t WBYTES = (FREE > sizeof(Swork<N>) ? FREE : sizeof(Swork<N>)) / 8 * 8;
  Smem<N, G> s0, s1;  // G: Jacobians in HBM
  alignas(8) unsigned char w[WBYTES];
};
template <int N, int HI, bool G = true> constexpr bool duo_fits() {
  return sizeof(Duo<N, HI, G>) <= duo_lds_bytes<HI>();
}
// slots per system in a Duo's shared work area
template <int N, int HI, bool G = true> constexpr int duo_ch() {
  return (int)(Duo<N, HI, G>::WBYTES / sizeof(PairSlot)) / 2;
}
// both systems' force-gradient stages at once (phi_build2): the columns and
// pairs of one system fit half a wave and two work areas fit the Duo's
template <int N, int HI, bool G = true> constexpr bool duo_phi2() {
  return 4 * N <= 32 && Swork<N>::NP <= 32 && 2 * sizeof(Swork<N>) <= Duo<N, HI, G>::WBYTES;
}
// the whole step of both systems at once (ahl21_step2h): every column loop
// of one system (7n columns) fits half a wave, and two records with their
// Jacobians in LDS fit beside two work areas
template <int N, int HI> constexpr bool duo_half() {
  return 7 * N <= 32 && duo_fits<N, HI, false>() && duo_phi2<N, HI, false>() && duo_ch<N, HI, false>() >= N / 2 &&
         duo_ch<N, HI, false>() >= 1;
}

// The two systems' sweeps follow the same level schedule (and neither has a
// massless pair, which keeps the sequential order).  Wave-uniform.
template <int NB, bool G> __device__ inline bool same_schedule(const Smem<NB, G>& a, const Smem<NB, G>& b) {
  bool same = a.seq == 0 && b.seq == 0;
  for (int dir = 0; dir < 2 && same; ++dir) {
    same = a.nlv[dir] == b.nlv[dir];
    int np = 0;
    for (int l = 0; l < a.nlv[dir] && same; ++l) {
      same = a.lvw[dir][l] == b.lvw[dir][l];
      np += a.lvw[dir][l];
    }
    for (int q = 0; q < 2 * np && same; ++q) same = a.plist[dir][q] == b.plist[dir][q];
  }
  return __builtin_amdgcn_readfirstlane(same ? 1 : 0) != 0;
}

// The step loop of run_system for MODE_TRANSITS, steps [is, iend).
template <int NB, bool G>
__device__ void transit_steps(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, int64_t is,
                              int64_t iend, int lane) {
  const bool anyKick = D.any_kick != 0;
  const bool grad = D.grad != 0;
  double* ws = B.ws + D.off_ws;
#pragma unroll 1
  for (; is < iend; ++is) {
    const bool last = (is + 1 == D.nsteps);
    ahl21_step<NB>(S, W, D.h, grad, grad && last, anyKick, lane);
    if (lane == 0) S.t = D.t0 + ((double)(is + 1) * D.h);
    wsync();
    detect_transits<NB>(S, W, D, B, lane, ws);
  }
}

// A system's place in its run for the Duo's step-by-step loop: the next
// step is a main step (ACT_MAIN), a Newton iteration of a transit search
// (ACT_NEWTON) or the search's final gradient step (ACT_JSTEP); ACT_DONE once
// the visit's main steps and their searches are done.  The loop walks
// run_system's control flow -- the main step, detect_transits!'s body loop,
// findtransit!'s iterations -- one step at a time, so that the steps of two
// systems (of any kind) can share one combined step.
enum : int { ACT_MAIN = 0, ACT_NEWTON = 1, ACT_JSTEP = 2, ACT_DONE = 3 };
struct RunCursor {
  int64_t is, iend;  // next main step, end of the visit
  int act;
  int body;          // next body of detect_transits' loop
  int jb;            // body of the search in progress
  int iter;          // its Newton iterations
  int saved;         // s_prior taken after this main step
  double gi, dt0, stmp, tt1, tt2;
};

// detect_transits! (timing.jl:3-29) from body c.body on, as detect_transits
// runs it, until a transit starts a search (act = ACT_NEWTON) or the bodies
// are done (act = the next main step, or ACT_DONE).
template <int NB, bool G>
__device__ void cursor_detect(Smem<NB, G>& S, const SysDesc& D, RunCursor& c, int lane) {
  const int ti = D.ti;
  int act = c.is < c.iend ? ACT_MAIN : ACT_DONE;
  int body = NB;
#pragma unroll 1
  for (int i = c.body; i < NB; ++i) {
    if (i == ti) continue;
    const double gi = gsky<NB>(S, i, ti);
    const double ri = __builtin_sqrt((S.x[3 * i] * S.x[3 * i] + S.x[3 * i + 1] * S.x[3 * i + 1]) + S.x[3 * i + 2] * S.x[3 * i + 2]);
    if (gi > 0.0 && S.gsave[i] < 0.0 && -S.x[3 * i + 2] > 0.25 * ri && ri < 1e12) {
      const int64_t cnt = S.count[i] + 1;
      wsync();
      if (lane == 0) S.count[i] = cnt;
      wsync();
      if (cnt >= 1 && cnt <= D.ntt) {
        if (!c.saved) {
          save_prior<NB>(S, lane);
          c.saved = 1;
        }
        c.dt0 = ((-gi) * D.hstep) / (gi - S.gsave[i]);
        c.gi = gi;
        c.jb = i;
        c.iter = 0;
        c.stmp = 0.0;
        c.tt1 = c.dt0 + 1.0;
        c.tt2 = c.dt0 + 2.0;
        act = ACT_NEWTON;
        body = i;  // gsave[i] is written when the search ends
        break;
      }
    }
    wsync();
    if (lane == 0) S.gsave[i] = gi;
    wsync();
  }
  c.act = __builtin_amdgcn_readfirstlane(act);
  c.body = __builtin_amdgcn_readfirstlane(body);
}

// The step of the cursor's next action: its size and flags, and the state it
// starts from (findtransit! restores s_prior before each iteration, and keeps
// the Jacobian for the final gradient step, timing.jl:45-60).
template <int NB, bool G>
__device__ void cursor_prepare(Smem<NB, G>& S, const SysDesc& D, const SysBufs& B, RunCursor& c, double& h, bool& nJ,
                               bool& nQ, int lane) {
  const bool grad = D.grad != 0;
  if (c.act == ACT_MAIN) {
    h = D.h;
    nJ = grad;
    nQ = grad && c.is + 1 == D.nsteps;  // dq/dt of the run's last step, as run_system forms it
    return;
  }
  restore_prior<NB>(S, lane);
  h = c.dt0;
  nQ = true;
  if (c.act == ACT_NEWTON) {
    c.tt2 = c.tt1;
    c.tt1 = c.dt0;
    nJ = false;
    return;
  }
  constexpr int N6 = 6 * NB, N7 = 7 * NB;
  const auto J_ = jj<NB>(S);
  double* ws = B.ws + D.off_ws;
  for (int col = lane; col < N7; col += WAVE)
    for (int r = 0; r < N6; ++r) {
      ws[r * N7 + col] = J_[(unsigned)(r * N7 + col)];
      ws[N6 * N7 + r * N7 + col] = J_[(unsigned)(Smem<NB, G>::JOFF + r * N7 + col)];
    }
  nJ = true;
}

// The transit's records once its search has converged (findtransit!'s end,
// timing.jl:61-110): tt (and b^2, v_sky), with the gradient step just taken
// the dt/dq rows, then the Jacobian restored; the search counters.
template <int NB, bool G>
__device__ void cursor_record(Smem<NB, G>& S, const SysDesc& D, const SysBufs& B, const RunCursor& c, int lane) {
  constexpr int N6 = 6 * NB, N7 = 7 * NB;
  const int n = NB, ti = D.ti, jb = c.jb, nb = D.ntbv;
  const double dt0 = c.dt0;
  const int64_t cc = S.count[jb] - 1;
  double* tt = B.tt + D.off_tt;
  if (D.grad) {
    const auto J_ = jj<NB>(S);
    double* ws = B.ws + D.off_ws;
    const double gd = gdot<NB>(S, ti, jb);
    const double dx = S.x[3 * jb] - S.x[3 * ti], dy = S.x[3 * jb + 1] - S.x[3 * ti + 1];
    const double dvx = S.v[3 * jb] - S.v[3 * ti], dvy = S.v[3 * jb + 1] - S.v[3 * ti + 1];
    double* dq = B.dq + D.off_dq;
    const int64_t ntt = D.ntt;
    for (int col = lane; col < N7; col += WAVE) {
      const int rj = 6 * jb, ri = 6 * ti;
      const double d0 = -((((J_[(unsigned)(rj * N7 + col)] - J_[(unsigned)(ri * N7 + col)]) * dvx +
                            (J_[(unsigned)((rj + 1) * N7 + col)] - J_[(unsigned)((ri + 1) * N7 + col)]) * dvy) +
                           (J_[(unsigned)((rj + 3) * N7 + col)] - J_[(unsigned)((ri + 3) * N7 + col)]) * dx) +
                          (J_[(unsigned)((rj + 4) * N7 + col)] - J_[(unsigned)((ri + 4) * N7 + col)]) * dy) /
                        gd;
      const int64_t base = (int64_t)jb + n * (cc + ntt * (int64_t)col);
      dq[0 + nb * base] = d0;
      if (nb == 3) {
        const double vsky = __builtin_sqrt(dvx * dvx + dvy * dvy);
        const double dvdt = (dvx * (S.dqdt[7 * jb + 3] - S.dqdt[7 * ti + 3]) +
                             dvy * (S.dqdt[7 * jb + 4] - S.dqdt[7 * ti + 4])) / vsky;
        dq[1 + nb * base] = ((J_[(unsigned)((rj + 3) * N7 + col)] - J_[(unsigned)((ri + 3) * N7 + col)]) * dvx +
                             (J_[(unsigned)((rj + 4) * N7 + col)] - J_[(unsigned)((ri + 4) * N7 + col)]) * dvy) / vsky +
                            dvdt * d0;
        dq[2 + nb * base] = 2.0 * ((J_[(unsigned)(rj * N7 + col)] - J_[(unsigned)(ri * N7 + col)]) * dx +
                                   (J_[(unsigned)((rj + 1) * N7 + col)] - J_[(unsigned)((ri + 1) * N7 + col)]) * dy);
      }
    }
    if (lane == 0) {
      tt[0 + nb * (jb + n * cc)] = S.t + dt0;
      if (nb == 3) {
        tt[1 + nb * (jb + n * cc)] = __builtin_sqrt(dvx * dvx + dvy * dvy);
        tt[2 + nb * (jb + n * cc)] = dx * dx + dy * dy;
      }
    }
    for (int col = lane; col < N7; col += WAVE)
      for (int r = 0; r < N6; ++r) {
        J_[(unsigned)(r * N7 + col)] = ws[r * N7 + col];
        J_[(unsigned)(Smem<NB, G>::JOFF + r * N7 + col)] = ws[N6 * N7 + r * N7 + col];
      }
  } else if (lane == 0) {
    tt[0 + nb * (jb + n * cc)] = S.t + dt0;
    if (nb == 3) {
      const double dx = S.x[3 * jb] - S.x[3 * ti], dy = S.x[3 * jb + 1] - S.x[3 * ti + 1];
      const double dvx = S.v[3 * jb] - S.v[3 * ti], dvy = S.v[3 * jb + 1] - S.v[3 * ti + 1];
      tt[1 + nb * (jb + n * cc)] = __builtin_sqrt(dvx * dvx + dvy * dvy);
      tt[2 + nb * (jb + n * cc)] = dx * dx + dy * dy;
    }
  }
  if (lane == 0) {
    int64_t* ints = B.ints + D.off_int;
    ints[NB + 1] += 1;
    ints[NB + 2] += c.iter;
  }
  wsync();
}

// After the cursor's step: the main step's time and transit tests, or the
// Newton update (timing.jl:49-56) and, once converged, the transit's records,
// s_prior restored and the next bodies' tests.
template <int NB, bool G>
__device__ void cursor_finish(Smem<NB, G>& S, const SysDesc& D, const SysBufs& B, RunCursor& c, int lane) {
  if (c.act == ACT_MAIN) {
    if (lane == 0) S.t = D.t0 + ((double)(c.is + 1) * D.h);
    wsync();
    c.is += 1;
    c.saved = 0;
    c.body = 0;
    cursor_detect<NB>(S, D, c, lane);
    return;
  }
  if (c.act == ACT_NEWTON) {
    const double gs = gsky<NB>(S, D.ti, c.jb);
    const double gd = gdot<NB>(S, D.ti, c.jb);
    const double dt = -gs / gd;
    kahan(c.dt0, c.stmp, dt);
    c.iter += 1;
    if (!(c.iter >= 20 || c.dt0 == c.tt1 || c.dt0 == c.tt2)) return;  // another iteration
    if (D.grad) {
      c.act = ACT_JSTEP;
      return;
    }
  }
  cursor_record<NB>(S, D, B, c, lane);
  restore_prior<NB>(S, lane);
  wsync();
  if (lane == 0) S.gsave[c.jb] = c.gi;
  wsync();
  c.body = c.jb + 1;
  cursor_detect<NB>(S, D, c, lane);
}

// Two MODE_TRANSITS systems of NB bodies (or resume them for at most their
// chunks of steps), step by step: every pass takes the next step of each
// system that has one -- a main step, a Newton iteration of a transit search
// or a search's gradient step -- both in one combined step when both have
// one, so that one system's searches overlap the other's steps; each ends as
// run_system ends it.  Every system's results are exactly its run_system
// results: each step runs the same operations on its own record, and each
// system's sequence of steps, tests and records is run_system's.
template <int NB, int HI, bool G>
__device__ void run_pair(Duo<NB, HI, G>& P, const SysDesc& D0, const SysDesc& D1, const SysBufs& B, int64_t chunk0,
                         int64_t chunk1, int lane) {
  static_assert(duo_ch<NB, HI, G>() >= NB / 2, "a level of pairs must fit a system's slots");
  Smem<NB, G>& S0 = P.s0;
  Smem<NB, G>& S1 = P.s1;
  Swork<NB>& W = *reinterpret_cast<Swork<NB>*>(P.w);
  load_system<NB>(S0, D0, B, lane);
  load_system<NB>(S1, D1, B, lane);
  int64_t* ints0 = B.ints + D0.off_int;
  int64_t* ints1 = B.ints + D1.off_int;
  const int64_t st0 = ints0[NB], st1 = ints1[NB];
  const bool on0 = st0 <= D0.nsteps, on1 = st1 <= D1.nsteps;  // not finished earlier
  const int64_t e0 = !on0 ? st0 : (chunk0 > 0 && st0 + chunk0 < D0.nsteps) ? st0 + chunk0 : D0.nsteps;
  const int64_t e1 = !on1 ? st1 : (chunk1 > 0 && st1 + chunk1 < D1.nsteps) ? st1 + chunk1 : D1.nsteps;
  // gsave[i] = g!(i, ti, x, v) before the loop (Transits.jl:146-149)
  if (lane == 0) {
    if (on0 && st0 == 0)
      for (int i = 0; i < NB; ++i)
        if (i != D0.ti) S0.gsave[i] = gsky<NB>(S0, i, D0.ti);
    if (on1 && st1 == 0)
      for (int i = 0; i < NB; ++i)
        if (i != D1.ti) S1.gsave[i] = gsky<NB>(S1, i, D1.ti);
  }
  wsync();
  if (on0 && on1 && same_schedule<NB>(S0, S1)) {
    // step by step: each pass takes one step of each system with work left
    // -- a main step, a Newton iteration or a transit's gradient step, of any
    // mix -- as one combined step when both have one
    const bool k0 = D0.any_kick != 0, k1 = D1.any_kick != 0;
    const bool fact = D0.rowupd != 0;  // the launch's form of the row update (both records)
    RunCursor c0, c1;
    c0.is = st0; c0.iend = e0; c0.act = st0 < e0 ? ACT_MAIN : ACT_DONE; c0.body = NB; c0.saved = 0;
    c1.is = st1; c1.iend = e1; c1.act = st1 < e1 ? ACT_MAIN : ACT_DONE; c1.body = NB; c1.saved = 0;
#pragma unroll 1
    while (c0.act != ACT_DONE || c1.act != ACT_DONE) {
      const bool a0 = c0.act != ACT_DONE, a1 = c1.act != ACT_DONE;
      double h0 = 0.0, h1 = 0.0;
      bool nJ0 = false, nQ0 = false, nJ1 = false, nQ1 = false;
      if (a0) cursor_prepare<NB>(S0, D0, B, c0, h0, nJ0, nQ0, lane);
      if (a1) cursor_prepare<NB>(S1, D1, B, c1, h1, nJ1, nQ1, lane);
      if (a0 && a1) {
        bool done = false;
        if constexpr (!G) {
          if (!k0 && !k1) {
            Swork<NB>& W1 = *reinterpret_cast<Swork<NB>*>(P.w + sizeof(Swork<NB>));
            if (fact)
              ahl21_step2h<NB, true>(S0, S1, W, W1, W.slot, duo_ch<NB, HI, G>(), h0, h1, nJ0, nQ0, nJ1, nQ1, lane);
            else
              ahl21_step2h<NB, false>(S0, S1, W, W1, W.slot, duo_ch<NB, HI, G>(), h0, h1, nJ0, nQ0, nJ1, nQ1, lane);
            done = true;
          }
        }
        if (!done) {
          constexpr bool PHI2 = duo_phi2<NB, HI, G>();
          if (fact)
            ahl21_step2<NB, G, PHI2, true>(S0, S1, W, duo_ch<NB, HI, G>(), h0, h1, nJ0, nQ0, nJ1, nQ1, k0, k1, lane);
          else
            ahl21_step2<NB, G, PHI2, false>(S0, S1, W, duo_ch<NB, HI, G>(), h0, h1, nJ0, nQ0, nJ1, nQ1, k0, k1, lane);
        }
      } else if (a0) {
        ahl21_step<NB>(S0, W, h0, nJ0, nQ0, k0, lane);
      } else {
        ahl21_step<NB>(S1, W, h1, nJ1, nQ1, k1, lane);
      }
      if (a0) cursor_finish<NB>(S0, D0, B, c0, lane);
      if (a1) cursor_finish<NB>(S1, D1, B, c1, lane);
    }
  } else {
    transit_steps<NB>(S0, W, D0, B, st0, e0, lane);
    transit_steps<NB>(S1, W, D1, B, st1, e1, lane);
  }
  if (on0) {
    store_system<NB>(S0, D0, B, lane);
    if (lane == 0) ints0[NB] = (e0 == D0.nsteps) ? D0.nsteps + 1 : e0;
    if (e0 == D0.nsteps && D0.grad) dtdelements<NB>(S0, D0, B, lane);
  }
  if (on1) {
    store_system<NB>(S1, D1, B, lane);
    if (lane == 0) ints1[NB] = (e1 == D1.nsteps) ? D1.nsteps + 1 : e1;
    if (e1 == D1.nsteps && D1.grad) dtdelements<NB>(S1, D1, B, lane);
  }
}

// Run D0 and D1 (equal body count) as a Duo; false when that count is not
// paired in a launch of class HI.
template <int N, int HI>
__device__ __forceinline__ bool dispatch_pair(unsigned char* raw, const SysDesc& D0, const SysDesc& D1,
                                              const SysBufs& B, int64_t chunk0, int64_t chunk1, int lane) {
  if constexpr (N <= HI) {
    if (D0.n == N) {
      if constexpr (duo_half<N, HI>()) {  // Jacobians in LDS, every stage on half-waves
        run_pair<N, HI, false>(*reinterpret_cast<Duo<N, HI, false>*>(raw), D0, D1, B, chunk0, chunk1, lane);
        return true;
      } else if constexpr (duo_fits<N, HI>()) {
        run_pair<N, HI, true>(*reinterpret_cast<Duo<N, HI, true>*>(raw), D0, D1, B, chunk0, chunk1, lane);
        return true;
      } else {
        return false;
      }
    }
    return dispatch_pair<N + 1, HI>(raw, D0, D1, B, chunk0, chunk1, lane);
  }
  return false;
}

// ---------------------------------------------------------------------------
// Cooperative workgroups: two waves per system, for the longest systems of a
// batch.  Wave 0 integrates; at each detected transit it pushes a job -- the
// body, the Newton start dt0, the transit's index, a snapshot of the state at
// the step's end and of J/Jerr -- into a ring in LDS (its J part into the
// workgroup's HBM slots) and goes on.  Wave 1 runs each job's findtransit on
// its own LDS layout, from the snapshot, exactly as the one-wave kernel runs
// it from its saved prior state, so the results are the same bits; the
// transit search leaves the integrating wave's critical path.  Ring protocol:
// wave 0 owns `tail` and `done`, wave 1 owns `head`; data are published with
// a workgroup-scope release before the counter store and read after an
// acquire.  Every wait has a partner that always makes progress: wave 1
// consumes jobs unconditionally, and wave 0 sets `done` after its last step.
// Each wait is also bounded (CoopBufs::spin_* polls of ~500 cycles; the host
// passes ~1 s for a free slot, which frees within one search, and ~2 min for
// a job, a whole integration without transits): should one ever run out, the
// wave records the failure in `err` (the host turns it into an error) and
// goes on, so that every workgroup still finishes.
// ---------------------------------------------------------------------------
template <int HI> struct TransitRing {
  int head, tail, done, pad;
  int jb[COOP_RQ];
  int64_t cidx[COOP_RQ];
  double dt0[COOP_RQ];
  double st[COOP_RQ][19 * HI + 1];  // x, v, xe, ve (3n each), dqdt (7n), t
};

__device__ __forceinline__ int ring_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void ring_store(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wave 0: hand one transit search to wave 1.
template <int NB, int HI, bool G>
__device__ void push_transit(Smem<NB, G>& S, const SysDesc& D, TransitRing<HI>& R, double* jobJ, int jb, double dt0,
                             int64_t c, int lane, unsigned long long* err, int64_t spin_slot) {
  constexpr int N3 = 3 * NB, N6 = 6 * NB, N7 = 7 * NB;
  const int tail = R.tail;  // only this wave writes it
  long long spins = 0;
  while (tail - ring_load(&R.head) >= COOP_RQ) {
    if (++spins > spin_slot) {
      if (lane == 0) atomicOr(err, 1ull);
      return;
    }
    __builtin_amdgcn_s_sleep(4);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // wave 1 is done with the slot
  const int slot = tail % COOP_RQ;
  double* st = R.st[slot];
  for (int q = lane; q < N3; q += WAVE) {
    st[q] = S.x[q];
    st[N3 + q] = S.v[q];
    st[2 * N3 + q] = S.xe[q];
    st[3 * N3 + q] = S.ve[q];
  }
  for (int q = lane; q < N7; q += WAVE) st[4 * N3 + q] = S.dqdt[q];
  if (lane == 0) {
    st[4 * N3 + N7] = S.t;
    R.jb[slot] = jb;
    R.cidx[slot] = c;
    R.dt0[slot] = dt0;
  }
  if (D.grad) {
    const auto J_ = jj<NB>(S);
    double* dst = jobJ + (int64_t)slot * 2 * N6 * N7;
    for (int col = lane; col < N7; col += WAVE)
      for (int r = 0; r < N6; ++r) {
        dst[r * N7 + col] = J_[(unsigned)(r * N7 + col)];
        dst[N6 * N7 + r * N7 + col] = J_[(unsigned)(Smem<NB>::JOFF + r * N7 + col)];
      }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (lane == 0) ring_store(&R.tail, tail + 1);
  wsync();
}

// Wave 0: the step loop of run_system (MODE_TRANSITS; the whole run, or at
// most `chunk` steps of it) with the transit searches handed to wave 1
// (detect_transits' tests and counting).  The kernel sets `done` after it
// returns, so the searches of the visit are all finished when it ends.
template <int NB, int HI, bool G>
__device__ void coop_integrate(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, TransitRing<HI>& R, double* jobJ,
                               int64_t chunk, int lane, unsigned long long* err, int64_t spin_slot) {
  load_system<NB>(S, D, B, lane);
  int64_t* ints = B.ints + D.off_int;
  const int64_t istep0 = ints[NB];
  const bool anyKick = D.any_kick != 0;
  const bool grad = D.grad != 0;
  if (istep0 <= D.nsteps) {
    const int64_t iend = (chunk > 0 && istep0 + chunk < D.nsteps) ? istep0 + chunk : D.nsteps;
    if (istep0 == 0) {
      if (lane == 0)
        for (int i = 0; i < NB; ++i)
          if (i != D.ti) S.gsave[i] = gsky<NB>(S, i, D.ti);
      wsync();
    }
#pragma unroll 1
    for (int64_t is = istep0; is < iend; ++is) {
      const bool last = (is + 1 == D.nsteps);
      ahl21_step<NB>(S, W, D.h, grad, grad && last, anyKick, lane);  // dq/dt as run_system forms it
      if (lane == 0) S.t = D.t0 + ((double)(is + 1) * D.h);
      wsync();
      const int ti = D.ti;
      for (int i = 0; i < NB; ++i) {
        if (i == ti) continue;
        const double gi = gsky<NB>(S, i, ti);
        const double ri =
            __builtin_sqrt((S.x[3 * i] * S.x[3 * i] + S.x[3 * i + 1] * S.x[3 * i + 1]) + S.x[3 * i + 2] * S.x[3 * i + 2]);
        if (gi > 0.0 && S.gsave[i] < 0.0 && -S.x[3 * i + 2] > 0.25 * ri && ri < 1e12) {
          const int64_t cnt = S.count[i] + 1;
          wsync();
          if (lane == 0) S.count[i] = cnt;
          wsync();
          if (cnt >= 1 && cnt <= D.ntt) {
            const double dt0 = ((-gi) * D.hstep) / (gi - S.gsave[i]);
            push_transit<NB, HI>(S, D, R, jobJ, i, dt0, cnt - 1, lane, err, spin_slot);
          }
        }
        wsync();
        if (lane == 0) S.gsave[i] = gi;
        wsync();
      }
    }
    store_system<NB>(S, D, B, lane);
    if (lane == 0) ints[NB] = (iend == D.nsteps) ? D.nsteps + 1 : iend;
  }
}

// Wave 1: run the pushed transit searches until wave 0 is done and the ring
// is empty.  Its layout holds the system's masses, pair mask and level
// schedule; state and J come from each job.
template <int NB, int HI, bool G>
__device__ void coop_search(Smem<NB, G>& S, Swork<NB>& W, const SysDesc& D, const SysBufs& B, TransitRing<HI>& R, double* jobJ,
                            int lane, unsigned long long* err, int64_t spin_job) {
  constexpr int N3 = 3 * NB, N6 = 6 * NB, N7 = 7 * NB;
  load_system<NB>(S, D, B, lane, false);
  double* ws = B.ws + D.off_ws;
  int head = 0;
#pragma unroll 1
  while (true) {
    int tail = ring_load(&R.tail);
    long long spins = 0;
    bool stop = false;
    while (tail == head) {
      if (ring_load(&R.done)) {
        tail = ring_load(&R.tail);
        break;
      }
      if (++spins > spin_job) {
        if (lane == 0) atomicOr(err, 2ull);
        stop = true;
        break;
      }
      __builtin_amdgcn_s_sleep(8);
      tail = ring_load(&R.tail);
    }
    if (stop || tail == head) break;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int slot = head % COOP_RQ;
    const double* st = R.st[slot];
    for (int q = lane; q < N3; q += WAVE) {
      S.px[q] = st[q];
      S.pv[q] = st[N3 + q];
      S.pxe[q] = st[2 * N3 + q];
      S.pve[q] = st[3 * N3 + q];
    }
    for (int q = lane; q < N7; q += WAVE) S.pdqdt[q] = st[4 * N3 + q];
    const int jb = R.jb[slot];
    const int64_t c = R.cidx[slot];
    const double dt0 = R.dt0[slot];
    if (lane == 0) {
      S.pt = st[4 * N3 + N7];
      S.count[jb] = c + 1;
    }
    double* src = jobJ + (int64_t)slot * 2 * N6 * N7;
    if constexpr (G) {
      if (lane == 0) S.jg = src;  // the snapshot is this search's Jacobian store
    } else if (D.grad) {
      for (int col = lane; col < N7; col += WAVE)
        for (int r = 0; r < N6; ++r) {
          S.JJ[r * N7 + col] = src[r * N7 + col];
          S.JJ[Smem<NB>::JOFF + r * N7 + col] = src[N6 * N7 + r * N7 + col];
        }
    }
    wsync();
    findtransit<NB>(S, W, D, B, D.ti, jb, dt0, lane, ws);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) ring_store(&R.head, head + 1);
    head += 1;
    wsync();
  }
}

template <int N, int HI>
__device__ __forceinline__ void dispatch_coop(unsigned char* raw, int wave, const SysDesc& D, const SysBufs& B,
                                              TransitRing<HI>& R, double* jobJ, int64_t chunk, int lane,
                                              unsigned long long* err, const CoopBufs& C) {
  if constexpr (N <= HI) {
    if (D.n == N) {
      Solo<N>& L = *reinterpret_cast<Solo<N>*>(raw);
      if (wave == 0) coop_integrate<N, HI>(L.s, L.w, D, B, R, jobJ, chunk, lane, err, C.spin_slot);
      else coop_search<N, HI>(L.s, L.w, D, B, R, jobJ, lane, err, C.spin_job);
      return;
    }
    dispatch_coop<N + 1, HI>(raw, wave, D, B, R, jobJ, chunk, lane, err, C);
  }
}

template <int N, int HI>
__device__ __forceinline__ void dispatch_dtdel(unsigned char* raw, const SysDesc& D, const SysBufs& B, int lane) {
  if constexpr (N <= HI) {
    if (D.n == N) {
      dtdelements<N>(reinterpret_cast<Solo<N>*>(raw)->s, D, B, lane);
      return;
    }
    dispatch_dtdel<N + 1, HI>(raw, D, B, lane);
  }
}

// Work accounting at the end of one visit of a system (lane 0): the steps
// it took in this visit plus, with gradients, one unit per transit it
// recorded (its final gradient step) go to units[n]; the transits and the
// Newton iterations of the visit to units[0] and units[1].  A system that
// reached its end in this visit is appended to the done list.
__device__ inline void visit_done(const SysDesc& D, const int64_t* ints, int64_t before, int64_t tr0, int64_t it0,
                                  const SysBufs& B, unsigned long long* units) {
  if (before > D.nsteps) return;  // finished in an earlier visit: nothing ran
  const int64_t after = ints[D.n];
  const int64_t steps = (after > D.nsteps ? D.nsteps : after) - before;
  const int64_t tr = ints[D.n + 1] - tr0, it = ints[D.n + 2] - it0;
  atomicAdd(&units[D.n], (unsigned long long)(steps + (D.grad ? tr : 0)));
  if (tr) atomicAdd(&units[0], (unsigned long long)tr);
  if (it) atomicAdd(&units[1], (unsigned long long)it);
  if (after > D.nsteps && B.done_list) B.done_list[atomicAdd(B.done_count, 1)] = D.tag;
}

// Persistent cooperative kernel: each two-wave workgroup takes systems
// (MODE_TRANSITS, at most HI bodies) from the queues of its group list in
// order and runs each to completion (or for `chunk` steps), as nbg_run does.
// units[0..12] counts work as visit_done does; units[13] collects wait
// failures.
template <int HI>
__global__ __launch_bounds__(2 * WAVE) void nbg_run_coop(GroupList gl, SysBufs B, CoopBufs C, int64_t chunk,
                                                         unsigned long long* __restrict__ units) {
  __shared__ __attribute__((aligned(16))) unsigned char raw[2][lds_bytes_upto<HI>()];
  __shared__ TransitRing<HI> R;
  __shared__ int sys;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int lane = threadIdx.x % WAVE;
  double* jobJ = C.jobs + (int64_t)blockIdx.x * C.stride;
  int g = 0;
#pragma unroll 1
  while (g < gl.ng) {
    if (threadIdx.x == 0) {
      sys = atomicAdd(gl.queue[g], 1);
      R.head = 0;
      R.tail = 0;
      R.done = 0;
    }
    __syncthreads();
    const int idx = __builtin_amdgcn_readfirstlane(sys);
    if (idx >= gl.nsys[g]) {
      ++g;
      __syncthreads();  // thread 0 writes `sys` again
      continue;
    }
    SysDesc D = gl.descs[g][idx];
    D.rowupd = gl.rowupd;
    const int64_t* ints = B.ints + D.off_int;
    const int64_t before = ints[D.n];
    const int64_t tr0 = ints[D.n + 1], it0 = ints[D.n + 2];
    __syncthreads();
    dispatch_coop<2, HI>(raw[wave], wave, D, B, R, jobJ, D.chunk > 0 ? D.chunk : chunk, lane, units + 13, C);
    if (wave == 0) {  // unconditionally, so that wave 1 always leaves its loop
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) ring_store(&R.done, 1);
    }
    __syncthreads();  // every search is written
    const bool finished = before <= D.nsteps && ints[D.n] > D.nsteps;  // during this visit
    if (wave == 0 && D.grad && finished) dispatch_dtdel<2, HI>(raw[0], D, B, lane);
    __syncthreads();
    if (threadIdx.x == 0) visit_done(D, ints, before, tr0, it0, B, units);
    __syncthreads();
  }
}

// Persistent kernel of launch class HI: every workgroup (one wave) takes
// systems from the queues of its group list in order -- its own class first,
// then classes of fewer bodies, each ordered longest first -- and runs each
// to completion (or for `chunk` steps).  LDS is sized for HI bodies; Smem<N>
// grows with N, so every smaller layout fits in it.  `units` (13 entries)
// collects the work of every visit (visit_done): by body count, main steps
// plus one per recorded transit when gradients are on; transits and Newton
// iterations in entries 0 and 1.
template <int HI>
__global__ __launch_bounds__(WAVE) void nbg_run(GroupList gl, SysBufs B, int64_t chunk,
                                                unsigned long long* __restrict__ units) {
  __shared__ __attribute__((aligned(16))) unsigned char raw[lds_bytes_upto<HI>()];
  const int lane = threadIdx.x;
  int g = 0;
  while (g < gl.ng) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(gl.queue[g], 1);
    // readfirstlane (lane 0: every lane is active here) makes the system
    // index, its descriptor and every branch on it wave-uniform (scalar
    // registers, scalar branches).  With a lane shuffle instead, the compiler
    // treats them as divergent and builds exec-masked regions around the
    // per-system code; one such region's exit once dropped a register copy
    // carried to the next system (it ran with an empty exec mask).
    idx = __builtin_amdgcn_readfirstlane(idx);
    if (idx >= gl.nsys[g]) {
      ++g;
      continue;
    }
    SysDesc D = gl.descs[g][idx];
    D.rowupd = gl.rowupd;
    const int64_t* ints = B.ints + D.off_int;
    const int64_t before = ints[D.n];  // step cursor (lane 0 reads and writes it)
    const int64_t tr0 = ints[D.n + 1], it0 = ints[D.n + 2];
    dispatch_n<2, HI>(raw, D, B, D.chunk > 0 ? D.chunk : chunk, lane);
    wsync();
    if (lane == 0) visit_done(D, ints, before, tr0, it0, B, units);
  }
}

// Persistent kernel of launch class HI with two systems per wave: every
// workgroup takes one unit of its queue at a time -- one system, or two
// systems the host paired (equal body count) -- and runs a pair as a Duo when
// it qualifies, else one system after the other as nbg_run does.  LDS and the
// work accounting are nbg_run's.
template <int HI>
__global__ __launch_bounds__(WAVE) void nbg_run_duo(GroupList gl, SysBufs B, int64_t chunk,
                                                    unsigned long long* __restrict__ units) {
  static_assert(duo_lds_bytes<HI>() >= lds_bytes_upto<HI>(), "the one-system layouts must fit");
  __shared__ __attribute__((aligned(16))) unsigned char raw[duo_lds_bytes<HI>()];
  const int lane = threadIdx.x;
  int g = 0;
  while (g < gl.ng) {
    int u = 0;
    if (lane == 0) u = atomicAdd(gl.queue[g], 1);
    u = __builtin_amdgcn_readfirstlane(u);  // wave-uniform, as in nbg_run
    if (u >= gl.nunits[g]) {
      ++g;
      continue;
    }
    const int idx = __builtin_amdgcn_readfirstlane(gl.ustart[g][u]);
    const bool two = __builtin_amdgcn_readfirstlane(gl.ustart[g][u + 1]) - idx == 2;
    SysDesc D0 = gl.descs[g][idx];
    SysDesc D1 = gl.descs[g][two ? idx + 1 : idx];
    D0.rowupd = gl.rowupd;
    D1.rowupd = gl.rowupd;
    const int64_t* ints0 = B.ints + D0.off_int;
    const int64_t* ints1 = B.ints + D1.off_int;
    const int64_t before0 = ints0[D0.n], tr00 = ints0[D0.n + 1], it00 = ints0[D0.n + 2];
    const int64_t before1 = ints1[D1.n], tr01 = ints1[D1.n + 1], it01 = ints1[D1.n + 2];
    const int64_t c0 = D0.chunk > 0 ? D0.chunk : chunk, c1 = D1.chunk > 0 ? D1.chunk : chunk;
    const bool pairable = two && D0.n == D1.n && D0.mode == MODE_TRANSITS && D1.mode == MODE_TRANSITS;
    if (!(pairable && dispatch_pair<2, HI>(raw, D0, D1, B, c0, c1, lane))) {
      dispatch_n<2, HI>(raw, D0, B, c0, lane);
      wsync();
      if (two) dispatch_n<2, HI>(raw, D1, B, c1, lane);
    }
    wsync();
    if (lane == 0) {
      visit_done(D0, ints0, before0, tr00, it00, B, units);
      if (two) visit_done(D1, ints1, before1, tr01, it01, B, units);
    }
  }
}

}  // namespace nbg
