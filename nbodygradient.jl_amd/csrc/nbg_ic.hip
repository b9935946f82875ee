// nbg_ic.hip -- initial conditions on the GPU: orbital elements and a
// hierarchy matrix to Cartesian x, v and the element Jacobian jac_init
// (src/ics/init_nbody.jl:13-162 with kepler_init.jl:66-210, kepler.jl:1-41),
// for every system of a batch, written straight into the batch's records.
//
// One workgroup (one wave) per system:
//  1. lane 0 walks the hierarchy (which elements row and which mass sum each
//     orbit uses, init_nbody.jl:58-75);
//  2. lane i solves orbit i (kepler_init with its 7x7 Jacobian);
//  3. the A matrix, its inverse (Gauss-Jordan with partial pivoting, one lane
//     per column of the augmented matrix) and d(A^-1)/dm, one lane per entry;
//  4. jac_init one lane per column, x and v one lane per entry.
// Every sum runs in the host restatement's order.  sin, cos and cbrt come
// from nbg_libm.h, atan2 from the device math library: the results agree
// with a host computation to rounding, not bit for bit.
#include "nbg_libm.h"
#include "nbg_types.h"

namespace nbg {

constexpr int IC_MAX = 12;  // bodies per system
constexpr int IC_WAVE = 64;
constexpr double NBG_PI = 3.141592653589793;                       // pi as Float64
constexpr double IC_GNEWT = 39.4845 / (365.242 * 365.242);         // NbodyGradient.jl:15

struct IcSmem {
  double eps[IC_MAX * IC_MAX], m[IC_MAX], el[IC_MAX * 7];
  double A[IC_MAX * IC_MAX], a[IC_MAX * IC_MAX], b[IC_MAX * IC_MAX];
  double mu[IC_MAX];
  int row[IC_MAX];
  double r[IC_MAX * 3], rd[IC_MAX * 3], j21[IC_MAX * 49];
  double dAdm[IC_MAX * IC_MAX * IC_MAX], dAinv[IC_MAX * IC_MAX * IC_MAX], tmp[IC_MAX * IC_MAX];
  double inv;
  int piv;
};

__device__ inline void ic_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 3x3 column-major products in the reference's matmul order
__device__ inline void mm3(const double* A, const double* B, double* C) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) C[r + 3 * c] = (A[r] * B[3 * c] + A[r + 3] * B[1 + 3 * c]) + A[r + 6] * B[2 + 3 * c];
}
__device__ inline void mv3(const double* A, const double* x, double* y) {
  for (int r = 0; r < 3; ++r) y[r] = (A[r] * x[0] + A[r + 3] * x[1]) + A[r + 6] * x[2];
}
__device__ inline void scale3(const double* A, double s, double* B) {
  for (int q = 0; q < 9; ++q) B[q] = A[q] * s;
}

// Eccentric anomaly from mean anomaly m (kepler.jl:1-41).  Julia's mod(m, 2pi)
// takes the divisor's sign.
__device__ double ekepler_dev(double m, double ecc) {
  if (m == 0.0) return 0.0;
  const double pi2 = 2.0 * NBG_PI;
  double ms = ::fmod(m, pi2);
  if (ms != 0.0 && ((ms < 0.0) != (pi2 < 0.0))) ms += pi2;
  double de0 = (ecc * 0.85) * nbg::sign(ms);
  double de1 = 2.0 * de0, de2 = 3.0 * de0;
  for (int it = 1;; ++it) {
    de2 = de1;
    de1 = de0;
    double s, c;
    nbg::sincos(de0 + ms, &s, &c);
    const double f3 = ecc * c, f2 = ecc * s;
    de0 = (f2 - de1 * f3) / (1.0 - f3);
    if (it >= 20 || de0 == de1 || de0 == de2) break;
  }
  return de0 + m;
}

// One orbit: elements (P, t0, ecosw, esinw, I, Omega) and the mass sum to the
// relative position and velocity at `time`, with their 7x7 Jacobian in
// (P, t0, ecosw, esinw, I, Omega, m) (kepler_init.jl:66-210).
__device__ void kepler_init_dev(double time, double mass, const double* el, double* x, double* v, double* jac) {
  const double PI = NBG_PI, th = 1.0 / 3.0;
  const double period = el[0], t0 = el[1], ecosw = el[2], esinw = el[3], inc = el[4], capom = el[5];
  const double n = (2.0 * PI) / period;
  const double semi = nbg::cbrt(((((IC_GNEWT * mass) * (period * period)) / 4.0) / (PI * PI)));
  const double dsemidp = ((2.0 * th) * semi) / period;
  const double dsemidm = (th * semi) / mass;
  const double ecc = __builtin_sqrt(esinw * esinw + ecosw * ecosw);
  const bool ecc0 = (ecc == 0.0);
  const double deccdecos = ecc0 ? 0.0 : ecosw / ecc;
  const double deccdesin = ecc0 ? 0.0 : esinw / ecc;
  const double sq = __builtin_sqrt(1.0 - ecc * ecc);
  const double den1 = (esinw - ecosw) - ecc;
  double tp;
  if (ecc0) {
    tp = t0 - (3.0 * period) / 4.0;
  } else {
    tp = (t0 - ((sq / n) * ecosw) / (1.0 - esinw)) -
         (2.0 / n) * ::atan2(__builtin_sqrt(1.0 - ecc) * ((esinw + ecosw) + ecc), __builtin_sqrt(1.0 + ecc) * den1);
  }
  const double dtpdp = (tp - t0) / period;
  const double fac = __builtin_sqrt((1.0 - ecc) / (1.0 + ecc));
  const double den2 = 1.0 / (den1 * den1);
  const double theta = (fac * ((esinw + ecosw) + ecc)) / den1;
  const double dthetadecc =
      ((((ecc + ecosw) * (ecc + ecosw)) + ((2.0 * (1.0 - ecc * ecc)) * esinw)) - esinw * esinw) / (sq * (1.0 + ecc)) *
      den2;
  const double dthetadecos = ((2.0 * fac) * esinw) * den2;
  const double dthetadesin = ((-(2.0 * fac)) * (ecosw + ecc)) * den2;
  const double q2 = (2.0 / n) / (1.0 + theta * theta);
  const double dtpdecc = (((ecc / sq) / n) * ecosw) / (1.0 - esinw) - q2 * dthetadecc;
  const double dtpdecos = (dtpdecc * deccdecos - ((sq / n) / (1.0 - esinw))) - q2 * dthetadecos;
  const double dtpdesin =
      (dtpdecc * deccdesin - (((sq / n) * ecosw) / ((1.0 - esinw) * (1.0 - esinw)))) - q2 * dthetadesin;
  const double mean = n * (time - tp);
  const double dmdp = (-mean) / period;
  const double dmdtp = -n;
  const double ekep = ekepler_dev(mean, ecc);
  double sinE, cosE, sinO, cosO, sinI, cosI;
  nbg::sincos(ekep, &sinE, &cosE);
  nbg::sincos(capom, &sinO, &cosO);
  nbg::sincos(inc, &sinI, &cosI);
  const double r = semi * (1.0 - ecc * cosE);
  const double denom = semi / r;
  const double dekepdecos = (sinE * denom) * deccdecos;
  const double dekepdesin = (sinE * denom) * deccdesin;
  const double dekepdm = denom;
  const double cosw = ecc0 ? 1.0 : ecosw / ecc;
  const double sinw = ecc0 ? 0.0 : esinw / ecc;
  const double P1[9] = {cosw, sinw, 0.0, -sinw, cosw, 0.0, 0.0, 0.0, 1.0};
  const double P2[9] = {1.0, 0.0, 0.0, 0.0, cosI, sinI, 0.0, -sinI, cosI};
  const double P3[9] = {cosO, sinO, 0.0, -sinO, cosO, 0.0, 0.0, 0.0, 1.0};
  double P32[9], P321[9], Ps[9], A1[9], A2[9];
  mm3(P3, P2, P32);
  mm3(P32, P1, P321);
  const double xp[3] = {semi * (cosE - ecc), semi * (sq * sinE), semi * 0.0};
  const double vp[3] = {-sinE, sq * cosE, 0.0};
  mv3(P321, xp, x);
  double dxda[3], dxdE[3], dxdecc[3], dxdecos[3], dxdesin[3], dxdinc[3], dxdcom[3], t3[3];
  for (int q = 0; q < 3; ++q) dxda[q] = x[q] / semi;
  scale3(P321, semi, Ps);
  mv3(Ps, vp, dxdE);
  for (int q = 0; q < 9; ++q) Ps[q] = ((-P321[q]) * semi) / ecc;
  {
    const double w[3] = {cosE, sinE / sq, 0.0};
    mv3(Ps, w, dxdecc);
  }
  for (int q = 0; q < 9; ++q) Ps[q] = P32[q] / ecc;
  mv3(Ps, xp, t3);
  for (int q = 0; q < 3; ++q) dxdecos[q] = dxdecc[q] * deccdecos + t3[q];
  {
    const double w[3] = {-xp[1], xp[0], 0.0};
    mv3(Ps, w, t3);
  }
  for (int q = 0; q < 3; ++q) dxdesin[q] = dxdecc[q] * deccdesin + t3[q];
  const double Dinc[9] = {0.0, 0.0, 0.0, 0.0, -sinI, cosI, 0.0, -cosI, -sinI};
  const double Dcom[9] = {-sinO, cosO, 0.0, -cosO, -sinO, 0.0, 0.0, 0.0, 0.0};
  mm3(P3, Dinc, A1);
  mm3(A1, P1, A2);
  mv3(A2, xp, dxdinc);
  mm3(Dcom, P2, A1);
  mm3(A1, P1, A2);
  mv3(A2, xp, dxdcom);
  const double nsd = (n * semi) * denom;
  double vv[3], tv[3];
  for (int q = 0; q < 9; ++q) Ps[q] = ((P321[q] * n) * semi) * denom;
  mv3(Ps, vp, v);
  for (int q = 0; q < 3; ++q) vv[q] = v[q];
  {
    const double w[3] = {-cosE, (-sq) * sinE, 0.0};
    mv3(Ps, w, tv);
  }
  double dvda[3], dvdp[3], dvdE[3], dvdecc[3], dvdecos[3], dvdesin[3], dvdinc[3], dvdcom[3];
  for (int q = 0; q < 3; ++q) {
    dvda[q] = vv[q] / semi;
    dvdp[q] = (-vv[q]) / period;
    dvdE[q] = ((((-vv[q]) * ecc) * sinE) * denom) + tv[q];
  }
  {
    const double w[3] = {0.0, ((-ecc) / sq) * cosE, 0.0};
    mv3(Ps, w, t3);
  }
  for (int q = 0; q < 3; ++q) dvdecc[q] = (((-vv[q]) / ecc) + ((vv[q] * cosE) * denom)) + t3[q];
  for (int q = 0; q < 9; ++q) Ps[q] = (((P32[q] * n) * semi) * denom) / ecc;
  mv3(Ps, vp, t3);
  for (int q = 0; q < 3; ++q) dvdecos[q] = dvdecc[q] * deccdecos + t3[q];
  {
    const double w[3] = {-vp[1], vp[0], 0.0};
    mv3(Ps, w, t3);
  }
  for (int q = 0; q < 3; ++q) dvdesin[q] = dvdecc[q] * deccdesin + t3[q];
  {
    double w[3];
    for (int q = 0; q < 3; ++q) w[q] = vp[q] * nsd;
    mm3(P3, Dinc, A1);
    mm3(A1, P1, A2);
    mv3(A2, w, dvdinc);
    mm3(Dcom, P2, A1);
    mm3(A1, P1, A2);
    mv3(A2, w, dvdcom);
  }
  for (int q = 0; q < 49; ++q) jac[q] = 0.0;
  const double mP = dmdp + dmdtp * dtpdp, kT = dekepdm * dmdtp;
  for (int q = 0; q < 3; ++q) {
    jac[q + 0] = dxda[q] * dsemidp + ((dxdE[q] * dekepdm) * mP);
    jac[q + 7] = ((dxdE[q] * dekepdm) * dmdtp) * 1.0;  // x dtp/dt0 = 1
    jac[q + 14] = ecc0 ? 0.0 : dxdecos[q] + dxdE[q] * (kT * dtpdecos + dekepdecos);
    jac[q + 21] = ecc0 ? 0.0 : dxdesin[q] + dxdE[q] * (kT * dtpdesin + dekepdesin);
    jac[q + 28] = dxdinc[q];
    jac[q + 35] = dxdcom[q];
    jac[q + 42] = dxda[q] * dsemidm;
    jac[3 + q + 0] = (dvdp[q] + dvda[q] * dsemidp) + ((dvdE[q] * dekepdm) * mP);
    jac[3 + q + 7] = ((dvdE[q] * dekepdm) * dmdtp) * 1.0;
    jac[3 + q + 14] = ecc0 ? 0.0 : dvdecos[q] + dvdE[q] * (kT * dtpdecos + dekepdecos);
    jac[3 + q + 21] = ecc0 ? 0.0 : dvdesin[q] + dvdE[q] * (kT * dtpdesin + dekepdesin);
    jac[3 + q + 28] = dvdinc[q];
    jac[3 + q + 35] = dvdcom[q];
    jac[3 + q + 42] = dvda[q] * dsemidm;
  }
  jac[48] = 1.0;
}

__global__ __launch_bounds__(IC_WAVE) void nbg_ic_run(const IcDesc* __restrict__ descs, const double* __restrict__ in,
                                                   double* __restrict__ state, double* __restrict__ jac) {
  __shared__ IcSmem S;
  const int lane = threadIdx.x;
  const IcDesc D = descs[blockIdx.x];
  const int n = D.n, n7 = 7 * n;
  const double* el = in + D.off_el;
  const double* H = el + 7 * n;
  for (int q = lane; q < n * n; q += IC_WAVE) S.eps[q] = H[q];
  for (int q = lane; q < 7 * n; q += IC_WAVE) S.el[q] = el[q];
  for (int q = lane; q < n; q += IC_WAVE) S.m[q] = el[q];
  ic_sync();
  auto EPS = [&](int i, int j) { return S.eps[i + n * j]; };
  // 1. the orbits of the hierarchy (init_nbody.jl:58-75)
  if (lane == 0) {
    int b = 0;
    for (int i = 0; i < n - 1; ++i) {
      double mu = 0.0;
      for (int l = 0; l < n; ++l)
        if (EPS(i, l) != 0.0) mu += S.m[l];
      if (EPS(i, 0) == 0.0) b += 1;
      S.row[i] = i + 1 + b;
      S.mu[i] = mu;
      if (b > 0) b -= 2;
      else if (b < 0) b = 0;
    }
  }
  ic_sync();
  // 2. one lane per orbit
  if (lane < n - 1) {
    double e6[6];
    for (int q = 0; q < 6; ++q) e6[q] = S.el[S.row[lane] + n * (1 + q)];
    double x[3], v[3], J[49];
    kepler_init_dev(D.t0, S.mu[lane], e6, x, v, J);
    for (int q = 0; q < 3; ++q) {
      S.r[lane + n * q] = x[q];
      S.rd[lane + n * q] = v[q];
    }
    for (int q = 0; q < 49; ++q) S.j21[49 * lane + q] = J[q];
  } else if (lane == n - 1) {
    // the last row of the orbit positions and velocities is zero
    // (init_nbody.jl:52-53 allocates them zeroed; n - 1 orbits fill the rest)
    for (int q = 0; q < 3; ++q) {
      S.r[lane + n * q] = 0.0;
      S.rd[lane + n * q] = 0.0;
    }
  }
  // 3. A (init_nbody.jl:176-210) and dA/dm (:120-140)
  auto sigma = [&](int i, int j) {
    double s = 0.0;
    for (int l = 0; l < n; ++l) s += S.m[l] * ((EPS(i, j) == EPS(i, l)) ? 1.0 : 0.0);
    return s;
  };
  for (int q = lane; q < n * n; q += IC_WAVE) {
    const int i = q % n, j = q / n;
    S.A[q] = (EPS(i, j) * S.m[j]) / sigma(i, j);
    S.a[q] = S.A[q];
    S.b[q] = (i == j) ? 1.0 : 0.0;
  }
  for (int q = lane; q < n * n * n; q += IC_WAVE) {
    const int i = q % n, j = (q / n) % n, k = q / (n * n);
    const double eij = EPS(i, j), sm = sigma(i, j);
    const double t1 = ((k == j ? 1.0 : 0.0) * eij) / sm;
    const double t2 = (((eij == EPS(i, k) ? 1.0 : 0.0) * eij) * S.m[j]) / (sm * sm);
    S.dAdm[q] = t1 - t2;
  }
  ic_sync();
  // inverse of A: Gauss-Jordan with partial pivoting; lane q < n updates
  // column q of a, lane n + q column q of b
  for (int c = 0; c < n; ++c) {
    if (lane == 0) {
      int piv = c;
      double best = __builtin_fabs(S.a[c + n * c]);
      for (int r = c + 1; r < n; ++r)
        if (__builtin_fabs(S.a[r + n * c]) > best) {
          best = __builtin_fabs(S.a[r + n * c]);
          piv = r;
        }
      S.piv = piv;
    }
    ic_sync();
    const int piv = S.piv;
    double* M = lane < n ? S.a : S.b;
    const int col = lane < n ? lane : lane - n;
    const bool act = lane < 2 * n;
    if (act && piv != c) {
      const double t = M[c + n * col];
      M[c + n * col] = M[piv + n * col];
      M[piv + n * col] = t;
    }
    ic_sync();
    const double inv = 1.0 / S.a[c + n * c];
    ic_sync();
    if (act) M[c + n * col] *= inv;
    ic_sync();
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = S.a[r + n * c];
      ic_sync();
      if (f != 0.0 && act) M[r + n * col] -= f * M[c + n * col];
      ic_sync();
    }
  }
  // d(A^-1)/dm_k = (-A^-1 dA/dm_k) A^-1, sums in index order
  const double* Ai = S.b;
  for (int k = 0; k < n; ++k) {
    for (int q = lane; q < n * n; q += IC_WAVE) {
      const int r = q % n, c = q / n;
      double s = 0.0;
      for (int p = 0; p < n; ++p) s += (-Ai[r + n * p]) * S.dAdm[p + n * (c + n * k)];
      S.tmp[q] = s;
    }
    ic_sync();
    for (int q = lane; q < n * n; q += IC_WAVE) {
      const int r = q % n, c = q / n;
      double s = 0.0;
      for (int p = 0; p < n; ++p) s += S.tmp[r + n * p] * Ai[p + n * c];
      S.dAinv[r + n * (c + n * k)] = s;
    }
    ic_sync();
  }
  // 4. jac_init (d_dm, init_nbody.jl:141-160), one lane per column l: the
  // orbit Jacobians jk (6n x 7n) enter through A^-1, the mass columns add the
  // d(A^-1)/dm terms
  double* J = jac + D.off_jinit;
  for (int l = lane; l < n7; l += IC_WAVE) {
    const int lb = l / 7, lk = l - 7 * lb;
    for (int ii = 0; ii < n; ++ii) {
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) {
          // jk[6k + j, l]: orbit k's Jacobian in the columns of body k + 1,
          // its mass column in the mass columns of the bodies it holds
          double jkv = 0.0;
          if (k < n - 1) {
            if (lk < 6 && lb == k + 1) jkv = S.j21[49 * k + j + 7 * lk];
            else if (lk == 6 && EPS(k, lb) != 0.0) jkv = S.j21[49 * k + j + 7 * 6];
          }
          s += Ai[ii + n * k] * jkv;
        }
        if (lk == 6) {
          const double* rr = j < 3 ? S.r : S.rd;
          const int jj = j < 3 ? j : j - 3;
          double sx = 0.0;
          for (int q = 0; q < n; ++q) sx += S.dAinv[ii + n * (q + n * lb)] * rr[q + n * jj];
          s += sx;
        }
        J[(7 * ii + j) + (int64_t)n7 * l] = s;
      }
      J[(7 * ii + 6) + (int64_t)n7 * l] = (lk == 6 && lb == ii) ? 1.0 : 0.0;
    }
  }
  // x = (A^-1 r)^T, v = (A^-1 rdot)^T
  double* st = state + D.off_state;
  for (int q = lane; q < 3 * n; q += IC_WAVE) {
    const int ii = q / 3, c = q - 3 * ii;
    double sx = 0.0, sv = 0.0;
    for (int k = 0; k < n; ++k) {
      sx += Ai[ii + n * k] * S.r[k + n * c];
      sv += Ai[ii + n * k] * S.rd[k + n * c];
    }
    st[off_x(n) + q] = sx;
    st[off_v(n) + q] = sv;
  }
}

}  // namespace nbg

// Initial conditions of `nsys` systems: descs (device) describe each one.
extern "C" hipError_t nbg_launch_ic(const void* descs, int nsys, const double* in, double* state, double* jac,
                                    hipStream_t stream) {
  hipLaunchKernelGGL(nbg::nbg_ic_run, dim3(nsys), dim3(nbg::IC_WAVE), 0, stream,
                     static_cast<const nbg::IcDesc*>(descs), in, state, jac);
  return hipGetLastError();
}
