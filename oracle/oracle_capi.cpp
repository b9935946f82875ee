// oracle_capi.cpp -- extern "C" surface of the CPU oracle for the test suite.
// TEST INFRASTRUCTURE ONLY (see nbg_oracle.hpp).  Arrays are Julia
// column-major with 0-based indices; all host memory.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>

#include "nbg_oracle.hpp"

using namespace oracle;

namespace {

template <class T> void load_state(State<T>& s, int n, const double* x, const double* v, const double* xe,
                                   const double* ve, const double* m, const uint8_t* pair, const double* jac,
                                   const double* jace, const double* dqdt, double t) {
  s.resize(n);
  for (int q = 0; q < 3 * n; ++q) {
    s.x[q] = T(x[q]); s.v[q] = T(v[q]);
    s.xerror[q] = xe ? T(xe[q]) : T(0);
    s.verror[q] = ve ? T(ve[q]) : T(0);
  }
  for (int q = 0; q < n; ++q) s.m[q] = T(m[q]);
  if (pair)
    for (int q = 0; q < n * n; ++q) s.pair[q] = pair[q];
  const size_t NN = (size_t)49 * n * n;
  if (jac)
    for (size_t q = 0; q < NN; ++q) s.jac_step[q] = T(jac[q]);
  if (jace)
    for (size_t q = 0; q < NN; ++q) s.jac_error[q] = T(jace[q]);
  if (dqdt)
    for (int q = 0; q < 7 * n; ++q) s.dqdt[q] = T(dqdt[q]);
  s.t = T(t);
}

template <class T> void store_state(const State<T>& s, double* x, double* v, double* xe, double* ve, double* jac,
                                    double* jace, double* dqdt) {
  const int n = s.n;
  for (int q = 0; q < 3 * n; ++q) {
    x[q] = (double)s.x[q]; v[q] = (double)s.v[q];
    if (xe) xe[q] = (double)s.xerror[q];
    if (ve) ve[q] = (double)s.verror[q];
  }
  const size_t NN = (size_t)49 * n * n;
  if (jac)
    for (size_t q = 0; q < NN; ++q) jac[q] = (double)s.jac_step[q];
  if (jace)
    for (size_t q = 0; q < NN; ++q) jace[q] = (double)s.jac_error[q];
  if (dqdt)
    for (int q = 0; q < 7 * n; ++q) dqdt[q] = (double)s.dqdt[q];
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- libm
// fn: 0 sin, 1 cos, 2 exp, 3 sinh, 4 cosh, 5 cbrt.  ulp[i] = |f(x)-F(x)|/ulp(F(x))
// with F evaluated in binary128.
void orc_libm_eval(int fn, int count, const double* xin, double* out, double* ulp) {
  for (int q = 0; q < count; ++q) {
    double x = xin[q], y = 0;
    f128 Y = 0, X = (f128)x;
    switch (fn) {
      case 0: y = nbg::sin(x); Y = sinq(X); break;
      case 1: y = nbg::cos(x); Y = cosq(X); break;
      case 2: y = nbg::exp(x); Y = expq(X); break;
      case 3: y = nbg::sinh(x); Y = sinhq(X); break;
      case 4: y = nbg::cosh(x); Y = coshq(X); break;
      default: y = nbg::cbrt(x); Y = cbrtq(X); break;
    }
    out[q] = y;
    double Yd = (double)Y;
    double u = std::fabs(std::nextafter(Yd, INFINITY) - Yd);
    if (u == 0 || !std::isfinite(u)) u = 4.9406564584124654e-324;
    ulp[q] = std::isfinite(Yd) ? (double)(fabsq((f128)y - Y) / (f128)u) : (y == Yd ? 0.0 : 1e300);
  }
}

// ---------------------------------------------------------------- utils
void orc_utils_eval(int fn, int count, const double* a, const double* b, const double* c, double* out) {
  for (int q = 0; q < count; ++q) {
    switch (fn) {
      case 0: out[q] = cubic1<double>(a[q], b[q], c[q]); break;
      case 1: out[q] = G3<double>(a[q], b[q], c[q]); break;
      case 2: out[q] = H1<double>(a[q], b[q]); break;
      case 3: out[q] = H2<double>(a[q], b[q], c[q]); break;
      case 4: out[q] = H3<double>(a[q], b[q], c[q]); break;
      case 5: out[q] = H5<double>(a[q], b[q], c[q]); break;
      case 6: out[q] = H6<double>(a[q], b[q]); break;
      case 7: out[q] = G3_series<double>(a[q], b[q], c[q]); break;
      case 8: out[q] = H1_series<double>(a[q], b[q]); break;
      case 9: out[q] = H2_series<double>(a[q], b[q], c[q]); break;
      case 10: out[q] = H3_series<double>(a[q], b[q], c[q]); break;
      case 11: out[q] = H5_series<double>(a[q], b[q], c[q]); break;
      case 12: out[q] = H6_series<double>(a[q], b[q]); break;
      default: out[q] = 0;
    }
  }
}
// Same functions in binary128 (inputs promoted), rounded back to double.
void orc_utils_eval_q(int fn, int count, const double* a, const double* b, const double* c, double* out) {
  for (int q = 0; q < count; ++q) {
    f128 A = a[q], B = b[q], C = c ? (f128)c[q] : (f128)0;
    f128 r = 0;
    switch (fn) {
      case 1: r = G3<f128>(A, B, C); break;
      case 2: r = H1<f128>(A, B); break;
      case 3: r = H2<f128>(A, B, C); break;
      case 4: r = H3<f128>(A, B, C); break;
      case 5: r = H5<f128>(A, B, C); break;
      case 6: r = H6<f128>(A, B); break;
      default: r = 0;
    }
    out[q] = (double)r;
  }
}
void orc_comp_sum(int count, double* sv, double* se, const double* add) {
  for (int q = 0; q < count; ++q) comp_sum(sv[q], se[q], add[q]);
}

// ---------------------------------------------------------------- IC stage
int orc_init_nbody(int n, double t0, const double* elements, const double* H, double* x, double* v, double* jac_init) {
  init_nbody<double>(n, t0, elements, H, x, v, jac_init);
  return 0;
}
void orc_hierarchy_nested(int n, double* H) {
  auto h = hierarchy_nested<double>(n);
  std::copy(h.begin(), h.end(), H);
}
void orc_kepler_init(double t, double mass, const double* el, double* x, double* v, double* jac) {
  kepler_init<double>(t, mass, el, x, v, jac);
}
// Central finite differences of kepler_init in binary128: jac_num (7x7) of
// (x, v, mass) with respect to (P, t0, ecosw, esinw, I, Omega, mass).
void orc_fd_kepler_init(double t, double mass, const double* el, double dq, double* jac_num) {
  for (int q = 0; q < 49; ++q) jac_num[q] = 0;
  for (int p = 0; p < 7; ++p) {
    f128 elp[6], elm[6], mp = mass, mm = mass;
    for (int k = 0; k < 6; ++k) elp[k] = elm[k] = el[k];
    f128 step;
    if (p < 6) {
      step = (el[p] != 0.0) ? (f128)dq * fabsq((f128)el[p]) : (f128)dq;
      elp[p] += step; elm[p] -= step;
    } else {
      step = (f128)dq * (f128)mass;
      mp += step; mm -= step;
    }
    f128 xp[3], vp[3], xm[3], vm[3], jj[49];
    kepler_init<f128>((f128)t, mp, elp, xp, vp, jj);
    kepler_init<f128>((f128)t, mm, elm, xm, vm, jj);
    for (int k = 0; k < 3; ++k) {
      jac_num[k + 7 * p] = (double)((xp[k] - xm[k]) / (2 * step));
      jac_num[3 + k + 7 * p] = (double)((vp[k] - vm[k]) / (2 * step));
    }
  }
  jac_num[6 + 7 * 6] = 1.0;
}
// Finite differences of init_nbody (x, v, m w.r.t. elements and masses).
void orc_fd_init_nbody(int n, double t0, const double* elements, const double* H, double dq, double* jac_num) {
  const int sevn = 7 * n;
  std::vector<f128> el((size_t)n * 7), Hq((size_t)n * n);
  for (int q = 0; q < n * 7; ++q) el[q] = elements[q];
  for (int q = 0; q < n * n; ++q) Hq[q] = H[q];
  std::fill(jac_num, jac_num + (size_t)sevn * sevn, 0.0);
  std::vector<f128> xp(3 * n), vp(3 * n), xm(3 * n), vm(3 * n), ji((size_t)sevn * sevn);
  for (int jb = 0; jb < n; ++jb)
    for (int iq = 0; iq < 7; ++iq) {
      // element column ivary (1..6 = P..Omega, 0 = mass) of body jb
      int ivary = (iq == 6) ? 0 : iq + 1;
      f128 base = el[jb + (size_t)n * ivary];
      f128 step = (base != 0) ? (f128)dq * fabsq(base) : (f128)dq;
      std::vector<f128> ep = el, em = el;
      ep[jb + (size_t)n * ivary] += step;
      em[jb + (size_t)n * ivary] -= step;
      init_nbody<f128>(n, (f128)t0, ep.data(), Hq.data(), xp.data(), vp.data(), ji.data());
      init_nbody<f128>(n, (f128)t0, em.data(), Hq.data(), xm.data(), vm.data(), ji.data());
      const int col = 7 * jb + iq;
      for (int b = 0; b < n; ++b)
        for (int k = 0; k < 3; ++k) {
          jac_num[(7 * b + k) + (size_t)sevn * col] = (double)((xp[k + 3 * b] - xm[k + 3 * b]) / (2 * step));
          jac_num[(7 * b + 3 + k) + (size_t)sevn * col] = (double)((vp[k + 3 * b] - vm[k + 3 * b]) / (2 * step));
        }
      if (iq == 6) jac_num[(7 * jb + 6) + (size_t)sevn * col] = 1.0;
    }
}

// ---------------------------------------------------------------- operators
// kepler_driftij_gamma! with Jacobian on a copy of the state (in place x, v).
int orc_kepler_driftij(int n, double* x, double* v, const double* m, int i, int j, double h, int drift_first,
                       double* jac_ij, double* dqdt_ij) {
  State<double> s(n);
  load_state(s, n, x, v, nullptr, nullptr, m, nullptr, nullptr, nullptr, nullptr, 0.0);
  Derivatives<double> d(n);
  kepler_driftij_gamma(s, d, i, j, h, drift_first != 0);
  for (int q = 0; q < 196; ++q) jac_ij[q] = d.jac_ij[q];
  for (int q = 0; q < 14; ++q) dqdt_ij[q] = d.dqdt_ij[q];
  store_state(s, x, v, nullptr, nullptr, nullptr, nullptr, nullptr);
  return 0;
}
// Central differences (binary128) of the no-grad pair map: returns the full
// 14x14 map Jacobian (identity included) and the time derivative.
void orc_fd_kepler_driftij(int n, const double* x, const double* v, const double* m, int i, int j, double h,
                           int drift_first, double dlnq, double* jac_num, double* dqdt_num) {
  auto run = [&](const std::vector<f128>& X, const std::vector<f128>& V, const std::vector<f128>& Mm, f128 hh,
                 f128* out) {
    State<f128> s(n);
    s.x = X; s.v = V; s.m = Mm;
    kepler_driftij_gamma_ng(s, i, j, hh, drift_first != 0);
    for (int k = 0; k < 3; ++k) {
      out[k] = s.X(k, i); out[3 + k] = s.V(k, i); out[6] = s.m[i];
      out[7 + k] = s.X(k, j); out[10 + k] = s.V(k, j); out[13] = s.m[j];
    }
  };
  std::vector<f128> X(3 * n), V(3 * n), Mm(n);
  for (int q = 0; q < 3 * n; ++q) { X[q] = x[q]; V[q] = v[q]; }
  for (int q = 0; q < n; ++q) Mm[q] = m[q];
  f128 op[14], om[14];
  const int bodies[2] = {i, j};
  for (int q = 0; q < 196; ++q) jac_num[q] = 0;
  for (int bb = 0; bb < 2; ++bb) {
    const int b = bodies[bb];
    for (int c = 0; c < 7; ++c) {
      std::vector<f128> Xp = X, Vp = V, Mp = Mm, Xm = X, Vm = V, Mn = Mm;
      f128* basep; f128* basem; f128 val;
      if (c < 3) { basep = &Xp[c + 3 * b]; basem = &Xm[c + 3 * b]; val = X[c + 3 * b]; }
      else if (c < 6) { basep = &Vp[c - 3 + 3 * b]; basem = &Vm[c - 3 + 3 * b]; val = V[c - 3 + 3 * b]; }
      else { basep = &Mp[b]; basem = &Mn[b]; val = Mm[b]; }
      f128 dq = (val != 0) ? (f128)dlnq * val : (f128)dlnq;
      *basep = val + dq; *basem = val - dq;
      run(Xp, Vp, Mp, (f128)h, op);
      run(Xm, Vm, Mn, (f128)h, om);
      for (int r = 0; r < 14; ++r) jac_num[r + 14 * (7 * bb + c)] = (double)((op[r] - om[r]) / (2 * dq));
    }
  }
  f128 hq = h, dq = (f128)dlnq * hq;
  run(X, V, Mm, hq + dq, op);
  run(X, V, Mm, hq - dq, om);
  for (int r = 0; r < 14; ++r) dqdt_num[r] = (double)((op[r] - om[r]) / (2 * dq));
}

// One operator (kickfast 0, phic 1, phisalpha 2) with Jacobian.
// Returns jac_op (7n x 7n, identity NOT included) and dqdt_op (7n); x, v in place.
int orc_operator(int which, int n, double* x, double* v, const double* m, const uint8_t* pair, double h,
                 double* jac_op, double* dqdt_op) {
  State<double> s(n);
  load_state(s, n, x, v, nullptr, nullptr, m, pair, nullptr, nullptr, nullptr, 0.0);
  Derivatives<double> d(n);
  const size_t NN = (size_t)49 * n * n;
  if (which == 0) {
    kickfast(s, d, h);
    for (size_t q = 0; q < NN; ++q) jac_op[q] = d.jac_kick[q];
    for (int q = 0; q < 7 * n; ++q) dqdt_op[q] = d.dqdt_kick[q];
  } else {
    if (which == 1) phic(s, d, h); else phisalpha(s, d, h, 2.0);
    for (size_t q = 0; q < NN; ++q) jac_op[q] = d.jac_phi[q];
    for (int q = 0; q < 7 * n; ++q) dqdt_op[q] = d.dqdt_phi[q];
  }
  store_state(s, x, v, nullptr, nullptr, nullptr, nullptr, nullptr);
  return 0;
}
// Central differences (binary128) of the same no-grad operator.
void orc_fd_operator(int which, int n, const double* x, const double* v, const double* m, const uint8_t* pair,
                     double h, double dlnq, double* jac_num, double* dqdt_num) {
  const int sevn = 7 * n;
  auto run = [&](std::vector<f128>& X, std::vector<f128>& V, std::vector<f128>& Mm, f128 hh, std::vector<f128>& out) {
    State<f128> s(n);
    s.x = X; s.v = V; s.m = Mm;
    for (int q = 0; q < n * n; ++q) s.pair[q] = pair[q];
    if (which == 0) kickfast_ng(s, hh);
    else if (which == 1) phic_ng(s, hh);
    else phisalpha_ng(s, hh, (f128)2);
    for (int b = 0; b < n; ++b) {
      for (int k = 0; k < 3; ++k) { out[7 * b + k] = s.X(k, b); out[7 * b + 3 + k] = s.V(k, b); }
      out[7 * b + 6] = s.m[b];
    }
  };
  std::vector<f128> X(3 * n), V(3 * n), Mm(n), op(sevn), om(sevn);
  for (int q = 0; q < 3 * n; ++q) { X[q] = x[q]; V[q] = v[q]; }
  for (int q = 0; q < n; ++q) Mm[q] = m[q];
  for (int b = 0; b < n; ++b)
    for (int c = 0; c < 7; ++c) {
      std::vector<f128> Xp = X, Vp = V, Mp = Mm, Xm = X, Vm = V, Mn = Mm;
      f128 val = c < 3 ? X[c + 3 * b] : (c < 6 ? V[c - 3 + 3 * b] : Mm[b]);
      f128 dq = (val != 0) ? (f128)dlnq * val : (f128)dlnq;
      if (c < 3) { Xp[c + 3 * b] += dq; Xm[c + 3 * b] -= dq; }
      else if (c < 6) { Vp[c - 3 + 3 * b] += dq; Vm[c - 3 + 3 * b] -= dq; }
      else { Mp[b] += dq; Mn[b] -= dq; }
      run(Xp, Vp, Mp, (f128)h, op);
      run(Xm, Vm, Mn, (f128)h, om);
      for (int r = 0; r < sevn; ++r) jac_num[r + (size_t)sevn * (7 * b + c)] = (double)((op[r] - om[r]) / (2 * dq));
    }
  f128 hq = h, dq = (f128)dlnq * hq;
  run(X, V, Mm, hq + dq, op);
  run(X, V, Mm, hq - dq, om);
  for (int r = 0; r < sevn; ++r) dqdt_num[r] = (double)((op[r] - om[r]) / (2 * dq));
}


// Scalar pair Kepler formulas (jac_delxv_gamma! + compute_jacobian_gamma!):
// out = delxv[6], jac_kepler[48], jac_mass[6], iterations.
void orc_pair_kepler(const double* x0, const double* v0, double k, double h, int drift_first, double* out) {
  State<double> s(2);
  for (int q = 0; q < 3; ++q) { s.x0[q] = x0[q]; s.v0[q] = v0[q]; }
  KParams<double> P = jac_delxv_gamma(s, k, h, drift_first != 0);
  double kep[48] = {0}, jm[6] = {0};
  compute_jacobian_gamma(P, s.x0, s.v0, kep, jm, drift_first != 0);
  for (int q = 0; q < 6; ++q) out[q] = s.delxv[q];
  for (int q = 0; q < 48; ++q) out[6 + q] = kep[q];
  for (int q = 0; q < 6; ++q) out[54 + q] = jm[q];
  out[60] = (double)P.iters;
}

// ---------------------------------------------------------------- stepping
// Test knob: see oracle::g_correction_scale.
void orc_set_correction_scale(double kicked, double kepler) {
  g_correction_scale[0] = kicked;
  g_correction_scale[1] = kepler;
}

int orc_ahl21_step(int n, double h, int grad, double* x, double* v, double* xe, double* ve, const double* m,
                   const uint8_t* pair, double* jac, double* jace, double* dqdt) {
  State<double> s(n);
  load_state(s, n, x, v, xe, ve, m, pair, jac, jace, dqdt, 0.0);
  if (grad) {
    Derivatives<double> d(n);
    ahl21_grad(s, d, h);
  } else {
    ahl21_nograd(s, h);
  }
  store_state(s, x, v, xe, ve, grad ? jac : nullptr, grad ? jace : nullptr, grad ? dqdt : nullptr);
  return 0;
}
// (intr)(s, N; grad) and (intr)(s, time; grad); *t is the state time in/out.
int orc_integrate(int mode, int n, double h, long long nsteps, double time, int grad, double* t, double* x, double* v,
                  double* xe, double* ve, const double* m, const uint8_t* pair, double* jac, double* jace,
                  double* dqdt) {
  State<double> s(n);
  load_state(s, n, x, v, xe, ve, m, pair, jac, jace, dqdt, *t);
  if (mode == 0) integrate_steps(s, h, nsteps, grad != 0);
  else integrate_to(s, h, time, grad != 0);
  store_state(s, x, v, xe, ve, jac, jace, dqdt);
  *t = (double)s.t;
  return 0;
}
// CartesianOutput run; records as arrays over the step index (last axis):
// t[ns], x/v/xe/ve[3n x ns], jac/jace[7n x 7n x ns] (null pointers: skipped).
int orc_cartesian(int n, double h, double tmax, long long nstep, int grad, double* t, double* x, double* v, double* xe,
                  double* ve, const double* m, const uint8_t* pair, double* jac, double* jace, double* dqdt,
                  double* ot, double* ox, double* ov, double* oxe, double* ove, double* ojac, double* ojace) {
  State<double> s(n);
  load_state(s, n, x, v, xe, ve, m, pair, jac, jace, dqdt, *t);
  const size_t n3 = 3 * (size_t)n, nn = 49 * (size_t)n * n;
  cartesian_output(s, h, tmax, nstep, grad != 0, [&](long long k, const State<double>& c) {
    ot[k] = c.t;
    for (size_t q = 0; q < n3; ++q) {
      ox[n3 * k + q] = c.x[q];
      ov[n3 * k + q] = c.v[q];
      if (oxe) oxe[n3 * k + q] = c.xerror[q];
      if (ove) ove[n3 * k + q] = c.verror[q];
    }
    for (size_t q = 0; q < nn; ++q) {
      if (ojac) ojac[nn * k + q] = c.jac_step[q];
      if (ojace) ojace[nn * k + q] = c.jac_error[q];
    }
  });
  store_state(s, x, v, xe, ve, jac, jace, dqdt);
  *t = s.t;
  return 0;
}
// TransitSnapshot run (see transit_snapshot).
int orc_snapshot(int n, double h, int nt, const double* times, int grad, double* t, double* x, double* v, double* xe,
                 double* ve, const double* m, const uint8_t* pair, double* jac, double* jace, double* dqdt,
                 const double* jac_init, double* vsky, double* bsky2, double* dbvdq0, double* dbvdel) {
  State<double> s(n);
  load_state(s, n, x, v, xe, ve, m, pair, jac, jace, dqdt, *t);
  for (size_t q = 0; q < (size_t)49 * n * n; ++q) s.jac_init[q] = jac_init[q];
  transit_snapshot(s, h, times, nt, grad != 0, vsky, bsky2, dbvdq0, dbvdel);
  store_state(s, x, v, xe, ve, jac, jace, dqdt);
  *t = s.t;
  return 0;
}
// Binary128 no-grad integration of N steps (finite-difference pinning).
void orc_integrate_steps_q_fd(int n, double h, long long nsteps, const double* x, const double* v, const double* m,
                              const uint8_t* pair, double dlnq, double* jac_num) {
  const int sevn = 7 * n;
  std::vector<f128> X(3 * n), V(3 * n), Mm(n);
  for (int q = 0; q < 3 * n; ++q) { X[q] = x[q]; V[q] = v[q]; }
  for (int q = 0; q < n; ++q) Mm[q] = m[q];
  auto run = [&](std::vector<f128>& Xi, std::vector<f128>& Vi, std::vector<f128>& Mi, std::vector<f128>& out) {
    State<f128> s(n);
    s.x = Xi; s.v = Vi; s.m = Mi;
    if (pair)
      for (int q = 0; q < n * n; ++q) s.pair[q] = pair[q];
    for (long long k = 0; k < nsteps; ++k) ahl21_nograd(s, (f128)h);
    for (int b = 0; b < n; ++b) {
      for (int k = 0; k < 3; ++k) { out[7 * b + k] = s.X(k, b); out[7 * b + 3 + k] = s.V(k, b); }
      out[7 * b + 6] = s.m[b];
    }
  };
#pragma omp parallel for schedule(dynamic)
  for (int col = 0; col < sevn; ++col) {
    const int b = col / 7, c = col % 7;
    std::vector<f128> Xp = X, Vp = V, Mp = Mm, Xm = X, Vm = V, Mn = Mm, op(sevn), om(sevn);
    f128 val = c < 3 ? X[c + 3 * b] : (c < 6 ? V[c - 3 + 3 * b] : Mm[b]);
    f128 dq = (val != 0) ? (f128)dlnq * val : (f128)dlnq;
    if (c < 3) { Xp[c + 3 * b] += dq; Xm[c + 3 * b] -= dq; }
    else if (c < 6) { Vp[c - 3 + 3 * b] += dq; Vm[c - 3 + 3 * b] -= dq; }
    else { Mp[b] += dq; Mn[b] -= dq; }
    run(Xp, Vp, Mp, op);
    run(Xm, Vm, Mn, om);
    for (int r = 0; r < sevn; ++r) jac_num[r + (size_t)sevn * col] = (double)((op[r] - om[r]) / (2 * dq));
  }
}
// dq/dt of one step by central differences in binary128.
void orc_step_dqdt_q_fd(int n, double h, const double* x, const double* v, const double* m, const uint8_t* pair,
                        double dlnq, double* dqdt_num) {
  const int sevn = 7 * n;
  auto run = [&](f128 hh, std::vector<f128>& out) {
    State<f128> s(n);
    for (int q = 0; q < 3 * n; ++q) { s.x[q] = x[q]; s.v[q] = v[q]; }
    for (int q = 0; q < n; ++q) s.m[q] = m[q];
    if (pair)
      for (int q = 0; q < n * n; ++q) s.pair[q] = pair[q];
    ahl21_nograd(s, hh);
    for (int b = 0; b < n; ++b) {
      for (int k = 0; k < 3; ++k) { out[7 * b + k] = s.X(k, b); out[7 * b + 3 + k] = s.V(k, b); }
      out[7 * b + 6] = s.m[b];
    }
  };
  std::vector<f128> op(sevn), om(sevn);
  f128 hq = h, dq = (f128)dlnq * hq;
  run(hq + dq, op);
  run(hq - dq, om);
  for (int r = 0; r < sevn; ++r) dqdt_num[r] = (double)((op[r] - om[r]) / (2 * dq));
}

// ---------------------------------------------------------------- transits
// (intr)(s, tt; grad) with TransitTiming (ntbv = 1) or TransitParameters (3).
// tt: ntbv x n x ntt; dtdq0 / dtdelements: ntbv x n x ntt x 7 x n; count: n.
int orc_integrate_transits(int n, double h, double tmax, int ti, int grad, int ntbv, double* t, double* x, double* v,
                           double* xe, double* ve, const double* m, const uint8_t* pair, double* jac, double* jace,
                           double* dqdt, const double* jac_init, long long ntt, double* tt, double* dtdq0,
                           double* dtdelements, long long* count) {
  State<double> s(n);
  load_state(s, n, x, v, xe, ve, m, pair, jac, jace, dqdt, *t);
  for (size_t q = 0; q < (size_t)49 * n * n; ++q) s.jac_init[q] = jac_init[q];
  Transits<double> T(n, ntt, ti, ntbv);
  T.s_prior = s;
  integrate_transits(s, T, h, tmax, grad != 0);
  store_state(s, x, v, xe, ve, jac, jace, dqdt);
  *t = s.t;
  std::copy(T.tt.begin(), T.tt.end(), tt);
  if (dtdq0) std::copy(T.dtdq0.begin(), T.dtdq0.end(), dtdq0);
  if (dtdelements) std::copy(T.dtdelements.begin(), T.dtdelements.end(), dtdelements);
  for (int q = 0; q < n; ++q) count[q] = T.count[q];
  return 0;
}

// Central finite differences (binary128 ICs + integration, grad = false) of
// the transit times with respect to the elements (test_transit_timing.jl):
// dtde_num[i, k, iq, jq] (n x ntt x 7 x n); element iq = 0..5 (P..Omega),
// 6 = mass, as in the reference test.  ntbv = 1.
// Central finite differences in binary128 of every transit output with
// respect to the elements, as test_transit_timing.jl and
// test_transit_parameters.jl:46-80 form them: element (jq, ivary) +- dq0
// (the mass is varied last, ivary = 1), both runs without gradients, output
// ntbv x n x ntt x 7 x n (column-major; ntbv = 1 times, 3 times, v_sky and b^2).
void orc_fd_dtbvdelements(int n, double t0, const double* elements, const double* H, double h, double tmax, int ti,
                          double dq0, long long ntt, int ntbv, double* out) {
  const size_t tot = (size_t)ntbv * n * ntt * 7 * n;
  std::fill(out, out + tot, 0.0);
  std::vector<f128> el((size_t)n * 7), Hq((size_t)n * n);
  for (int q = 0; q < 7 * n; ++q) el[q] = elements[q];
  for (int q = 0; q < n * n; ++q) Hq[q] = H[q];
  auto run = [&](std::vector<f128>& E, Transits<f128>& T) {
    State<f128> s(n);
    init_nbody<f128>(n, (f128)t0, E.data(), Hq.data(), s.x.data(), s.v.data(), s.jac_init.data());
    for (int q = 0; q < n; ++q) s.m[q] = E[q];
    s.t = t0;
    T.s_prior = s;
    integrate_transits(s, T, (f128)h, (f128)tmax, false);
  };
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (int jq = 0; jq < n; ++jq)
    for (int iq = 0; iq < 7; ++iq) {
      int ivary = (iq == 6) ? 0 : iq + 1;
      std::vector<f128> ep = el, em = el;
      ep[jq + (size_t)n * ivary] += (f128)dq0;
      em[jq + (size_t)n * ivary] -= (f128)dq0;
      Transits<f128> Tp(n, ntt, ti, ntbv), Tm(n, ntt, ti, ntbv);
      run(ep, Tp);
      run(em, Tm);
      for (int i : Tm.occs)
        for (long long k = 0; k < std::min(Tm.count[i], ntt); ++k)
          for (int b = 0; b < ntbv; ++b)
            out[(size_t)b + (size_t)ntbv * ((size_t)i + (size_t)n * ((size_t)k + (size_t)ntt * ((size_t)iq + 7 * (size_t)jq)))] =
                (double)((Tp.tt[Tp.itt(b, i, k)] - Tm.tt[Tm.itt(b, i, k)]) / (2 * (f128)dq0));
    }
}

void orc_fd_dtdelements(int n, double t0, const double* elements, const double* H, double h, double tmax, int ti,
                        double dq0, long long ntt, double* dtde_num) {
  orc_fd_dtbvdelements(n, t0, elements, H, h, tmax, ti, dq0, ntt, 1, dtde_num);
}

// CPU baseline: integrate `nsys` independent systems (same layout as the
// batched GPU API) with OpenMP, one system per thread.  Returns wall seconds.
double orc_batch_transits(int nsys, const int* nb, const long long* off_state, const long long* off_jac,
                          const long long* off_tt, const long long* off_dq, const double* hs, const double* tmaxs,
                          const long long* ntts, const double* t0s, const double* x, const double* v, const double* m,
                          const double* jac_init, double* tt, double* dtdq0, double* dtdelements, long long* count,
                          int nthreads, long long* steps_done) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  long long total = 0;
  auto t_start = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
  for (int q = 0; q < nsys; ++q) {
    const int n = nb[q];
    State<double> s(n);
    for (int k = 0; k < 3 * n; ++k) { s.x[k] = x[off_state[q] + k]; s.v[k] = v[off_state[q] + k]; }
    for (int k = 0; k < n; ++k) s.m[k] = m[off_state[q] / 3 + k];
    for (size_t k = 0; k < (size_t)49 * n * n; ++k) s.jac_init[k] = jac_init[off_jac[q] + k];
    s.t = t0s[q];
    Transits<double> T(n, ntts[q], 0, 1);
    T.s_prior = s;
    integrate_transits(s, T, hs[q], tmaxs[q], true);
    std::copy(T.tt.begin(), T.tt.end(), tt + off_tt[q]);
    std::copy(T.dtdq0.begin(), T.dtdq0.end(), dtdq0 + off_dq[q]);
    std::copy(T.dtdelements.begin(), T.dtdelements.end(), dtdelements + off_dq[q]);
    long long tr = 0;
    for (int k = 0; k < n; ++k) { count[(long long)off_state[q] / 3 + k] = T.count[k]; tr += std::min(T.count[k], ntts[q]); }
    total += std::llabs(jround<double>(tmaxs[q] / hs[q])) + tr;
  }
  auto t_end = std::chrono::steady_clock::now();
  *steps_done = total;
  return std::chrono::duration<double>(t_end - t_start).count();
}

}  // extern "C"
