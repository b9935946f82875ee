// nbg_pair.h -- scalar formulas of one pair Kepler drift (universal variable
// gamma) with its analytic 6x8 Jacobian, for gfx950 device code.
//
// Follows NbodyGradient.jl src/utils.jl:103-396 (cubic1, G3, H1, H2, H3, H5,
// H6 and their series) and src/integrator/ahl21/ahl21.jl:766-1139
// (jac_delxv_gamma!, compute_jacobian_gamma!): the same operations in the
// same order (Julia folds n-ary * and + from the left, x^2 and x^3 are
// repeated products, x^4 is Julia's compensated power).  Build with
// -ffp-contract=off.  Compiled for the host too, so the test suite can compare
// these formulas bit for bit with the CPU restatement without a GPU.
#pragma once

#include "nbg_libm.h"

namespace nbg {

constexpr double NBG_GNEWT = 39.4845 / (365.242 * 365.242);  // NbodyGradient.jl:15
constexpr double NBG_THIRD = 1.0 / 3.0;                       // NbodyGradient.jl:16

NBG_HD double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
NBG_HD double dot3(const double* a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

// Julia's Float64 x^4 (Base.Math.pow_body: squaring with compensation).
NBG_HD double cube(double x) { return x * x * x; }

NBG_HD double pow4(double x) {
  const double x1 = x * x;
  const double xn1 = fmad(x, x, -x1);
  const double err = x1 * 2 * xn1;
  const double x2 = x1 * x1;
  const double xn2 = fmad(x1, x1, -x2) + err;
  return x2 + xn2;
}

// cubic1: utils.jl:103-122
NBG_HD double cubic1(double a, double b, double c) {
  double a3 = a * NBG_THIRD;
  double Q = a3 * a3 - b * NBG_THIRD;
  double R = a3 * a3 * a3 + (double)(0.5) * ((-a3) * b + c);
  double R2 = R * R;
  double Q3 = Q * Q * Q;
  // Each lane returns one of the two formulas; a formula is evaluated only
  // when some lane of the wave needs it (on the host: when this lane does).
  const bool three = R2 < Q3;
  double res = 0.0;
  if (NBG_ANY(three)) {
    const double v = -c / b;
    if (three) res = v;
  }
  if (NBG_ANY(!three)) {
    const double w = three ? 1.0 : __builtin_fabs(R) + __builtin_sqrt(R2 - Q3);
    double A = (-nbg::sign(R)) * nbg::cbrt(w);
    double B = (A == (double)(0)) ? (double)(0) : Q / A;
    const double v = (A + B) - a3;
    if (!three) res = v;
  }
  return res;
}

// G3 / G3_series: utils.jl:124-165
NBG_HD double G3_series(double gamma, double beta, double sqb) {
  double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  double term = (double)(1), g3 = (double)(1), g31 = (double)(2) * g3, g32 = (double)(2) * g3;
  int n = 0, iter = 0;
  while (true) {
    g32 = g31; g31 = g3;
    n += 1;
    term *= x2 / double((2 * n + 3) * (2 * n + 2));
    g3 += term;
    iter += 1;
    if (iter >= 100 || g3 == g32 || g3 == g31) break;
  }
  g3 *= ((-x2) * gamma) / (((double)(6) * beta) * sqb);
  return g3;
}
NBG_HD double G3(double gamma, double beta, double sqb) {
  if (gamma < (double)(0.5)) return G3_series(gamma, beta, sqb);
  if (beta >= (double)(0)) return (gamma - nbg::sin(gamma)) / (sqb * beta);
  return (gamma - nbg::sinh(gamma)) / (sqb * beta);
}
// H1: utils.jl:167-209
NBG_HD double H1_series(double gamma, double beta) {
  double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  double term = (double)(1), h1 = (double)(1), h11 = (double)(2) * h1, h12 = (double)(2) * h1;
  int n = 0, iter = 0;
  while (true) {
    h12 = h11; h11 = h1;
    n += 1;
    term *= x2 * (double)(n + 1);
    term /= double((2 * n + 4) * (2 * n + 3) * n);
    h1 += term;
    iter += 1;
    if (iter >= 100 || h1 == h12 || h1 == h11) break;
  }
  h1 *= (x2 * x2) / ((double)(12) * (beta * beta));
  return h1;
}
NBG_HD double H1(double gamma, double beta) {
  if (gamma < (double)(0.5)) return H1_series(gamma, beta);
  if (beta >= (double)(0)) {
    double s = nbg::sin((double)(0.5) * gamma);
    return ((double)(4) * (s * s) - gamma * nbg::sin(gamma)) / (beta * beta);
  }
  double s = nbg::sinh((double)(0.5) * gamma);
  return ((-(double)(4)) * (s * s) + gamma * nbg::sinh(gamma)) / (beta * beta);
}
// H2: utils.jl:211-254
NBG_HD double H2_series(double gamma, double beta, double sqb) {
  double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  double term = (double)(1), h2 = (double)(1), h21 = (double)(2) * h2, h22 = (double)(2) * h2;
  int n = 0, iter = 0;
  while (true) {
    h22 = h21; h21 = h2;
    n += 1;
    term *= x2;
    term /= double((4 * n + 6) * n);
    h2 += term;
    iter += 1;
    if (iter >= 100 || h2 == h22 || h2 == h21) break;
  }
  h2 *= ((-x2) * gamma) / (((double)(3) * beta) * sqb);
  return h2;
}
NBG_HD double H2(double gamma, double beta, double sqb) {
  if (gamma < (double)(0.5)) return H2_series(gamma, beta, sqb);
  if (beta >= (double)(0)) return (nbg::sin(gamma) - gamma * nbg::cos(gamma)) / (sqb * beta);
  return (nbg::sinh(gamma) - gamma * nbg::cosh(gamma)) / (sqb * beta);
}
// H3: utils.jl:256-303
NBG_HD double H3_series(double gamma, double beta, double sqb) {
  double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  double term = (double)(1) / (double)(30);   // convert(double, 1//30)
  double h3 = (double)(1) / (double)(10);     // convert(double, 1//10)
  double h31 = (double)(2) * h3, h32 = (double)(2) * h3;
  int n = 0, iter = 0;
  double four2n = (double)(1) * (double)(4);
  while (true) {
    h32 = h31; h31 = h3;
    n += 1;
    term *= x2;
    term /= double((2 * n + 4) * (2 * n + 5));
    four2n *= (double)(4);
    h3 += term * (four2n - (double)(1));
    iter += 1;
    if (iter >= 100 || h3 == h32 || h3 == h31) break;
  }
  h3 *= ((-(x2 * x2)) * gamma) / (beta * sqb);
  return h3;
}
NBG_HD double H3(double gamma, double beta, double sqb) {
  if (gamma < (double)(0.5)) return H3_series(gamma, beta, sqb);
  if (beta >= (double)(0)) {
    double sg = nbg::sin(gamma);
    return (((double)(4) * sg - sg * nbg::cos(gamma)) - (double)(3) * gamma) / (beta * sqb);
  }
  double sg = nbg::sinh(gamma);
  return (((double)(4) * sg - sg * nbg::cosh(gamma)) - (double)(3) * gamma) / (beta * sqb);
}
// H5: utils.jl:305-349
NBG_HD double H5_series(double gamma, double beta, double sqb) {
  double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  double term = (double)(1) / (double)(60);
  double h5 = term, h51 = (double)(2) * h5, h52 = (double)(2) * h5;
  int n = 0, iter = 0;
  while (true) {
    h52 = h51; h51 = h5;
    n += 1;
    term *= x2 * (double)(n + 1);
    term /= double((2 * n + 5) * (2 * n + 4) * n);
    h5 += term;
    iter += 1;
    if (iter >= 100 || h5 == h52 || h5 == h51) break;
  }
  h5 *= ((-(x2 * x2)) * gamma) / (beta * sqb);
  return h5;
}
NBG_HD double H5(double gamma, double beta, double sqb) {
  if (gamma < (double)(0.5)) return H5_series(gamma, beta, sqb);
  if (beta >= (double)(0))
    return (((double)(3) * nbg::sin(gamma) - (double)(2) * gamma) - gamma * nbg::cos(gamma)) / (beta * sqb);
  return (((double)(3) * nbg::sinh(gamma) - (double)(2) * gamma) - gamma * nbg::cosh(gamma)) / (beta * sqb);
}
// H6: utils.jl:351-396
NBG_HD double H6_series(double gamma, double beta) {
  double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  double term = (double)(1) / (double)(360);  // convert(double, 1//360)
  double h6 = (double)(1) / (double)(40);     // convert(double, 1//40)
  double h61 = (double)(2) * h6, h62 = (double)(2) * h6;
  int n = 0, iter = 0;
  double four2n = (double)(1) * (double)(16);
  while (true) {
    h62 = h61; h61 = h6;
    n += 1;
    term *= x2;
    term /= double((2 * n + 5) * (2 * n + 6));
    four2n *= (double)(4);
    h6 += term * ((four2n - (double)(3 * n)) - (double)(7));
    iter += 1;
    if (iter >= 100 || h6 == h62 || h6 == h61) break;
  }
  h6 *= (-(x2 * x2 * x2)) / (beta * beta);
  return h6;
}
NBG_HD double H6(double gamma, double beta) {
  if (gamma < (double)(0.5)) return H6_series(gamma, beta);
  if (beta >= (double)(0))
    return ((((double)(9) - (double)(8) * nbg::cos(gamma)) - nbg::cos((double)(2) * gamma)) - ((double)(6) * gamma) * nbg::sin(gamma)) /
           ((double)(2) * (beta * beta));
  return ((((double)(9) - (double)(8) * nbg::cosh(gamma)) - nbg::cosh((double)(2) * gamma)) + ((double)(6) * gamma) * nbg::sinh(gamma)) /
         ((double)(2) * (beta * beta));
}

// The series branches (gamma < 0.5) of G3, H1, H2, H3, H5 and H6 evaluated
// together: each series keeps its own recurrence, iteration count and
// stopping test exactly as in its single-function form above, and stops
// updating once its own test fires; running them side by side only lets
// their independent divisions overlap.  MASK selects the series wanted.
enum : int { SER_G3 = 1, SER_H1 = 2, SER_H2 = 4, SER_H3 = 8, SER_H5 = 16, SER_H6 = 32 };
struct SeriesOut {
  double g3, h1, h2, h3, h5, h6;
};
template <int MASK>
NBG_HD SeriesOut series_set(double gamma, double beta, double sqb) {
  const double x2 = (-nbg::sign(beta)) * (gamma * gamma);
  // G3: term, sum, previous two sums, done
  double tg = 1.0, sg = 1.0, sg1 = 2.0 * sg, sg2 = 2.0 * sg;
  double t1 = 1.0, s1 = 1.0, s11 = 2.0 * s1, s12 = 2.0 * s1;
  double t2 = 1.0, s2 = 1.0, s21 = 2.0 * s2, s22 = 2.0 * s2;
  double t3 = (double)(1) / (double)(30), s3 = (double)(1) / (double)(10), s31 = 2.0 * s3, s32 = 2.0 * s3;
  double f3 = (double)(1) * (double)(4);
  double t5 = (double)(1) / (double)(60), s5 = t5, s51 = 2.0 * s5, s52 = 2.0 * s5;
  double t6 = (double)(1) / (double)(360), s6 = (double)(1) / (double)(40), s61 = 2.0 * s6, s62 = 2.0 * s6;
  double f6 = (double)(1) * (double)(16);
  bool dg = !(MASK & SER_G3), d1 = !(MASK & SER_H1), d2 = !(MASK & SER_H2);
  bool d3 = !(MASK & SER_H3), d5 = !(MASK & SER_H5), d6 = !(MASK & SER_H6);
  // n is each running series' iteration count.  A series that has stopped
  // keeps its values (the candidate update is computed and discarded), so
  // the loop can run until every series of every lane of the wave is done.
  for (int n = 1;; ++n) {
    if (MASK & SER_G3) {
      const double t = tg * (x2 / double((2 * n + 3) * (2 * n + 2)));
      const double v = sg + t;
      const bool u = !dg;
      sg2 = u ? sg1 : sg2; sg1 = u ? sg : sg1; tg = u ? t : tg; sg = u ? v : sg;
      dg = dg || (n >= 100 || sg == sg2 || sg == sg1);
    }
    if (MASK & SER_H1) {
      const double t = (t1 * (x2 * (double)(n + 1))) / double((2 * n + 4) * (2 * n + 3) * n);
      const double v = s1 + t;
      const bool u = !d1;
      s12 = u ? s11 : s12; s11 = u ? s1 : s11; t1 = u ? t : t1; s1 = u ? v : s1;
      d1 = d1 || (n >= 100 || s1 == s12 || s1 == s11);
    }
    if (MASK & SER_H2) {
      const double t = (t2 * x2) / double((4 * n + 6) * n);
      const double v = s2 + t;
      const bool u = !d2;
      s22 = u ? s21 : s22; s21 = u ? s2 : s21; t2 = u ? t : t2; s2 = u ? v : s2;
      d2 = d2 || (n >= 100 || s2 == s22 || s2 == s21);
    }
    if (MASK & SER_H3) {
      const double t = (t3 * x2) / double((2 * n + 4) * (2 * n + 5));
      const double f = f3 * (double)(4);
      const double v = s3 + t * (f - (double)(1));
      const bool u = !d3;
      s32 = u ? s31 : s32; s31 = u ? s3 : s31; t3 = u ? t : t3; f3 = u ? f : f3; s3 = u ? v : s3;
      d3 = d3 || (n >= 100 || s3 == s32 || s3 == s31);
    }
    if (MASK & SER_H5) {
      const double t = (t5 * (x2 * (double)(n + 1))) / double((2 * n + 5) * (2 * n + 4) * n);
      const double v = s5 + t;
      const bool u = !d5;
      s52 = u ? s51 : s52; s51 = u ? s5 : s51; t5 = u ? t : t5; s5 = u ? v : s5;
      d5 = d5 || (n >= 100 || s5 == s52 || s5 == s51);
    }
    if (MASK & SER_H6) {
      const double t = (t6 * x2) / double((2 * n + 5) * (2 * n + 6));
      const double f = f6 * (double)(4);
      const double v = s6 + t * ((f - (double)(3 * n)) - (double)(7));
      const bool u = !d6;
      s62 = u ? s61 : s62; s61 = u ? s6 : s61; t6 = u ? t : t6; f6 = u ? f : f6; s6 = u ? v : s6;
      d6 = d6 || (n >= 100 || s6 == s62 || s6 == s61);
    }
    if (NBG_ALL(dg && d1 && d2 && d3 && d5 && d6)) break;
  }
  SeriesOut o;
  o.g3 = sg * (((-x2) * gamma) / (((double)(6) * beta) * sqb));
  o.h1 = s1 * ((x2 * x2) / ((double)(12) * (beta * beta)));
  o.h2 = s2 * (((-x2) * gamma) / (((double)(3) * beta) * sqb));
  o.h3 = s3 * (((-(x2 * x2)) * gamma) / (beta * sqb));
  o.h5 = s5 * (((-(x2 * x2)) * gamma) / (beta * sqb));
  o.h6 = s6 * ((-(x2 * x2 * x2)) / (beta * beta));
  return o;
}
// The functions of MASK at (gamma, beta): the joint series below 0.5, the
// closed forms above (each function's own branch).
template <int MASK>
NBG_HD SeriesOut gfuncs(double gamma, double beta, double sqb) {
  const bool ser = gamma < (double)(0.5);
  SeriesOut o = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (NBG_ANY(ser)) {  // lanes of the closed forms run the series on a dummy argument
    const SeriesOut v = series_set<MASK>(ser ? gamma : 0.25, beta, sqb);
    if (ser) o = v;
  }
  if (NBG_ANY(!ser)) {
    const double g = ser ? 1.0 : gamma;
    SeriesOut v;
    v.g3 = (MASK & SER_G3) ? G3(g, beta, sqb) : 0.0;
    v.h1 = (MASK & SER_H1) ? H1(g, beta) : 0.0;
    v.h2 = (MASK & SER_H2) ? H2(g, beta, sqb) : 0.0;
    v.h3 = (MASK & SER_H3) ? H3(g, beta, sqb) : 0.0;
    v.h5 = (MASK & SER_H5) ? H5(g, beta, sqb) : 0.0;
    v.h6 = (MASK & SER_H6) ? H6(g, beta) : 0.0;
    if (!ser) o = v;
  }
  return o;
}

// (sin, cos)(xx) on lanes with ell, (sinh, e^-xx + sinh)(xx) on the others:
// the two branches of the Kepler solve's G-function evaluation.  Each is
// evaluated only if some lane of the wave needs it, with a harmless argument
// on the lanes that do not, so no lane's inputs push the others off the
// libm fast paths.
NBG_HD void sc_pair(bool ell, double xx, double* sx, double* cx) {
#if defined(__HIP_DEVICE_COMPILE__)
  // Every lane's argument is on the polynomial paths of both branches (sincos
  // without reduction, sinh's series, exp without reduction: what the general
  // calls below would evaluate for it): one straight block, so the four
  // polynomial chains can overlap.
  {
    const double ye = ell ? xx : 0.125, yh = ell ? 0.125 : xx;
    if (__all(nbg::fabsd(ye * TWO_OVER_PI) < 0.5 && nbg::fabsd(ye) >= 0x1p-27 && nbg::fabsd(yh * LOG2E) < 0.5)) {
      const double s1 = nbg::ksin(ye, 0.0), c1 = nbg::kcos(ye, 0.0);
      const double s2 = nbg::sinh_poly(yh);
      const double c2 = nbg::exp_poly(-yh) + s2;
      *sx = ell ? s1 : s2;
      *cx = ell ? c1 : c2;
      return;
    }
  }
#endif
  double s = 0.0, c = 0.0;
  if (NBG_ANY(ell)) {
    double s1, c1;
    nbg::sincos(ell ? xx : 0.125, &s1, &c1);
    if (ell) { s = s1; c = c1; }
  }
  if (NBG_ANY(!ell)) {
    const double y = ell ? 0.125 : xx;
    const double s2 = nbg::sinh(y);
    const double c2 = nbg::exp(-y) + s2;
    if (!ell) { s = s2; c = c2; }
  }
  *sx = s;
  *cx = c;
}

struct KParams {
  double gamma, g0, g1, g2, g3, h1, h2, dfdt, fm1, gmh, dgdtm1, r0, r, r0inv, rinv, k, h, beta, betainv, eta, sqb, zeta;
  int iters;
};

NBG_HD KParams jac_delxv_gamma(const double* x0, const double* v0, double* delxv, double k, double h,
                               bool drift_first) {
  KParams P;
  double rtmp[3];
  rtmp[0] = x0[0] - h * v0[0];
  rtmp[1] = x0[1] - h * v0[1];
  rtmp[2] = x0[2] - h * v0[2];
  double r0 = drift_first ? __builtin_sqrt(dot3(rtmp)) : __builtin_sqrt(dot3(x0));
  double r0inv = (double)(1) / r0;
  double beta0 = ((double)(2) * k) * r0inv - dot3(v0, v0);
  double beta0inv = (double)(1) / beta0;
  double signb = nbg::sign(beta0);
  double sqb = __builtin_sqrt(signb * beta0);
  double zeta = k - r0 * beta0;
  double gamma_guess = (double)(0);
  double eta = drift_first ? dot3(rtmp, v0) : dot3(x0, v0);
  const bool zn = zeta != (double)(0);
  if (NBG_ANY(zn)) {
    double zinv = (double)(6) / zeta;
    const double g = cubic1((((double)(0.5) * eta) * sqb) * zinv, ((r0 * signb) * beta0) * zinv,
                            ((((-h) * signb) * beta0) * sqb) * zinv);
    if (zn) gamma_guess = g;
  }
  if (NBG_ANY(!zn)) {
    double g;
    if (eta != (double)(0)) {
      double reta = r0 / eta;
      double disc = reta * reta + ((double)(2) * h) / eta;
      g = (disc > (double)(0)) ? sqb * ((-reta) + __builtin_sqrt(disc)) : (h * r0inv) * sqb;
    } else {
      g = (h * r0inv) * sqb;
    }
    if (!zn) gamma_guess = g;
  }
  double gamma = gamma_guess;
  double gamma1 = (double)(2) * gamma;
  double gamma2 = (double)(3) * gamma;
  int iter = 0;
  const double c2 = (-(double)(2)) * zeta;
  const double c3 = (((double)(2) * eta) * signb) * sqb;
  const double c4 = ((-sqb) * h) * beta0;
  const double c5 = (((double)(2) * eta) * signb) * sqb;
  const bool ell = beta0 > (double)(0);
  double sx, cx, xx;
  // Newton on gamma (at most 20 iterations, until gamma repeats one of the
  // previous two).  A lane that has stopped keeps its values while the
  // others finish, so the loop runs until the whole wave is done.
  bool run = true;
  do {
    xx = (double)(0.5) * gamma;
    sc_pair(ell, xx, &sx, &cx);
    const double gn = gamma - (((k * gamma + (c2 * sx) * cx) + c3 * (sx * sx)) + c4) /
                                  (((((double)(2) * signb) * zeta) * (sx * sx) + (c5 * sx) * cx) + r0 * beta0);
    if (run) {
      gamma2 = gamma1;
      gamma1 = gamma;
      gamma = gn;
      iter += 1;
      run = !(gamma == gamma2 || gamma == gamma1) && iter < 20;
    }
  } while (NBG_ANY(run));
  for (int j = 0; j < 6; ++j) delxv[j] = (double)(0);
  xx = (double)(0.5) * gamma;
  sc_pair(ell, xx, &sx, &cx);
  double g1bs = (((double)(2) * sx) * cx) / sqb;
  double g2bs = (((double)(2) * signb) * (sx * sx)) * beta0inv;
  double g0bs = (double)(1) - beta0 * g2bs;
  double g3bs, h1 = (double)(0), h2 = (double)(0);
  if (drift_first) {
    g3bs = gfuncs<SER_G3>(gamma, beta0, sqb).g3;
  } else {
    const SeriesOut so = gfuncs<SER_G3 | SER_H1 | SER_H2>(gamma, beta0, sqb);
    g3bs = so.g3;
    h1 = so.h1;
    h2 = so.h2;
  }
  double r = (r0 * g0bs + eta * g1bs) + k * g2bs;
  double rinv = (double)(1) / r;
  double dfdt = (((-k) * g1bs) * rinv) * r0inv;
  double fm1, gmh, dgdtm1;
  if (drift_first) {
    fm1 = ((-k) * r0inv) * g2bs;
    gmh = (k * r0inv) * (h * g2bs - r0 * g3bs);
  } else {
    fm1 = (k * rinv) * (g2bs - (k * r0inv) * h1);
    gmh = (k * rinv) * (r0 * h2 + eta * h1);
  }
  if (drift_first) {
    dgdtm1 = ((k * r0inv) * rinv) * (h * g1bs - r0 * g2bs);
  } else {
    dgdtm1 = ((-k) * rinv) * g2bs;
  }
  for (int j = 0; j < 3; ++j) delxv[j] = fm1 * x0[j] + gmh * v0[j];
  for (int j = 0; j < 3; ++j) delxv[3 + j] = dfdt * x0[j] + dgdtm1 * v0[j];
  P.gamma = gamma; P.g0 = g0bs; P.g1 = g1bs; P.g2 = g2bs; P.g3 = g3bs; P.h1 = h1; P.h2 = h2;
  P.dfdt = dfdt; P.fm1 = fm1; P.gmh = gmh; P.dgdtm1 = dgdtm1; P.r0 = r0; P.r = r;
  P.r0inv = r0inv; P.rinv = rinv; P.k = k; P.h = h; P.beta = beta0; P.betainv = beta0inv;
  P.eta = eta; P.sqb = sqb; P.zeta = zeta; P.iters = iter;
  return P;
}

// --------------------------------------------------------------------------
// compute_jacobian_gamma!: ahl21.jl:896-1139 (debug = false).  Pinned by the
// f128 central finite differences of test_kepler_driftij_gamma.jl:4-282
// (both drift_first, elliptic and hyperbolic).
// --------------------------------------------------------------------------
NBG_HD
void compute_jacobian_gamma(const KParams& P, const double* x0, const double* v0, double* delxv_jac /*6x8*/, double* jac_mass,
                            bool drift_first) {
  const double gamma = P.gamma, g0 = P.g0, g1 = P.g1, g2 = P.g2, g3 = P.g3;
  double h1 = P.h1, h2 = P.h2;
  const double dfdt = P.dfdt, fm1 = P.fm1, gmh = P.gmh, dgdtm1 = P.dgdtm1, r0 = P.r0, r = P.r;
  const double r0inv = P.r0inv, rinv = P.rinv, k = P.k, h = P.h, beta = P.beta, betainv = P.betainv;
  const double eta = P.eta, sqb = P.sqb, zeta = P.zeta;
  const double GN = NBG_GNEWT;
  auto JK = [&](int rr, int cc) -> double& { return delxv_jac[rr + 6 * cc]; };
  double r0inv2 = r0inv * r0inv;
  double r0inv3 = r0inv2 * r0inv;
  double rinv2 = rinv * rinv;
  double rinv3 = rinv2 * rinv;
  double hsq = h * h;
  double ksq = k * k;
  if (drift_first) {
    double d = ((h + eta * g2) + ((double)(2) * k) * g3) * betainv;
    double c1 = d - r0 * g3;
    double c2 = eta * g0 + g1 * zeta;
    double c3 = d * k + g1 * (r0 * r0);
    double c13 = g1 * h - g2 * r0;
    double c9 = (((double)(2) * g2) * h) - (((double)(3) * g3) * r0);
    double c10 = (k * (r0inv2 * r0inv2)) * ((((-g2) * r0) * h + ((k * c9) * betainv)) - ((c3 * c13) * rinv));
    double c24 = r0inv3 * ((r0 * (((double)(2) * k) * r0inv - beta)) * betainv - ((g1 * c3) * rinv) / g2);
    const SeriesOut so = gfuncs<SER_H1 | SER_H2 | SER_H3 | SER_H5 | SER_H6>(gamma, beta, sqb);
    double h6 = so.h6;
    double dfm1dxx = fm1 * c24;
    double dfm1dxv = (-fm1) * (g1 * rinv + h * c24);
    double dfm1dvx = dfm1dxv;
    double dfm1dvv = (fm1 * rinv) * (((-r0) * g2 + ((k * h6) * betainv) / g2) + h * ((double)(2) * g1 + ((h * r) * c24)));
    double dfm1dh = fm1 * ((g1 * rinv) * (((double)(1) / g2 + ((double)(2) * k) * r0inv) - beta) - eta * c24);
    double dfm1dk = fm1 * (((double)(1) / k + ((((g1 * c1) * rinv) * r0inv) / g2)) - (((double)(2) * betainv) * r0inv));
    double h4 = (-so.h1) * beta;
    double h5 = so.h5;
    double dfm1dk2 = r0 * h4 + k * h6;
    double dgmhdxx = c10;
    double dgmhdxv = (((((-g2) * k) * c13) * rinv) * r0inv) - h * c10;
    double dgmhdvx = dgmhdxv;
    double h3 = so.h3;
    double h8 = (-((double)(2) * h3)) + (double)(3) * h5;
    double dgmhdvv = ((((((((double)(2) * g2) * h) * k) * c13) * rinv) * r0inv) + hsq * c10) +
                (((k * betainv) * rinv) * r0inv) *
                    ((((r0 * r0) * h8) - ((((beta * h) * r0) * (g2 * g2)))) + ((h * k + eta * r0) * h6));
    double dgmhdh = ((((g2 * k) * r0inv) + ((((k * c13) * rinv) * r0inv))) +
                (((((g2 * k) * (((double)(2) * k) * r0inv - beta)) * c13) * rinv) * r0inv)) -
               eta * c10;
    double dgmhdk = r0inv * ((((((((k * c1) * c13) * rinv) * r0inv) + g2 * h) - g3 * r0) - (((k * c9) * betainv) * r0inv)));
    double dgmhdk2 = (((((h6 * g3) * ksq) + ((eta * r0) * (h6 + g2 * h4))) + (((r0 * r0) * g0) * h5)) +
                 (((k * eta) * g2) * h6)) +
                (((g1 * h6 + g3 * h4) * k) * r0);
    for (int j = 0; j < 3; ++j) {
      JK(j, j) = fm1;
      JK(j, 3 + j) = gmh;
      for (int i = 0; i < 3; ++i) {
        JK(j, i) += (dfm1dxx * x0[i] + dfm1dxv * v0[i]) * x0[j] + (dgmhdxx * x0[i] + dgmhdxv * v0[i]) * v0[j];
        JK(j, 3 + i) += (dfm1dvx * x0[i] + dfm1dvv * v0[i]) * x0[j] + (dgmhdvx * x0[i] + dgmhdvv * v0[i]) * v0[j];
      }
      JK(j, 6) = dfm1dk * x0[j] + dgmhdk * v0[j];
      JK(j, 7) = dfm1dh * x0[j] + dgmhdh * v0[j];
      double gr = GN * r0inv;
      jac_mass[j] = (((gr * gr) * betainv) * rinv) * (dfm1dk2 * x0[j] - dgmhdk2 * v0[j]);
    }
    double c5 = ((r0 - k * g2) * rinv) / g1;
    double c6 = (r0 * g0 - k * g2) * betainv;
    double c7 = g2 * ((double)(1) / g1 + c2 * rinv);
    double c8 = ((k * c6 + r * r0) + c3 * c5) * r0inv3;
    double c12 = g0 * h - g1 * r0;
    double c17 = (r0 - r) - g2 * k;
    double c18 = eta * g1 + ((double)(2) * g2) * k;
    double c20 = k * (g2 * k + r) - ((g0 * r0) * zeta);
    double c21 = (((((((g2 * k - r0) * (beta * c3 - ((k * g1) * r))) * betainv) * rinv2) * r0inv3) / g1) +
             (((eta * g1) * rinv) * r0inv2)) -
            ((double)(2) * r0inv2);
    double c22 = rinv * (((-g1) - ((g0 * g2) / g1)) + ((g2 * c2) * rinv));
    double c25 = ((k * rinv) * r0inv2) *
            ((((((-g2) + ((((k * (c13 - g2 * r0)) * betainv) * r0inv2))) - c13 * r0inv) -
               ((((c12 * c3) * rinv) * r0inv2))) +
              (((((c13 * c2) * c3) * rinv2) * r0inv2))) -
             ((((c13 * (k * (g2 * k + r) - ((g0 * r0) * zeta))) * betainv) * rinv) * r0inv2));
    double c26 = ((k * rinv2) * r0inv) * ((((-g2) * c12) - g1 * c13) + (((g2 * c13) * c2) * rinv));
    double ddfdtdxx = dfdt * c21;
    double ddfdtdxv = dfdt * (c22 - h * c21);
    double ddfdtdvx = ddfdtdxv;
    double c34 = (((((((((-beta) * (eta * eta)) * (g2 * g2)) - ((eta * k) * h8)) - (h6 * ksq)) -
                 ((((((double)(2) * beta) * eta) * r0) * g1) * g2)) +
                (((((g2 * g2) - (((double)(3) * g1) * g3)) * beta) * k) * r0)) -
               ((beta * (g1 * g1)) * (r0 * r0))) *
              betainv * rinv2 +
              (((eta * (g2 * g2)) * rinv) / g1)) +
             ((((k * h8) * betainv) * rinv) / g1);
    double ddfdtdvv = dfdt * ((c34 - (((double)(2) * h) * c22)) + hsq * c21);
    double ddfdtdk = dfdt * (((((double)(1) / k - (betainv * r0inv)) - (((c17 * betainv) * rinv) * r0inv)) -
                         ((((c1 * (g1 * c2 - g0 * r)) * rinv2) * r0inv) / g1)));
    double ddfdtdk2 = (((-(g2 * k - r0)) * (((beta * r0) * (g3 - g1 * g2) - ((beta * eta) * (g2 * g2))) + k * h3)) * betainv) *
                 rinv2 * r0inv;
    double ddfdtdh = dfdt * ((((((g0 * rinv) / g1) - (c2 * rinv2)) - ((((double)(2) * k) * r0inv - beta) * c22)) - (eta * c21)));
    double dgdtmhdfdtm1dxx = c25;
    double dgdtmhdfdtm1dxv = c26 - h * c25;
    double dgdtmhdfdtm1dvx = c26 - h * c25;
    h2 = so.h2;
    double c33 = ((((((d * k) * rinv3) * r0inv) * k) * (h * g2 - r0 * g3)) +
             ((((k * ((((((((-eta) * k) * g1) * (g2 * g2)) - (((g1 * g2) * g3) * ksq)) -
                         ((((r0 * eta) * beta) * g1) * (g2 * g2))) -
                        (((r0 * k) * g1) * h2)) -
                       (((beta * (g2 * g2)) * g0) * (r0 * r0)))) *
                 betainv) *
                rinv2) *
               r0inv));
    double dgdtmhdfdtm1dvv = (c33 - (((double)(2) * h) * c26)) + hsq * c25;
    double dgdtmhdfdtm1dk =
        (rinv * r0inv) *
        ((((((-k) * (c13 - g2 * r0)) * betainv) * r0inv + c13) - ((((k * c13) * c17) * betainv) * rinv) * r0inv) +
         ((((k * c1) * c12) * rinv) * r0inv) - (((((k * c1) * c2) * c13) * rinv2) * r0inv));
    double g2s = g2 * g2;
    double dgdtmhdfdtm1dk2 =
        (((k * betainv) * rinv2) * r0inv) *
        (((((((-beta) * (eta * eta)) * pow4(g2)) + ((eta * g2) * ((g1 * g2s + (g1 * g1) * g3) - (((double)(5) * g2) * g3))) * k) +
           (((g2 * g3) * h3) * ksq)) +
          ((((((double)(2) * eta) * r0) * beta) * g2s) * (g3 - g1 * g2))) +
         ((((((double)(4) * g3 - g0 * g3) - g1 * g2) * (g3 - g1 * g2)) * r0) * k) +
         (beta * ((((((double)(2) * g1) * g3) * g2) - (g1 * g1) * g2s) - (g3 * g3))) * (r0 * r0));
    double dgdtmhdfdtm1dh = ((((((g1 * k) * rinv) * r0inv) + (((k * c12) * rinv2) * r0inv)) -
                         ((((k * c2) * c13) * rinv3) * r0inv)) -
                        ((((double)(2) * k) * r0inv - beta) * c26)) -
                       (eta * c25);
    for (int j = 0; j < 3; ++j) {
      JK(3 + j, j) = dfdt;
      JK(3 + j, 3 + j) = dgdtm1;
      for (int i = 0; i < 3; ++i) {
        JK(3 + j, i) += (ddfdtdxx * x0[i] + ddfdtdxv * v0[i]) * x0[j] +
                        (dgdtmhdfdtm1dxx * x0[i] + dgdtmhdfdtm1dxv * v0[i]) * v0[j];
        JK(3 + j, 3 + i) += (ddfdtdvx * x0[i] + ddfdtdvv * v0[i]) * x0[j] +
                            (dgdtmhdfdtm1dvx * x0[i] + dgdtmhdfdtm1dvv * v0[i]) * v0[j];
      }
      JK(3 + j, 6) = ddfdtdk * x0[j] + dgdtmhdfdtm1dk * v0[j];
      JK(3 + j, 7) = ddfdtdh * x0[j] + dgdtmhdfdtm1dh * v0[j];
      jac_mass[3 + j] = (((GN * GN) * r0inv) * rinv) * (ddfdtdk2 * x0[j] + dgdtmhdfdtm1dk2 * v0[j]);
    }
    (void)c7; (void)c8; (void)c18; (void)c20;
  } else {
    double d = ((h + eta * g2) + ((double)(2) * k) * g3) * betainv;
    double c1 = d - r0 * g3;
    double c2 = eta * g0 + g1 * zeta;
    double c3 = d * k + g1 * (r0 * r0);
    double c6 = (r0 * g0 - k * g2) * betainv;
    double c9 = g2 * r - h1 * k;
    double c14 = r0 * g2 - k * h1;
    double c15 = eta * h1 + h2 * r0;
    double c16 = eta * h2 + ((g1 * gamma) * r0) / sqb;
    double c17 = (r0 - r) - g2 * k;
    double c18 = eta * g1 + ((double)(2) * g2) * k;
    double c19 = ((double)(4) * eta) * h1 + ((double)(3) * h2) * r0;
    double c23 = h2 * k - r0 * g1;
    const SeriesOut so = gfuncs<SER_H3 | SER_H5 | SER_H6>(gamma, beta, sqb);
    double h6 = so.h6;
    double h3 = so.h3;
    double h5 = so.h5;
    double h8 = (-((double)(2) * h3)) + (double)(3) * h5;
    double dfm1dxx = ((((k * rinv3) * betainv) * pow4(r0inv)) *
                 (((((((k * h1) * (r * r)) * r0) * (beta - ((double)(2) * k) * r0inv)) +
                    ((beta * c3) * (r * c23 + c14 * c2))) +
                   ((c14 * r) * (k * (r - g2 * k) + ((g0 * r0) * zeta))))));
    double dfm1dxv = ((k * rinv2) * r0inv) * (((k * (g2 * h2 + g1 * h1)) - ((((double)(2) * g1) * g2) * r0)) + (((g2 * c14) * c2) * rinv));
    double dfm1dvx = dfm1dxv;
    double dfm1dvv =
        (((k * r0inv) * rinv2) * betainv) *
        (((((((double)(2) * eta) * k) * (g2 * g3 - g1 * h1)) + ((((double)(3) * g3) * h2 - ((double)(4) * h1) * g2) * ksq)) +
          (((beta * g2) * r0) * (((double)(3) * h1) * k - g2 * r0))) +
         (c14 * rinv) * ((((((((-beta) * (g2 * g2)) * (eta * eta)) + ((eta * k) * (((double)(2) * g0) * g3 - h2))) - (h6 * ksq)) +
                           (((((-((double)(2) * eta)) * g1) * g2 + (k * (h1 - ((double)(2) * g1) * g3))) * beta) * r0)) -
                          ((beta * (g1 * g1)) * (r0 * r0)))));
    double dfm1dh = (((g1 * k) - ((h2 * ksq) * r0inv)) - ((((k * c14) * c2) * rinv) * r0inv)) * rinv2;
    double dfm1dk = (rinv * r0inv) *
               (((((((((((double)(4) * h1) * ksq) * betainv) * r0inv) - k * h1) - (((double)(2) * g2) * k) * betainv) + c14) -
                  ((((k * c14) * c17) * betainv) * rinv) * r0inv) +
                 ((((k * (g1 * r0 - k * h2)) * c1) * rinv) * r0inv)) -
                (((((k * c14) * c1) * c2) * rinv2) * r0inv));
    double dfm1dk2 =
        ((betainv * r0inv) * rinv2) *
        ((r * (((((((((double)(2) * eta) * k) * (g1 * h1 - g3 * g2)) + ((((double)(4) * g2) * h1 - ((double)(3) * g3) * h2) * ksq)) -
                   ((((eta * r0) * beta) * g1) * h1)) +
                  (((((g3 * h2) - (((double)(4) * g2) * h1)) * beta) * k) * r0)) +
                 ((((g2 * h1) * (beta * beta)) * (r0 * r0)))))) -
         (c14 * ((((((((-(eta * eta)) * beta) * (g2 * g2)) - ((k * eta) * h8)) - (ksq * h6)) -
                   (((eta * r0) * beta) * (g1 * g2 + g0 * g3))) +
                  (((((double)(2) * (h1 - g1 * g3)) * beta) * k) * r0)) -
                 (((g2 - ((beta * g1) * g3)) * beta) * (r0 * r0)))));
    double dgmhdxx = ((k * rinv) * r0inv) *
                ((((h2 + (((k * c19) * betainv) * r0inv2)) - (((c16 * c3) * rinv) * r0inv2)) +
                  (((c2 * c3) * c15) * ((rinv * r0inv) * (rinv * r0inv)))) -
                 ((((c15 * (k * (g2 * k + r) - ((g0 * r0) * zeta))) * betainv) * rinv) * r0inv2));
    double dgmhdxv = (k * rinv2) * ((((h1 * r) - (g2 * c16)) - (g1 * c15)) + (((g2 * c2) * c15) * rinv));
    double dgmhdvx = dgmhdxv;
    double dgmhdvv;
    {
      double A = ((double)(2) * (eta * eta)) * (g1 * h1 - g2 * g3);
      double B = (eta * k) * (((double)(4) * g2) * h1 - ((double)(3) * h2) * g3);
      double C = (r0 * eta) * (((double)(4) * g0) * h1 - ((double)(2) * g1) * g3);
      double D = (((double)(3) * r0) * k) * ((g1 + beta * g3) * h1 - g3 * g2);
      double E = (g0 * h8 - (beta * g1) * ((g2 * g2) + g1 * g3)) * (r0 * r0);
      double F2 = (((((beta * (g2 * g2)) * (eta * eta)) + (eta * k) * h8) + h6 * ksq) +
              ((((((double)(2) * eta) * g1) * g2 - k * ((g2 * g2) - ((double)(3) * g1) * g3)) * beta) * r0)) +
             (beta * (g1 * g1)) * (r0 * r0);
      double F = (c15 * rinv) * F2;
      dgmhdvv = ((k * betainv) * rinv2) * (((((A + B) + C) + D) + E) - F);
    }
    double dgmhdk = rinv * (((((((((k * c1) * c16) * rinv) * r0inv) + c15) - ((((k * c15) * c17) * betainv) * rinv) * r0inv) -
                         (((k * c19) * betainv) * r0inv)) -
                        (((((k * c1) * c2) * c15) * rinv2) * r0inv)));
    double h7 = ((beta * g1) * (g2 * g2)) - g0 * h8;
    double dgmhdk2;
    {
      double a1 = ((double)(2) * (eta * eta)) * (g3 * g2 - g1 * h1);
      double a2 = (eta * k) * (((double)(3) * g3) * h2 - ((double)(4) * g2) * h1);
      double a3 = (r0 * eta) * ((beta * g3) * (g1 * g2 + g0 * g3) - ((double)(2) * g0) * h6);
      double a4 = (((-h6) * (g1 + beta * g3) + g2 * ((double)(2) * g3 - h2)) * r0) * k;
      double a5 = (h7 - ((beta * beta) * g1) * (g3 * g3)) * (r0 * r0);
      double A = (((a1 + a2) + a3) + a4) + a5;
      double b1 = ((-beta) * (eta * eta)) * (g2 * g2);
      double b2 = (eta * k) * ((-h2) + ((double)(2) * g0) * g3);
      double b3 = h6 * ksq;
      double b4 = ((r0 * eta) * beta) * (h2 + ((double)(2) * g0) * g3);
      double b5 = ((((double)(2) * beta) * ((double)(2) * h1 - g2 * g2)) * r0) * k;
      double b6 = (beta * ((beta * g1) * g3 - g2)) * (r0 * r0);
      double B = ((((b1 + b2) - b3) - b4) + b5) + b6;
      dgmhdk2 = (betainv * rinv2) * (r * A - c15 * B);
    }
    double dgmhdh = (k * rinv3) * ((r * c16) - (c2 * c15));
    for (int j = 0; j < 3; ++j) {
      JK(j, j) = fm1;
      JK(j, 3 + j) = gmh;
      for (int i = 0; i < 3; ++i) {
        JK(j, i) += (dfm1dxx * x0[i] + dfm1dxv * v0[i]) * x0[j] + (dgmhdxx * x0[i] + dgmhdxv * v0[i]) * v0[j];
        JK(j, 3 + i) += (dfm1dvx * x0[i] + dfm1dvv * v0[i]) * x0[j] + (dgmhdvx * x0[i] + dgmhdvv * v0[i]) * v0[j];
      }
      JK(j, 6) = dfm1dk * x0[j] + dgmhdk * v0[j];
      JK(j, 7) = dfm1dh * x0[j] + dgmhdh * v0[j];
      jac_mass[j] = (((GN * GN) * rinv) * r0inv) * (dfm1dk2 * x0[j] + dgmhdk2 * v0[j]);
    }
    double c5 = ((r0 - k * g2) * rinv) / g1;
    double c7 = g2 * ((double)(1) / g1 + c2 * rinv);
    double c8 = ((k * c6 + r * r0) + c3 * c5) * r0inv3;
    double c12 = g0 * h - g1 * r0;
    double c20 = k * (g2 * k + r) - ((g0 * r0) * zeta);
    double ddfdtdxx = (dfdt * (((((((eta * g1) * rinv) - (double)(2)) - ((((g0 * c3) * rinv) * r0inv) / g1)) +
                            (((c2 * c3) * r0inv) * rinv2)) -
                           ((((k * (k * g2 - r0)) * betainv) * rinv) * r0inv)))) *
                  r0inv2;
    double ddfdtdxv = ((-dfdt) * (((g0 * g2) / g1) + ((r0 * g1 + eta * g2) * rinv))) * rinv;
    double ddfdtdvx = ddfdtdxv;
    double ddfdtdvv = ((((-k) * rinv3) * r0inv) * betainv) *
                 ((((((beta * eta) * (g2 * g2)) + (k * h8)) * (r0 * g0 + k * g2))) +
                  (g1 * ((((-h6) * ksq) + (((((-((double)(2) * eta)) * g1) * g2) + ((h1 - ((double)(2) * g1) * g3) * k)) * beta) * r0) -
                         ((beta * (g1 * g1)) * (r0 * r0)))));
    double ddfdtdk = dfdt * ((((double)(1) / k) + ((((c1 * (r0 - g2 * k)) * r0inv) * rinv2) / g1)) - ((betainv * r0inv) * ((double)(1) + c17 * rinv)));
    double ddfdtdk2 = ((((r0 - g2 * k) * betainv) * r0inv) * rinv2) *
                 ((((((-eta) * beta) * (g2 * g2)) + (h3 * k)) + (((g3 - g1 * g2) * beta) * r0)));
    double ddfdtdh = ((dfdt * (r0 - g2 * k)) * rinv2) / g1;
    double dgdotm1dxx = (rinv2 * r0inv3) * (((((eta * g2 + g1 * r0) * k) * c3) * rinv) +
                                        (((g2 * k) * (k * (g2 * k - r) - ((g0 * r0) * zeta))) * betainv));
    double dgdotm1dxv = ((k * g2) * rinv3) * ((r * g1 + r0 * g1) + eta * g2);
    double dgdotm1dvx = dgdotm1dxv;
    double dgdotm1dvv = ((k * betainv) * rinv3) *
                   ((((((((eta * eta) * beta) * cube(g2)) - (((eta * k) * g2) * h3)) + (((((double)(3) * r0) * eta) * beta) * g1) * (g2 * g2)) +
                      ((r0 * k) * (((-g0) * h6) + (((((double)(3) * beta) * g1) * g2) * g3)))) +
                     (((beta * g2) * (g0 * g2 + (g1 * g1))) * (r0 * r0))));
    double dgdotm1dk = (rinv * r0inv) * (((((-r0) * g2) + ((((g2 * k) * ((r + r0) - g2 * k)) * betainv) * rinv)) -
                                     (((k * g1) * c1) * rinv)) +
                                    ((((k * g2) * c1) * c2) * rinv2));
    double dgdotm1dk2 = (betainv * rinv2) *
                   (((((((-beta) * (eta * eta)) * cube(g2)) + (((eta * k) * g2) * h3)) +
                      ((((eta * r0) * beta) * g2) * (g3 - ((double)(2) * g1) * g2))) +
                     (((h6 - beta * cube(g2)) * r0) * k)) +
                    (((beta * g1) * (g3 - g1 * g2)) * (r0 * r0)));
    double dgdotm1dh = (k * rinv3) * ((g2 * c2) - (r * g1));
    for (int j = 0; j < 3; ++j) {
      JK(3 + j, j) = dfdt;
      JK(3 + j, 3 + j) = dgdtm1;
      for (int i = 0; i < 3; ++i) {
        JK(3 + j, i) += (ddfdtdxx * x0[i] + ddfdtdxv * v0[i]) * x0[j] + (dgdotm1dxx * x0[i] + dgdotm1dxv * v0[i]) * v0[j];
        JK(3 + j, 3 + i) += (ddfdtdvx * x0[i] + ddfdtdvv * v0[i]) * x0[j] + (dgdotm1dvx * x0[i] + dgdotm1dvv * v0[i]) * v0[j];
      }
      JK(3 + j, 6) = ddfdtdk * x0[j] + dgdotm1dk * v0[j];
      JK(3 + j, 7) = ddfdtdh * x0[j] + dgdotm1dh * v0[j];
      jac_mass[3 + j] = (((GN * GN) * rinv) * r0inv) * (ddfdtdk2 * x0[j] + dgdotm1dk2 * v0[j]);
    }
    (void)c7; (void)c8; (void)c9; (void)c18; (void)c20; (void)c5;
  }
}


}  // namespace nbg
