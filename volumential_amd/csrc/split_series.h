// Series form of the Helmholtz split remainder R = G_k - G_0, shared by the
// point-to-point correction and the plan's split blocks.  A split of order
// m > 1 drops the terms it tabulates (expansion_wrangler_fpnd.py:284-304):
// 2D the log parts s_n of n < m, 3D the even n <= 2 (m - 1).
//   2D  R = c0 + P_b(r^2) - log(r) / (2 pi) * P_a(r^2),
//       P_a(u) = sum_{n=1..N} s_n u^n,  s_n = (-1)^n (k^2/4)^n / (n!)^2,
//       P_b(u) = sum_n s_n ((H_n - log(k/2) - gamma) / (2 pi) + i/4) u^n,
//       c0 = i/4 - (log(k/2) + gamma) / (2 pi)  (the log term is 0 at r = 0);
//   3D  R = sum_{n=1..N} (i k)^n / (4 pi n!) r^(n-1)
// (volumential/expansion_wrangler_fpnd.py:212-312; at r = 0 both give the
// self-pair limit of 5293-5317).
#pragma once
#include <cmath>
#include <complex>
#include <vector>

#include <hip/hip_runtime.h>

namespace vl1 {

constexpr int kSplitMaxTerms = 96;
constexpr double kEulerGamma = 0.5772156649015329;

// ---- Closed forms of G_k - G_0 (the legacy subtraction, 3806-3822) ----
struct dcplx {
  double re, im;
};
__device__ inline dcplx cmul(dcplx a, dcplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ inline dcplx cdiv(dcplx a, dcplx b) {  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double t = b.im / b.re, d = b.re + b.im * t;
    return {(a.re + a.im * t) / d, (a.im - a.re * t) / d};
  }
  const double t = b.re / b.im, d = b.re * t + b.im;
  return {(a.re * t + a.im) / d, (a.im * t - a.re) / d};
}

// H0^(1)(z) for Im z >= 0, z != 0.  For |z| >= 8, Hankel's asymptotic
// expansion (DLMF 10.17.5, nu = 0: term_k = term_{k-1} i (-(2k-1)^2) /
// (8 k z)) cut at its smallest term, taken when 4 x that term x
// |sqrt(2 / (pi z)) e^{iz}| is below the power series' rounding scale
// 2e-16 e^|z| / sqrt(2 pi |z|); otherwise the power series of J0 and Y0
// (DLMF 10.8.1-2): J0 = sum_m t_m, t_m = (-z^2/4)^m / (m!)^2, Y0 = (2/pi)
// ((log(z/2) + gamma) J0 - sum_{m>=1} H_m t_m).  Absolute error below
// 2e-12 over the upper half plane (worst near |z| = 12 on the real axis).
__device__ inline dcplx hankel1_0(dcplx z) {
  const double az = hypot(z.re, z.im);
  const double arg = atan2(z.im, z.re);
  if (az >= 8.0) {
    dcplx s = {1.0, 0.0}, term = {1.0, 0.0};
    double tmin = 1.0;
    for (int k = 1; k < 64; ++k) {
      const double c = -(double)((2 * k - 1) * (2 * k - 1)) / (8.0 * k);
      const dcplx nt = cdiv(dcplx{-c * term.im, c * term.re}, z);  // term * (i c) / z
      const double an = hypot(nt.re, nt.im);
      if (an > tmin) break;
      term = nt;
      tmin = an;
      s.re += term.re;
      s.im += term.im;
      if (an < 1e-17) break;
    }
    const double pref = sqrt(2.0 / (M_PI * az)) * exp(-z.im);
    if (4.0 * tmin * pref < 2e-16 * exp(az) / sqrt(2.0 * M_PI * az)) {
      const double ph = z.re - 0.25 * M_PI - 0.5 * arg;
      return cmul(dcplx{pref * cos(ph), pref * sin(ph)}, s);
    }
  }
  const dcplx q = {0.25 * (z.re * z.re - z.im * z.im), 0.5 * z.re * z.im};
  const double aq = 0.25 * az * az;
  dcplx t = {1.0, 0.0}, J = {1.0, 0.0}, S = {0.0, 0.0};
  double H = 0.0;
  for (int m = 1; m < 100; ++m) {
    const double inv = -1.0 / ((double)m * m);
    t = cmul(t, dcplx{q.re * inv, q.im * inv});
    H += 1.0 / m;
    J.re += t.re;
    J.im += t.im;
    S.re -= t.re * H;
    S.im -= t.im * H;
    if ((double)m * m > aq && hypot(t.re, t.im) * (1.0 + H) < 1e-18 * fmax(hypot(J.re, J.im), 1.0))
      break;
  }
  const dcplx L = {log(0.5 * az) + kEulerGamma, arg};
  dcplx Y = cmul(L, J);
  Y.re = (Y.re + S.re) * (2.0 / M_PI);
  Y.im = (Y.im + S.im) * (2.0 / M_PI);
  return {J.re - Y.im, J.im + Y.re};
}

// H1^(1)(z) for Im z >= 0, z != 0, by the same rule as hankel1_0: the
// asymptotic expansion (nu = 1: term_k = term_{k-1} i (4 - (2k-1)^2) /
// (8 k z), phase z - 3 pi / 4) where it is the more accurate, otherwise
// J1 = (z/2) sum_m t_m, t_m = (-z^2/4)^m / (m! (m+1)!), and Y1 = (2/pi)
// log(z/2) J1 - 2 / (pi z) - (z / (2 pi)) sum_m (psi(m+1) + psi(m+2)) t_m
// (DLMF 10.8.1-2).  Error below 1e-12 (absolute, relative where |H1| > 1)
// over the upper half plane in a float64 restatement against scipy.
__device__ inline dcplx hankel1_1(dcplx z) {
  const double az = hypot(z.re, z.im);
  const double arg = atan2(z.im, z.re);
  if (az >= 8.0) {
    dcplx s = {1.0, 0.0}, term = {1.0, 0.0};
    double tmin = 1.0;
    for (int k = 1; k < 64; ++k) {
      const double c = (4.0 - (double)((2 * k - 1) * (2 * k - 1))) / (8.0 * k);
      const dcplx nt = cdiv(dcplx{-c * term.im, c * term.re}, z);  // term * (i c) / z
      const double an = hypot(nt.re, nt.im);
      if (an > tmin) break;
      term = nt;
      tmin = an;
      s.re += term.re;
      s.im += term.im;
      if (an < 1e-17) break;
    }
    const double pref = sqrt(2.0 / (M_PI * az)) * exp(-z.im);
    if (4.0 * tmin * pref < 2e-16 * exp(az) / sqrt(2.0 * M_PI * az)) {
      const double ph = z.re - 0.75 * M_PI - 0.5 * arg;
      return cmul(dcplx{pref * cos(ph), pref * sin(ph)}, s);
    }
  }
  const dcplx q = {0.25 * (z.re * z.re - z.im * z.im), 0.5 * z.re * z.im};
  const double aq = 0.25 * az * az;
  dcplx t = {1.0, 0.0}, J = {1.0, 0.0};
  dcplx S = {1.0 - 2.0 * kEulerGamma, 0.0};  // psi(1) + psi(2)
  double H = 0.0;
  for (int m = 1; m < 100; ++m) {
    const double inv = -1.0 / ((double)m * (m + 1));
    t = cmul(t, dcplx{q.re * inv, q.im * inv});
    H += 1.0 / m;
    const double psi = 2.0 * (H - kEulerGamma) + 1.0 / (m + 1);
    J.re += t.re;
    J.im += t.im;
    S.re += psi * t.re;
    S.im += psi * t.im;
    if ((double)m * m > aq &&
        4.0 * hypot(t.re, t.im) * (1.0 + H) < 1e-18 * fmax(hypot(J.re, J.im), 1.0))
      break;
  }
  const dcplx hz = {0.5 * z.re, 0.5 * z.im};
  const dcplx J1 = cmul(hz, J);
  const dcplx L = {log(0.5 * az), arg};
  const dcplx LJ = cmul(L, J1);
  const dcplx iz = cdiv(dcplx{1.0, 0.0}, z);
  const dcplx zS = cmul(hz, S);
  const dcplx Y1 = {(2.0 * LJ.re - 2.0 * iz.re - zS.re) / M_PI,
                    (2.0 * LJ.im - 2.0 * iz.im - zS.im) / M_PI};
  return {J1.re - Y1.im, J1.im + Y1.re};
}

// (G_k - G_0)'(r) / r for r > 0 from the closed forms: 3D [e^{ikr} (ikr -
// 1) + 1] / (4 pi r^3); 2D [-(ik/4) H1^(1)(kr) + 1 / (2 pi r)] / r.  k = 0
// gives 0.
template <int DIM>
__device__ inline void legacy_radial_derivative(double r2, double kr, double ki, double& gr,
                                                double& gi) {
  const double r = sqrt(r2);
  if (kr == 0.0 && ki == 0.0) {
    gr = gi = 0.0;
    return;
  }
  if (DIM == 3) {
    const double e = exp(-ki * r), c = cos(kr * r), sn = sin(kr * r);
    const double A = -ki * r - 1.0, B = kr * r;  // ikr - 1
    const double w = 0.25 / (M_PI * r2 * r);
    gr = (e * (c * A - sn * B) + 1.0) * w;
    gi = e * (sn * A + c * B) * w;
  } else {
    const dcplx h = hankel1_1(dcplx{kr * r, ki * r});
    gr = (0.25 * (ki * h.re + kr * h.im) + 0.5 / (M_PI * r)) / r;
    gi = 0.25 * (ki * h.im - kr * h.re) / r;
  }
}

// G_k(r) - G_0(r) for r > 0 from the closed forms: 3D (e^{ikr} - 1) / (4 pi
// r); 2D (i/4) H0^(1)(kr) + log(r) / (2 pi).  k = 0 gives 0.
template <int DIM>
__device__ inline void legacy_difference(double r2, double kr, double ki, double& vr,
                                         double& vi) {
  const double r = sqrt(r2);
  if (kr == 0.0 && ki == 0.0) {
    vr = vi = 0.0;
    return;
  }
  if (DIM == 3) {
    const double e = exp(-ki * r), w = 0.25 / (M_PI * r);
    vr = (e * cos(kr * r) - 1.0) * w;
    vi = e * sin(kr * r) * w;
  } else {
    const dcplx h = hankel1_0(dcplx{kr * r, ki * r});
    vr = -0.25 * h.im + 0.5 / M_PI * log(r);
    vi = 0.25 * h.re;
  }
}

// Coefficient layout (interleaved complex): 2D c0, then (s_n, b_n) for
// n = 1..N; 3D c_n for n = 1..N.
template <int DIM>
__device__ inline void remainder_value(double r2, const double* __restrict__ cf, int N, double& vr,
                                       double& vi) {
  if (DIM == 2) {
    // Horner in u = r^2 for P_a and P_b (coefficients n = N .. 1)
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    for (int n = N; n >= 1; --n) {
      const double* c = cf + 2 + 4 * (n - 1);
      const double nar = fma(ar, r2, c[0]), nai = fma(ai, r2, c[1]);
      const double nbr = fma(br, r2, c[2]), nbi = fma(bi, r2, c[3]);
      ar = nar;
      ai = nai;
      br = nbr;
      bi = nbi;
    }
    ar *= r2;
    ai *= r2;
    br *= r2;
    bi *= r2;
    const double lg = r2 > 0.0 ? -0.25 / M_PI * log(r2) : 0.0;  // -log(r) / (2 pi)
    vr = cf[0] + br + lg * ar;
    vi = cf[1] + bi + lg * ai;
  } else {
    const double r = sqrt(r2);
    double pr = 0.0, pi = 0.0;
    for (int n = N; n >= 1; --n) {
      const double* c = cf + 2 * (n - 1);
      const double npr = fma(pr, r, c[0]), npi = fma(pi, r, c[1]);
      pr = npr;
      pi = npi;
    }
    vr = pr;
    vi = pi;
  }
}

// R'(r) / r of the same series, for axis target / source derivative
// wrappers (the pair value is +-(x_a - y_a) R'(r) / r; the reference runs
// its P2P on the wrapped series kernel, 3884-3957):
//   2D  sum_{n=1..N} u^(n-1) (B'_n + lg A'_n),  lg = -log(r) / (2 pi),
//       A'_n = 2n s_n,  B'_n = 2n b_n - s_n / (2 pi)   (u = r^2);
//   3D  (1/r) sum_{m=0..N-2} (m + 1) c_{m+2} r^m.
// Coefficient layout (interleaved complex): 2D (A'_n, B'_n) n = 1..N; 3D
// the N - 1 values (m + 1) c_{m+2}.  Callers treat r = 0 (self pairs,
// value 0) themselves.
template <int DIM>
__device__ inline void remainder_radial_derivative(double r2, const double* __restrict__ cf, int N,
                                                   double& gr, double& gi) {
  if (DIM == 2) {
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    for (int n = N; n >= 1; --n) {
      const double* c = cf + 4 * (n - 1);
      const double nar = fma(ar, r2, c[0]), nai = fma(ai, r2, c[1]);
      const double nbr = fma(br, r2, c[2]), nbi = fma(bi, r2, c[3]);
      ar = nar;
      ai = nai;
      br = nbr;
      bi = nbi;
    }
    const double lg = -0.25 / M_PI * log(r2);
    gr = fma(lg, ar, br);
    gi = fma(lg, ai, bi);
  } else {
    const double r = sqrt(r2);
    double pr = 0.0, pi = 0.0;
    for (int m = N - 2; m >= 0; --m) {
      const double* c = cf + 2 * m;
      const double npr = fma(pr, r, c[0]), npi = fma(pi, r, c[1]);
      pr = npr;
      pi = npi;
    }
    const double ir = 1.0 / r;
    gr = pr * ir;
    gi = pi * ir;
  }
}

// Series coefficients (interleaved complex), the reference's formulas.
inline std::vector<double> split_coefficients(int dim, std::complex<double> k, int N,
                                              int order = 1) {
  using cd = std::complex<double>;
  std::vector<double> out;
  if (dim == 2) {
    const bool zero = std::abs(k) == 0.0;
    const cd L = zero ? cd(0.0) : std::log(0.5 * k) + kEulerGamma;
    const cd c0 = zero ? cd(0.0) : cd(0.0, 0.25) - L / (2.0 * M_PI);
    out.push_back(c0.real());
    out.push_back(c0.imag());
    double fact = 1.0, harm = 0.0;
    cd kp = 1.0;
    for (int n = 1; n <= N; ++n) {
      fact *= n;
      harm += 1.0 / n;
      kp *= -k * k / 4.0;  // (-1)^n (k^2/4)^n
      const cd s = zero ? cd(0.0) : kp / (fact * fact);
      const cd b = s * ((harm - L) / (2.0 * M_PI) + cd(0.0, 0.25));
      const cd sl = n < order ? cd(0.0) : s;  // log part, tabulated below order m
      out.push_back(sl.real());
      out.push_back(sl.imag());
      out.push_back(b.real());
      out.push_back(b.imag());
    }
  } else {
    double fact = 1.0;
    cd ik = cd(0.0, 1.0) * k, p = 1.0;
    for (int n = 1; n <= N; ++n) {
      fact *= n;
      p *= ik;
      const bool tabulated = n % 2 == 0 && n <= 2 * (order - 1);
      const cd c = tabulated ? cd(0.0) : p / (4.0 * M_PI * fact);
      out.push_back(c.real());
      out.push_back(c.imag());
    }
  }
  return out;
}

// Coefficients of remainder_radial_derivative, from those above.
inline std::vector<double> split_derivative_coefficients(int dim, std::complex<double> k, int N,
                                                         int order = 1) {
  // term by term from the value series of the same order (the terms an
  // order-m split tabulates are already dropped there)
  const std::vector<double> c = split_coefficients(dim, k, N, order);
  std::vector<double> out;
  if (dim == 2) {
    for (int n = 1; n <= N; ++n) {
      const double* sb = c.data() + 2 + 4 * (n - 1);  // s_n, b_n
      out.push_back(2.0 * n * sb[0]);
      out.push_back(2.0 * n * sb[1]);
      out.push_back(2.0 * n * sb[2] - sb[0] / (2.0 * M_PI));
      out.push_back(2.0 * n * sb[3] - sb[1] / (2.0 * M_PI));
    }
  } else {
    for (int m = 0; m + 2 <= N; ++m) {
      out.push_back((m + 1) * c[2 * (m + 1)]);
      out.push_back((m + 1) * c[2 * (m + 1) + 1]);
    }
  }
  return out;
}

}  // namespace vl1
