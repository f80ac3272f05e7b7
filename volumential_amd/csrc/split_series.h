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
inline std::vector<double> split_derivative_coefficients(int dim, std::complex<double> k, int N) {
  const std::vector<double> c = split_coefficients(dim, k, N);
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
