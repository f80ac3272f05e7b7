// Helmholtz split correction of list 1 (base quadrature).
//
// With the split, list 1 of a Helmholtz (or Yukawa, k = i lambda) potential
// is the Laplace-table evaluation plus an online point-to-point sum of the
// smooth remainder R = G_k - G_0 over every list-1 (target box, source box)
// pair (volumential/expansion_wrangler_fpnd.py:3491-3836):
//   corr[t] = sum_{s in list-1 boxes of t} R(|x_t - y_s|) * strength[s],
// strength = density x quadrature weight.  R is the reference's series
// (_HelmholtzSplitSeriesRemainderKernel, 212-312; split_series.h), by Horner.
// At r = 0 it gives the self-pair limit of 5293-5317, so self pairs need no
// special case.  N is the caller's series length (4337-4389).  For an axis
// target (source) derivative wrapper of the output kernel (3884-3957) the
// pair value is +(-)(x_a - y_a) R'(r) / r, and self pairs give 0 (the
// reference's limit for a wrapped chain, 5293-5300).
//
// GPU mapping: one workgroup per target box.  Each listed source box's
// point coordinates and complex strengths (a chunk of RC densities) are
// staged in LDS; a lane owns one target point and a share of the source
// points; R is evaluated once per (target, source) pair and applied to all
// RC densities; partial sums are reduced through LDS.  The series
// coefficients sit in LDS for the whole launch.
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <type_traits>
#include <vector>

#include "internal.h"
#include "split_series.h"

namespace vl1 {

constexpr int kSplitBlock = 256;

// Pair value: the remainder series, its axis derivative, or the closed
// forms of G_k - G_0 (legacy subtraction).
constexpr int kPairSeries = 0, kPairDeriv = 1, kPairLegacy = 2, kPairLegacyDeriv = 3;

struct SplitParams {
  int dim, n_rhs, nt, n_terms, chunk, ncoef;
  double k_re, k_im;   // wave number (legacy subtraction)
  int deriv_axis;      // derivative wrapper (DERIV kernels): axis and sign
  double deriv_sign;   // +1 target derivative, -1 source derivative
  const double* coef;  // split_coefficients / split_derivative_coefficients
  const int32_t* target_boxes;
  const int32_t* nbr_starts;
  const int32_t* nbr_lists;
  const int32_t* box_starts;
  const int32_t* box_counts;
  const double* points;  // [dim][n_points]
  int64_t n_points;
  const int32_t* src_starts;  // sources: the targets, or a separate set
  const int32_t* src_counts;
  const double* src_points;   // [dim][n_src]
  int64_t n_src;
  const double* strengths;  // [n_rhs][lds] interleaved complex
  int64_t lds;
  double* result;  // [n_rhs][ldr] interleaved complex
  int64_t ldr;
};

template <int DIM, int RC, int MODE>
__global__ void __launch_bounds__(kSplitBlock) split_correction_kernel(const SplitParams p) {
  constexpr bool DERIV = MODE == kPairDeriv || MODE == kPairLegacyDeriv;
  extern __shared__ double smem[];
  const int N = p.n_terms;
  const int ncoef = p.ncoef;
  const int P = kSplitBlock / p.nt;
  double* sC = smem;                 // series coefficients
  double* sX = sC + ncoef;           // source coordinates [DIM][chunk]
  const int chunk = p.chunk;         // source points staged per pass
  double* sF = sX + DIM * chunk;     // strengths [chunk][RC][2]
  double* red = sX;                  // part sums [256][RC][2], after the staging is done
  const int tid = threadIdx.x;
  const int lane_t = tid % p.nt;
  const int part = tid / p.nt;
  for (int i = tid; i < ncoef; i += kSplitBlock) sC[i] = p.coef[i];

  const int tbox = blockIdx.x;
  const int tb = p.target_boxes[tbox];
  const int tbeg = p.box_starts[tb];
  const int ntb = p.box_counts[tb];
  const int l0 = p.nbr_starts[tbox], l1 = p.nbr_starts[tbox + 1];

  for (int r0 = 0; r0 < p.n_rhs; r0 += RC) {
    const int rc = min(RC, p.n_rhs - r0);
    for (int t0 = 0; t0 < ntb; t0 += p.nt) {
      const int ti = t0 + lane_t;
      const bool tact = ti < ntb;
      double xt[DIM];
#pragma unroll
      for (int a = 0; a < DIM; ++a) xt[a] = tact ? p.points[(int64_t)a * p.n_points + tbeg + ti] : 0.0;
      double acc[RC], aci[RC];
#pragma unroll
      for (int r = 0; r < RC; ++r) acc[r] = aci[r] = 0.0;
      for (int li = l0; li < l1; ++li) {
        const int sb = p.nbr_lists[li];
        const int sbeg = p.src_starts[sb];
        const int ns = p.src_counts[sb];
        for (int c0 = 0; c0 < ns; c0 += chunk) {
          const int nc = min(chunk, ns - c0);
          __syncthreads();  // previous chunk done with LDS (and coefficients staged)
          for (int i = tid; i < DIM * nc; i += kSplitBlock) {
            const int a = i / nc, j = i % nc;
            sX[a * chunk + j] = p.src_points[(int64_t)a * p.n_src + sbeg + c0 + j];
          }
          for (int i = tid; i < nc * RC * 2; i += kSplitBlock) {
            const int c = i & 1, rj = i >> 1;
            const int j = rj / RC, r = rj % RC;
            sF[i] = r < rc ? p.strengths[(int64_t)(r0 + r) * p.lds + 2 * (int64_t)(sbeg + c0 + j) + c]
                           : 0.0;
          }
          __syncthreads();
          if (tact && part < P) {  // threads past P * nt (nt not dividing 256) idle
            for (int j = part; j < nc; j += P) {
              double r2 = 0.0, da = 0.0;
#pragma unroll
              for (int a = 0; a < DIM; ++a) {
                const double d = xt[a] - sX[a * chunk + j];
                r2 = fma(d, d, r2);
                if (DERIV && a == p.deriv_axis) da = d;
              }
              double vr, vi;
              if constexpr (DERIV) {
                // +-(x_a - y_a) R'(r) / r; the self pair (r = 0) gives 0
                vr = vi = 0.0;
                if (r2 > 0.0) {
                  if constexpr (MODE == kPairLegacyDeriv)
                    legacy_radial_derivative<DIM>(r2, p.k_re, p.k_im, vr, vi);
                  else
                    remainder_radial_derivative<DIM>(r2, sC, N, vr, vi);
                  const double f = p.deriv_sign * da;
                  vr *= f;
                  vi *= f;
                }
              } else if constexpr (MODE == kPairLegacy) {
                // closed forms off the diagonal; a self pair takes the
                // series at r = 0, the limit the reference adds back
                if (r2 > 0.0)
                  legacy_difference<DIM>(r2, p.k_re, p.k_im, vr, vi);
                else
                  remainder_value<DIM>(r2, sC, N, vr, vi);
              } else {
                remainder_value<DIM>(r2, sC, N, vr, vi);
              }
              const double* f = sF + (size_t)j * RC * 2;
#pragma unroll
              for (int r = 0; r < RC; ++r) {
                const double fr = f[2 * r], fi = f[2 * r + 1];
                acc[r] = fma(vr, fr, fma(-vi, fi, acc[r]));
                aci[r] = fma(vr, fi, fma(vi, fr, aci[r]));
              }
            }
          }
        }
      }
      if (P > 1) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RC; ++r) {
          red[((part * p.nt + lane_t) * RC + r) * 2] = acc[r];
          red[((part * p.nt + lane_t) * RC + r) * 2 + 1] = aci[r];
        }
        __syncthreads();
        if (part == 0)
          for (int q = 1; q < P; ++q)
#pragma unroll
            for (int r = 0; r < RC; ++r) {
              acc[r] += red[((q * p.nt + lane_t) * RC + r) * 2];
              aci[r] += red[((q * p.nt + lane_t) * RC + r) * 2 + 1];
            }
      }
      if (part == 0 && tact)
        for (int r = 0; r < rc; ++r) {
          p.result[(int64_t)(r0 + r) * p.ldr + 2 * (int64_t)(tbeg + ti)] = acc[r];
          p.result[(int64_t)(r0 + r) * p.ldr + 2 * (int64_t)(tbeg + ti) + 1] = aci[r];
        }
      __syncthreads();  // reduction buffer reused by the next target chunk
    }
  }
}

template <int DIM, int RC, int MODE>
static hipError_t launch_split(const SplitParams& prm, int grid, size_t smem, hipStream_t s) {
  auto fn = split_correction_kernel<DIM, RC, MODE>;
  hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kSplitBlock), smem, s, prm);
  return hipGetLastError();
}

}  // namespace vl1

using namespace vl1;

extern "C" int vl1_split_correction(const vl1_split_args* a, void* stream) {
  clear_error();
  if (!a) return fail(VL1_E_ARG, "vl1_split_correction: null argument");
  if (a->dim != 2 && a->dim != 3) return fail(VL1_E_ARG, "vl1_split_correction: dim must be 2 or 3");
  if (a->split_order < 1 || 2 * (a->split_order - 1) > kSplitMaxTerms)
    return fail(VL1_E_ARG, "vl1_split_correction: split_order must be 1..49");
  if (a->series_nmax < 1 || a->series_nmax > kSplitMaxTerms)
    return fail(VL1_E_ARG, "vl1_split_correction: series_nmax must be 1..96");
  if (a->n_tgt_boxes < 0 || a->n_rhs < 1 || a->n_points < 1 || a->max_box_points < 1)
    return fail(VL1_E_ARG, "vl1_split_correction: bad sizes");
  if (a->n_tgt_boxes == 0) return VL1_OK;
  if (!a->target_boxes || !a->nbr_starts || !a->nbr_lists || !a->box_starts || !a->box_counts ||
      !a->points || !a->strengths || !a->result)
    return fail(VL1_E_ARG, "vl1_split_correction: null device array");
  const bool sep = a->src_points != nullptr;
  if (sep && (!a->src_box_starts || !a->src_box_counts || a->n_src_points < 1 ||
              a->max_src_box_points < 1))
    return fail(VL1_E_ARG, "vl1_split_correction: incomplete separate source set");
  const int64_t n_src = sep ? a->n_src_points : a->n_points;
  if (a->n_rhs > 1 && (a->ld_strengths < 2 * n_src || a->ld_result < 2 * a->n_points))
    return fail(VL1_E_ARG, "vl1_split_correction: bad leading dimension");
  if (a->deriv_kind < VL1_SPLIT_DERIV_NONE || a->deriv_kind > VL1_SPLIT_DERIV_SOURCE)
    return fail(VL1_E_ARG, "vl1_split_correction: bad deriv_kind");
  const bool deriv = a->deriv_kind != VL1_SPLIT_DERIV_NONE;
  if (deriv && (a->deriv_axis < 0 || a->deriv_axis >= a->dim))
    return fail(VL1_E_ARG, "vl1_split_correction: deriv_axis out of range");

  if (a->legacy_subtraction != 0 && a->legacy_subtraction != 1)
    return fail(VL1_E_ARG, "vl1_split_correction: legacy_subtraction must be 0 or 1");
  const bool legacy = a->legacy_subtraction == 1;
  if (legacy && (a->coefficients || a->split_order != 1))
    return fail(VL1_E_UNSUPPORTED,
                "vl1_split_correction: the legacy subtraction takes split order 1 and no "
                "coefficients");
  if (legacy && a->k_im < 0.0)
    return fail(VL1_E_ARG, "vl1_split_correction: the legacy subtraction needs Im k >= 0");
  hipStream_t s = (hipStream_t)stream;
  const std::complex<double> k(a->k_re, a->k_im);
  std::vector<double> coef;
  if (a->coefficients) {
    // value layout, or with a derivative wrapper the R'(r)/r layout
    const int nc = deriv ? (a->dim == 2 ? 4 * a->series_nmax : 2 * (a->series_nmax - 1))
                         : (a->dim == 2 ? 2 + 4 * a->series_nmax : 2 * a->series_nmax);
    coef.assign(a->coefficients, a->coefficients + nc);
    for (double c : coef)
      if (!std::isfinite(c)) return fail(VL1_E_ARG, "vl1_split_correction: coefficients must be finite");
  } else {
    coef = deriv ? split_derivative_coefficients(a->dim, k, a->series_nmax, a->split_order)
                 : split_coefficients(a->dim, k, a->series_nmax, a->split_order);
  }
  if (coef.empty()) coef.assign(2, 0.0);  // 3D derivative with one term: R' = 0
  DevBuf dcoef;
  hipError_t e = upload(dcoef, coef.data(), coef.size() * sizeof(double), s);
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("vl1_split_correction: ") + hipGetErrorString(e));

  SplitParams prm{};
  prm.dim = a->dim;
  prm.n_rhs = a->n_rhs;
  prm.n_terms = a->series_nmax;
  prm.ncoef = (int)coef.size();
  prm.k_re = a->k_re;
  prm.k_im = a->k_im;
  prm.deriv_axis = deriv ? a->deriv_axis : -1;
  prm.deriv_sign = a->deriv_kind == VL1_SPLIT_DERIV_SOURCE ? -1.0 : 1.0;
  prm.coef = dcoef.as<double>();
  prm.target_boxes = a->target_boxes;
  prm.nbr_starts = a->nbr_starts;
  prm.nbr_lists = a->nbr_lists;
  prm.box_starts = a->box_starts;
  prm.box_counts = a->box_counts;
  prm.points = a->points;
  prm.n_points = a->n_points;
  prm.src_starts = sep ? a->src_box_starts : a->box_starts;
  prm.src_counts = sep ? a->src_box_counts : a->box_counts;
  prm.src_points = sep ? a->src_points : a->points;
  prm.n_src = n_src;
  prm.strengths = a->strengths;
  prm.lds = a->n_rhs > 1 ? a->ld_strengths : 2 * n_src;
  prm.result = a->result;
  prm.ldr = a->n_rhs > 1 ? a->ld_result : 2 * a->n_points;
  // a lane per target (the box's targets, at most a block), P = 256 / nt
  // source parts: 36 targets use 7 parts = 252 of the 256 threads
  const int nt = std::min(a->max_box_points, kSplitBlock);
  prm.nt = nt;
  const int P = kSplitBlock / nt;
  // the remainder is evaluated once per (target, source) pair and RC-chunk;
  // RC <= 8 keeps the LDS small enough for several workgroups per CU
  const int RC = a->n_rhs == 1 ? 1 : (a->n_rhs <= 4 ? 4 : 8);
  // source points staged per pass: the source box size, rounded up to the parts
  const int msrc = sep ? a->max_src_box_points : a->max_box_points;
  const int chunk = std::min(nt * P, (msrc + P - 1) / P * P);
  prm.chunk = chunk;
  const size_t stage = (size_t)a->dim * chunk + (size_t)chunk * RC * 2;
  const size_t reduce = P > 1 ? (size_t)kSplitBlock * RC * 2 : 0;  // aliases the staging
  const size_t smem = sizeof(double) * ((size_t)coef.size() + std::max(stage, reduce));
  if (smem > 160 * 1024) return fail(VL1_E_UNSUPPORTED, "vl1_split_correction: LDS budget exceeded");
  const int grid = a->n_tgt_boxes;
  const int mode = deriv ? (legacy ? kPairLegacyDeriv : kPairDeriv)
                         : (legacy ? kPairLegacy : kPairSeries);
  auto pick = [&](auto dimc, auto modec) {
    constexpr int D = decltype(dimc)::value, M = decltype(modec)::value;
    return RC == 1   ? launch_split<D, 1, M>(prm, grid, smem, s)
           : RC == 4 ? launch_split<D, 4, M>(prm, grid, smem, s)
                     : launch_split<D, 8, M>(prm, grid, smem, s);
  };
  using D2 = std::integral_constant<int, 2>;
  using D3 = std::integral_constant<int, 3>;
  using MS = std::integral_constant<int, kPairSeries>;
  using MD = std::integral_constant<int, kPairDeriv>;
  using ML = std::integral_constant<int, kPairLegacy>;
  using MLD = std::integral_constant<int, kPairLegacyDeriv>;
  auto by_mode = [&](auto dimc) {
    switch (mode) {
      case kPairDeriv: return pick(dimc, MD{});
      case kPairLegacy: return pick(dimc, ML{});
      case kPairLegacyDeriv: return pick(dimc, MLD{});
      default: return pick(dimc, MS{});
    }
  };
  e = a->dim == 2 ? by_mode(D2{}) : by_mode(D3{});
  // the coefficient buffer is freed when dcoef leaves scope: wait for the launch
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("vl1_split_correction: ") + hipGetErrorString(e));
  return VL1_OK;
}
