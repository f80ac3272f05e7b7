// Batched list-1 evaluation plans: tile schedule, dense FP64-MFMA form and
// Ozaki-split BF16-MFMA form.
//
// The reference evaluates list 1 once per density (volume_fmm.py:983-989,
// expansion_wrangler_fpnd.py:3222-3489).  Here the densities are the GEMM N
// dimension:  for every target box tb and list-1 source box sb,
//     Y[:, (tb, r)] += alpha(sb) * M_{case(tb,sb), lev(sb)} . F[:, (sb, r)]
//                      + disp(sb) * (nmlz . F[:, (sb, r)])           (2D log term)
// where M_c[t][s] is the dense per-case block of the orbit-reduced table
// (expanded once per plan).  Target boxes are grouped into tiles of B boxes
// (B * R_chunk <= 64 columns) sorted by their case signature; a tile is
// processed by one workgroup that keeps its FP64 output block in registers
// and walks the distinct (case, level) blocks of its boxes ("steps").
//
// Split form (Ozaki scheme I): rows of M_c and columns of F are scaled by
// powers of two into (-1, 1) and cut into S BF16 slices of 8 bits each
// (slice i stored as integer * 2^-8i, exact in BF16).  Slice products
// A_i X_j with i + j < S are accumulated by v_mfma_f32_16x16x32_bf16 into
// three FP32 accumulator sets (i+j = 0, 1, >= 2): the first two hold exact
// integer sums (|a||x| <= 2^14, K <= 256, at most two products), the third
// carries the small tail terms.  After every step the three sets are summed
// in FP64, scaled by the row/column exponents and the pair scaling, and
// added to the FP64 output ("cross-box sums in FP64").
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "internal.h"
#include "split_series.h"

namespace vl1 {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float floatx4;
typedef __attribute__((ext_vector_type(4))) double doublex4;

constexpr int kCols = 64;      // output columns per tile (boxes x RHS)
constexpr int kMaxSlices = 6;
constexpr int kMaxTileRows = 224;  // block rows per tile (row groups beyond)

static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// ------------------------------------------------------------------ setup
// M[b][t][s] (FP64, [n_blocks][MP][KP], zero padded) from the compact table.
// Complex tables (interleaved re, im) expand to the real 2n x 2n block
// [[Re T, -Im T], [Im T, Re T]] acting on [Re f; Im f] and giving
// [Re y; Im y], so both contraction kernels run unchanged on complex data
// (8 real flops per complex multiply-add, as the complex product itself).
// Blocks with more rows than a tile takes are cut into n_rg row groups of
// rg_size rows (all K columns each): plan block b is row group b % n_rg of
// (case, level) block b / n_rg.
__global__ void expand_blocks_kernel(int dim, int q, int n, int cplx, int sp_mask, int MP, int KP,
                                     int n_blocks, int n_rg, int rg_size,
                                     const int32_t* __restrict__ blk_case,
                                     const int32_t* __restrict__ blk_lev,
                                     const uint8_t* __restrict__ perm,
                                     const int8_t* __restrict__ sign,
                                     const uint8_t* __restrict__ group,
                                     const uint16_t* __restrict__ ranks,
                                     const int32_t* __restrict__ offs,
                                     const double* __restrict__ table, int64_t n_entries,
                                     double* __restrict__ M) {
  const int64_t total = (int64_t)n_blocks * MP * KP;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(i % KP);
    const int tl = (int)((i / KP) % MP);
    const int bg = (int)(i / ((int64_t)KP * MP));
    const int b = bg / n_rg, t = (bg % n_rg) * rg_size + tl;  // (case, level) block, its row
    double v = 0.0;
    const int rows = cplx ? 2 * n : n;
    if (s < rows && tl < rg_size && t < rows) {
      const int c = blk_case[b];
      const int ss = s % n, tt = t % n;
      const int64_t rel = compact_rel_row(dim, q, perm + c * dim, sign + c * dim, group + c * dim, ss, tt);
      const int64_t at = (int64_t)blk_lev[b] * n_entries + offs[ranks[c]] + rel;
      if (!cplx) {
        v = table[at];
      } else {
        const bool tr = t < n, sr = s < n;
        const double re = table[2 * at], im = table[2 * at + 1];
        v = (tr == sr) ? re : (tr ? -im : im);
      }
      // signed (axis-derivative) kernels: the entry sign of (s, t)
      if (entry_sign(dim, q, perm + c * dim, sign + c * dim, sp_mask, ss, tt) < 0) v = -v;
    }
    M[i] = v;
  }
}

// Dense-kind payloads: M[b][t][s] from the payload's expanded blocks
// D[lev][case][s][t] (a NaN marks an entry without data: 0), with the
// complex embedding of expand_blocks_kernel.
__global__ void expand_dense_blocks_kernel(int n, int n_cases, int cplx, int MP, int KP,
                                           int n_blocks, int n_rg, int rg_size,
                                           const int32_t* __restrict__ blk_case,
                                           const int32_t* __restrict__ blk_lev,
                                           const double* __restrict__ D, double* __restrict__ M) {
  const int64_t total = (int64_t)n_blocks * MP * KP;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(i % KP);
    const int tl = (int)((i / KP) % MP);
    const int bg = (int)(i / ((int64_t)KP * MP));
    const int b = bg / n_rg, t = (bg % n_rg) * rg_size + tl;
    double v = 0.0;
    const int rows = cplx ? 2 * n : n;
    if (s < rows && tl < rg_size && t < rows) {
      const int64_t at = (((int64_t)blk_lev[b] * n_cases + blk_case[b]) * n + s % n) * n + t % n;
      if (!cplx) {
        v = D[at];
      } else {
        const bool tr = t < n, sr = s < n;
        const double re = D[2 * at], im = D[2 * at + 1];
        v = (tr == sr) ? re : (tr ? -im : im);
      }
      if (v != v) v = 0.0;
    }
    M[i] = v;
  }
}

// Helmholtz split blocks (complex, real 2n x 2n embedding) per (case, source
// level): the Laplace table part with the canonical scaling folded in, plus
// the quadrature of the split remainder on the q^d source nodes,
//   V[t][s] = alpha T[row(case, s, t)] + disp nmlz[s]
//             + w_s h^d R(|x_t - y_s|),
// x_t - y_s = vec h / 4 + size h (xi_t - 1/2) - h (xi_s - 1/2) (source box
// extent h, target box extent size * h, nearfield_potential_table.py:
// 917-946), R the series of split_series.h.  This is, block by block, the
// sum the point-to-point correction forms for every list-1 pair of that
// case and level (expansion_wrangler_fpnd.py:3491-3836) added to the table
// part (3457-3489), so one contraction evaluates both.  With an axis
// derivative wrapper (deriv_axis >= 0, 3884-3957) T is the wrapped Laplace
// table (signed entries, sp_mask) and R(|d|) becomes deriv_sign d_a R'(r)/r.
// Non-table part of fused split blocks beyond split order 1 or with the
// smooth quadrature, per (case-level block b, source mode s, target node t),
// complex:
//   sum over the quadrature nodes y_j of R_m(x_t - y_j) W_j L_s(y_j)
//   + sum over the tabulated terms of coeff [ (h/H)^(d+p) T_p(case, s, t)
//       + (2D log terms) log(h/H) sum_j |x_t - y_j|^p W_j L_s(y_j) ].
// The nodes are the source box's own (ms = 0: y_j = its node s, L = 1, W =
// its Gauss weight) or the ms^d smooth nodes with the Lagrange basis of
// mode s there.  Target / source offsets are those of expand_split_blocks_kernel.
constexpr int kSplitMaxBlockTerms = 16;

template <int DIM>
__global__ void split_extra_kernel(int q, int n, int n_keys, const int32_t* __restrict__ blk_case,
                                   const double* __restrict__ blk_h, double H,
                                   const int32_t* __restrict__ case_vecs,
                                   const uint8_t* __restrict__ perm, const int8_t* __restrict__ sign,
                                   const uint8_t* __restrict__ group,
                                   const uint16_t* __restrict__ ranks,
                                   const int32_t* __restrict__ offs, int64_t n_entries,
                                   const double* __restrict__ unodes,
                                   const double* __restrict__ uweights,
                                   const double* __restrict__ coef, int N, int ms,
                                   const double* __restrict__ snodes,
                                   const double* __restrict__ sweights,
                                   const double* __restrict__ sinterp, int n_terms,
                                   const double* __restrict__ term_tables,
                                   const int32_t* __restrict__ term_powers,
                                   const int32_t* __restrict__ term_logs,
                                   const double* __restrict__ term_coefs,
                                   int deriv_axis, double deriv_sign, int sp_mask,
                                   double2* __restrict__ extra) {
  constexpr int kMaxTerms = kSplitMaxBlockTerms;
  // under an axis derivative wrapper (deriv_axis >= 0): the remainder's
  // +-(x_a - y_a) R'(r)/r (coef in the R'/r layout, self pairs 0), the
  // wrapped term tables' signed entries at (h/H)^(d + p - 1), and the log
  // parts' derivative +-(x_a - y_a) p r^(p-2)
  const bool deriv = deriv_axis >= 0;
  const int64_t total = (int64_t)n_keys * n * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int ss = (int)(i % n), tt = (int)((i / n) % n), b = (int)(i / ((int64_t)n * n));
    const int c = blk_case[b];
    const double h = blk_h[b];
    const int32_t* vec = case_vecs + c * DIM;
    int m = 0;
    for (int a = 0; a < DIM; ++a) m = max(m, abs((int)vec[a]));
    const double size = m == 0 ? 1.0 : 0.5 * (m - 2);
    double xt[DIM];
    int sa[DIM];
    {
      int sr = ss, tr = tt;
      for (int a = DIM - 1; a >= 0; --a) {
        sa[a] = sr % q;
        xt[a] = 0.25 * vec[a] * h + size * h * (unodes[tr % q] - 0.5);
        sr /= q;
        tr /= q;
      }
    }
    double rr = 0.0, ri = 0.0;           // remainder quadrature
    double lq[kMaxTerms];                // quadrature of r^p for the log terms
    for (int k = 0; k < n_terms && k < kMaxTerms; ++k) lq[k] = 0.0;
    const int nodes = ms > 0 ? (DIM == 2 ? ms * ms : ms * ms * ms) : 1;
    for (int j = 0; j < nodes; ++j) {
      double r2 = 0.0, wl = 1.0, da = 0.0;
      int jr = j;
      for (int a = DIM - 1; a >= 0; --a) {
        double y, wa;
        if (ms > 0) {
          const int ja = jr % ms;
          jr /= ms;
          y = h * (snodes[ja] - 0.5);
          wa = sweights[ja] * h * sinterp[ja * q + sa[a]];
        } else {
          y = h * (unodes[sa[a]] - 0.5);
          wa = uweights[sa[a]] * h;
        }
        const double d = xt[a] - y;
        r2 = fma(d, d, r2);
        wl *= wa;
        if (a == deriv_axis) da = d;
      }
      double vr = 0.0, vi = 0.0;
      if (!deriv) {
        remainder_value<DIM>(r2, coef, N, vr, vi);
      } else if (r2 > 0.0) {
        remainder_radial_derivative<DIM>(r2, coef, N, vr, vi);
        vr *= deriv_sign * da;
        vi *= deriv_sign * da;
      }
      rr = fma(wl, vr, rr);
      ri = fma(wl, vi, ri);
      for (int k = 0; k < n_terms && k < kMaxTerms; ++k) {
        if (!term_logs[k]) continue;
        const int pw = term_powers[k];
        const double tv = !deriv ? pow(r2, 0.5 * pw)
                                 : (r2 > 0.0 ? deriv_sign * da * pw * pow(r2, 0.5 * pw - 1.0) : 0.0);
        lq[k] = fma(wl, tv, lq[k]);
      }
    }
    if (n_terms > 0) {
      const int64_t rel = compact_rel_row(DIM, q, perm + c * DIM, sign + c * DIM, group + c * DIM, ss, tt);
      const int64_t at = offs[ranks[c]] + rel;
      const double lg = log(h / H);
      const double tsg =
          entry_sign(DIM, q, perm + c * DIM, sign + c * DIM, sp_mask, ss, tt) < 0 ? -1.0 : 1.0;
      for (int k = 0; k < n_terms && k < kMaxTerms; ++k) {
        double v = tsg * pow(h / H, (double)(DIM + term_powers[k] - (deriv ? 1 : 0))) *
                   term_tables[(int64_t)k * n_entries + at];
        if (term_logs[k]) v = fma(lg, lq[k], v);
        rr = fma(term_coefs[2 * k], v, rr);
        ri = fma(term_coefs[2 * k + 1], v, ri);
      }
    }
    extra[i] = make_double2(rr, ri);
  }
}

__global__ void expand_split_blocks_kernel(int dim, int q, int n, int sp_mask, int deriv_axis,
                                           double deriv_sign, int MP, int KP, int n_blocks,
                                           int n_rg, int rg_size,
                                           const int32_t* __restrict__ blk_case,
                                           const int32_t* __restrict__ blk_lev, int64_t n_entries,
                                           const double* __restrict__ blk_alpha,
                                           const double* __restrict__ blk_disp,
                                           const double* __restrict__ blk_h,
                                           const int32_t* __restrict__ case_vecs,
                                           const uint8_t* __restrict__ perm,
                                           const int8_t* __restrict__ sign,
                                           const uint8_t* __restrict__ group,
                                           const uint16_t* __restrict__ ranks,
                                           const int32_t* __restrict__ offs,
                                           const double* __restrict__ table,
                                           const double* __restrict__ nmlz,
                                           const double* __restrict__ unodes,
                                           const double* __restrict__ uweights,
                                           const double* __restrict__ coef, int N,
                                           const double2* __restrict__ extra,
                                           double* __restrict__ M) {
  const int64_t total = (int64_t)n_blocks * MP * KP;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(i % KP);
    const int tl = (int)((i / KP) % MP);
    const int bg = (int)(i / ((int64_t)KP * MP));
    const int b = bg / n_rg, t = (bg % n_rg) * rg_size + tl;  // (case, level) block, its row
    double v = 0.0;
    if (s < 2 * n && tl < rg_size && t < 2 * n) {
      const int c = blk_case[b];
      const int ss = s % n, tt = t % n;
      const int64_t rel =
          compact_rel_row(dim, q, perm + c * dim, sign + c * dim, group + c * dim, ss, tt);
      // the block's table level (0 for one table)
      double tv = table[(int64_t)blk_lev[b] * n_entries + offs[ranks[c]] + rel];
      if (entry_sign(dim, q, perm + c * dim, sign + c * dim, sp_mask, ss, tt) < 0) tv = -tv;
      double re = blk_alpha[b] * tv + blk_disp[b] * nmlz[ss];
      if (extra) {  // split order > 1 or smooth quadrature: split_extra_kernel
        const double2 x = extra[((int64_t)b * n + tt) * n + ss];
        const bool tre = t < n, sre = s < n;
        re += x.x;
        v = (tre == sre) ? re : (tre ? -x.y : x.y);
        M[i] = v;
        continue;
      }
      const double h = blk_h[b];
      const int32_t* vec = case_vecs + c * dim;
      int m = 0;
      for (int a = 0; a < dim; ++a) m = max(m, abs((int)vec[a]));
      const double size = m == 0 ? 1.0 : 0.5 * (m - 2);
      int sr = ss, tr = tt;
      double r2 = 0.0, w = 1.0, da = 0.0;
      for (int a = dim - 1; a >= 0; --a) {
        const int sa = sr % q, ta = tr % q;
        sr /= q;
        tr /= q;
        const double d = 0.25 * vec[a] * h + size * h * (unodes[ta] - 0.5) - h * (unodes[sa] - 0.5);
        r2 = fma(d, d, r2);
        w *= uweights[sa] * h;
        if (a == deriv_axis) da = d;
      }
      double rr = 0.0, ri = 0.0;
      if (deriv_axis < 0) {
        if (dim == 2)
          remainder_value<2>(r2, coef, N, rr, ri);
        else
          remainder_value<3>(r2, coef, N, rr, ri);
      } else if (r2 > 0.0) {  // coincident nodes (r = 0) give 0
        if (dim == 2)
          remainder_radial_derivative<2>(r2, coef, N, rr, ri);
        else
          remainder_radial_derivative<3>(r2, coef, N, rr, ri);
        rr *= deriv_sign * da;
        ri *= deriv_sign * da;
      }
      re = fma(w, rr, re);
      const double im = w * ri;
      const bool tre = t < n, sre = s < n;
      v = (tre == sre) ? re : (tre ? -im : im);
    }
    M[i] = v;
  }
}

// Per block row: exponent e with |row| * 2^-e < 1, BF16 slices
// a_i * 2^(-8 i), a_i integer in [-128, 128], and the row scale 2^(e - 7).
// The slices are stored in MFMA fragment order: for row tile rt (16 rows),
// K chunk kc (32 columns) and slice i, the 64 lanes' 8-element operands of
// v_mfma_f32_16x16x32_bf16 (lane l: row l % 16, columns 8 (l / 16) + 0..7)
// form one contiguous 1 KiB block, [b][rt][kc][i][lane][8], so a wave's
// fragment load is a single fully coalesced access.  One wavefront per row.
__global__ void split_rows_kernel(int MP, int KP, int n_blocks, int S,
                                  const double* __restrict__ M, __bf16* __restrict__ A,
                                  double* __restrict__ rowscale) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (row >= n_blocks * MP) return;
  const int b = row / MP, t = row % MP;
  const double* m = M + (int64_t)row * KP;
  double mx = 0.0;
  for (int s = lane; s < KP; s += 64) mx = fmax(mx, fabs(m[s]));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1)
  if (lane == 0) rowscale[row] = mx > 0.0 ? ldexp(1.0, e - 7) : 0.0;
  for (int s = lane; s < KP; s += 64) {
    double r = mx > 0.0 ? ldexp(m[s], -e) : 0.0;  // |r| < 1
    for (int i = 0; i < S; ++i) {
      const double d = rint(ldexp(r, 7 + 8 * i));  // integer in [-128, 128]
      r -= ldexp(d, -7 - 8 * i);
      const int rt = t / 16, lane_ = (s % 32) / 8 * 16 + t % 16, kc = s / 32;
      A[((((int64_t)b * (MP / 16) + rt) * (KP / 32) + kc) * S + i) * 512 + lane_ * 8 + s % 8] =
          (__bf16)(float)ldexp(d, -8 * i);
    }
  }
}

// FP64 blocks in the fragment order of v_mfma_f64_16x16x4f64 (lane l holds
// row l % 16, column 4 k4 + l / 16): MF[b][rt][k4][lane].
__global__ void frag_f64_kernel(int MP, int KP, int n_blocks, const double* __restrict__ M,
                                double* __restrict__ MF) {
  const int64_t total = (int64_t)n_blocks * MP * KP;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(o % 64);
    const int64_t g = o / 64;
    const int k4 = (int)(g % (KP / 4));
    const int64_t bt = g / (KP / 4);
    const int rt = (int)(bt % (MP / 16));
    const int b = (int)(bt / (MP / 16));
    const int t = 16 * rt + lane % 16, k = 4 * k4 + lane / 16;
    MF[o] = M[((int64_t)b * MP + t) * KP + k];
  }
}

// Densities -> per (box, rhs) column: BF16 slices X[box][j][r][k] (mode
// order, zero padded to KP), column scale 2^(e - 7) and, for the 2D log term,
// nm[box][r] = sum_sid nmlz[sid] f[sbeg + sid].  Half a wavefront per column,
// 256-wide windows of k: the lanes load the window coalesced (lane l holds
// k = l + 32 u), pass it through LDS, and then each lane owns 8 consecutive
// k, so every slice goes out as one 16-byte store per lane.
constexpr int kColK = 8;
constexpr int kColBlock = 256;  // threads per block: 8 columns
__global__ void __launch_bounds__(kColBlock) split_cols_kernel(
    int n, int cplx, int KP, int nboxes, int R, int S, const double* __restrict__ f, int64_t ldf,
    const int32_t* __restrict__ mode_point, const int32_t* __restrict__ box_starts,
    const double* __restrict__ nmlz, int log_disp, __bf16* __restrict__ X,
    double* __restrict__ colscale, double* __restrict__ nmdot) {
  constexpr int kWin = 32 * kColK;  // k per window
  __shared__ __align__(16) double win[kColBlock / 32][kWin];
  const int col = (blockIdx.x * blockDim.x + threadIdx.x) / 32;
  const int hl = threadIdx.x % 32;  // lane within the half wavefront
  double* wrow = win[threadIdx.x / 32];
  const bool live = col < nboxes * R;
  const int box = live ? col / R : 0, r = live ? col % R : 0;
  const double* fr = f + (int64_t)r * ldf;
  const int rows = cplx ? 2 * n : n;
  // window w, coalesced: c[u] = component w + hl + 32 u of the column
  // ([Re f; Im f] if complex, zero past the rows)
  auto load = [&](int w, double (&c)[kColK]) {
#pragma unroll
    for (int u = 0; u < kColK; ++u) {
      const int s = w + hl + 32 * u;
      c[u] = 0.0;
      if (live && s < rows) {
        const int pt = mode_point[(int64_t)box * n + s % n];
        c[u] = cplx ? fr[2 * pt + s / n] : fr[pt];
      }
    }
  };
  double c[kColK];
  double mx = 0.0, dot = 0.0;
  for (int w = 0; w < KP; w += kWin) {
    load(w, c);
#pragma unroll
    for (int u = 0; u < kColK; ++u) {
      const int s = w + hl + 32 * u;
      mx = fmax(mx, fabs(c[u]));
      if (live && log_disp && s < n) dot += nmlz[s] * fr[box_starts[box] + s];
    }
  }
  for (int o = 16; o > 0; o >>= 1) {  // within the half wavefront
    mx = fmax(mx, __shfl_xor(mx, o, 32));
    dot += __shfl_xor(dot, o, 32);
  }
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  if (live && hl == 0) {
    colscale[col] = mx > 0.0 ? ldexp(1.0, e - 7) : 0.0;
    nmdot[col] = dot;
  }
  const int s0 = kColK * hl;
  for (int w0 = 0; w0 < KP; w0 += kWin) {
    if (KP > kWin) load(w0, c);  // one window: c still holds it
    __syncthreads();               // the previous window's reads are done
#pragma unroll
    for (int u = 0; u < kColK; ++u) wrow[hl + 32 * u] = c[u];
    __syncthreads();
    double v[kColK];
#pragma unroll
    for (int u = 0; u < kColK; u += 2) {
      const double2 d2 = *reinterpret_cast<const double2*>(wrow + s0 + u);
      v[u] = d2.x;
      v[u + 1] = d2.y;
    }
    if (!live || w0 + s0 >= KP) continue;
#pragma unroll
    for (int u = 0; u < kColK; ++u) v[u] = mx > 0.0 ? ldexp(v[u], -e) : 0.0;
    for (int j = 0; j < S; ++j) {
      uint32_t wd[kColK / 2];
#pragma unroll
      for (int u = 0; u < kColK; ++u) {
        const double d = rint(ldexp(v[u], 7 + 8 * j));
        v[u] -= ldexp(d, -7 - 8 * j);
        const uint32_t bits = __builtin_bit_cast(uint16_t, (__bf16)(float)ldexp(d, -8 * j));
        wd[u / 2] = (u % 2) ? (wd[u / 2] | (bits << 16)) : bits;
      }
      *reinterpret_cast<uint4*>(X + (((int64_t)box * S + j) * R + r) * KP + w0 + s0) =
          make_uint4(wd[0], wd[1], wd[2], wd[3]);
    }
  }
}

// Densities in mode order for the FP64 form: XF[box][r][k] (zero padded).
__global__ void gather_cols_f64_kernel(int n, int cplx, int KP, int nboxes, int R,
                                       const double* __restrict__ f, int64_t ldf,
                                       const int32_t* __restrict__ mode_point,
                                       const int32_t* __restrict__ box_starts,
                                       const double* __restrict__ nmlz, int log_disp,
                                       double* __restrict__ XF, double* __restrict__ nmdot) {
  const int col = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (col >= nboxes * R) return;
  const int box = col / R, r = col % R;
  const double* fr = f + (int64_t)r * ldf;
  const int rows = cplx ? 2 * n : n;
  double dot = 0.0;
  for (int s = lane; s < KP; s += 64) {
    double v = 0.0;
    if (s < rows) {
      const int pt = mode_point[(int64_t)box * n + s % n];
      v = cplx ? fr[2 * pt + s / n] : fr[pt];
    }
    XF[((int64_t)box * R + r) * KP + s] = v;
    if (log_disp && s < n) dot += nmlz[s] * fr[box_starts[box] + s];
  }
  for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
  if (lane == 0) nmdot[col] = dot;
}

struct TileParams {
  int n, MP, KP, S;      // n: block rows in use (2 x modes for complex data)
  int n_rg, rg_size;     // row groups per block and their rows: tile t holds group t % n_rg
  int nm, cplx;          // modes per box; complex data (interleaved re, im)
  int B, Rc, R;          // boxes per tile, RHS per tile, total RHS in the staged arrays
  int r0;                // first RHS of this launch
  int n_tiles;
  const int32_t* tile_step_start;  // [n_tiles + 1]
  const int32_t* tile_boxes;       // [n_tiles][B] target box or -1
  const int32_t* step_block;       // [n_steps]
  const int32_t* step_src;         // [n_steps][B] source box or -1
  const double* step_alpha;        // [n_steps][B]
  const double* step_disp;         // [n_steps][B]
  const int32_t* mode_point;       // [nboxes][n]
  const __bf16* A;                 // split form
  const double* rowscale;          // [n_blocks][MP]
  const __bf16* X;                 // [nboxes][S][R][KP]
  const double* colscale;          // [nboxes][R]
  const double* M;                 // dense blocks [n_blocks][MP][KP]
  const double* MF;                // dense blocks in FP64 MFMA fragment order
  const double* XF;                // [nboxes][R][KP]
  const double* nmdot;             // [nboxes][R]
  double* y;
  int64_t ldy;
  int32_t* phase;                  // [8] block each XCD is working on (hint)
  int use_phase;                   // start tiles at their XCD's current block
  // int8 split form
  const int8_t* A8;                // table digit slices, fragment order
  const int8_t* X8;                // [nboxes][KP/32][S][R][32] density digits
  const double* rho;               // [n_rg][MP] row scales
  const double* sigma;             // [R] column scales
  const double* box_disp;          // [nboxes] 2D log-term factor of a source box
  double* est_part;                // [n_tiles][64][4] error-estimate partials (optional)
  int nslot;                       // density ring slots
  int log_disp;                    // 2D log term in use
};

// Output element of block row `row` of target box `tb`: the row's mode is
// row % nm, and for complex data row / nm selects the real or imaginary part.
__device__ inline int64_t out_index(const TileParams& p, int tb, int row) {
  const int mp = p.mode_point[(int64_t)tb * p.nm + row % p.nm];
  return p.cplx ? 2 * (int64_t)mp + row / p.nm : (int64_t)mp;
}

// Phase alignment of the step loop.  The steps of a tile may run in any
// order (their sum is order independent), and every tile walks the same
// sorted block list, so a tile starts at the block its XCD's workgroups are
// currently on: concurrently running workgroups then share one table block
// in that XCD's L2 instead of spreading over all of them.  The phase word is
// only a hint (plain racy stores); correctness does not depend on its value,
// but every wave of a workgroup must walk the steps in the same order (they
// stage the density ring together): other workgroups rewrite the word while
// this one starts, so one thread reads it and the rotation goes to the
// other waves through LDS.  Called by every thread of the workgroup, in
// workgroup-uniform control flow (it holds a barrier).
__device__ inline int phase_start(const TileParams& p, int st0, int st1, int xcd) {
  const int nst = st1 - st0;
  if (nst <= 1 || !p.use_phase) return 0;
  __shared__ int s_rot;
  if (threadIdx.x == 0) {
    const int want = __builtin_nontemporal_load(p.phase + xcd);
    int lo = st0, hi = st1;  // first step with block >= want (blocks sorted)
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (p.step_block[mid] < want) lo = mid + 1; else hi = mid;
    }
    s_rot = lo < st1 ? lo - st0 : 0;
  }
  __syncthreads();
  return s_rot;
}

// XCD-aware tile order: consecutive workgroup ids go round-robin over the 8
// XCDs, so give each XCD a contiguous range of (spatially sorted) tiles.
__device__ inline int xcd_tile(int wg, int n) {
  const int per = (n + 7) / 8;
  const int xcd = wg % 8, k = wg / 8;
  const int t = xcd * per + k;
  return t;
}

// Stage the step's density image into LDS: every thread first issues all
// of its global loads (so they are in flight together), then stores them.
// The image is V vectors of 16 bytes; thread t owns vectors t, t + T, ...
template <int NV, typename LoadFn, typename StoreFn>
__device__ inline void staged_copy(int nvec, LoadFn load, StoreFn store) {
  // rounds of up to 8 vectors per thread keep the in-flight buffer small
  constexpr int kRound = NV < 8 ? NV : 8;
#pragma unroll
  for (int u0 = 0; u0 < NV; u0 += kRound) {
    uint4 buf[kRound];
#pragma unroll
    for (int u = 0; u < kRound; ++u) {
      const int v = threadIdx.x + (u0 + u) * blockDim.x;
      buf[u] = (u0 + u < NV && v < nvec) ? load(v) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kRound; ++u) {
      const int v = threadIdx.x + (u0 + u) * blockDim.x;
      if (u0 + u < NV && v < nvec) store(v, buf[u]);
    }
  }
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [0, N).
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

#include "split_i8.h"

// Split form, software-pipelined.  Workgroup = 8 waves, two per SIMD.  Wave
// w owns row tiles w and w + 8 below MP/16 and all 64 columns, so every
// table fragment is fetched once per workgroup.
//
// The (step, K chunk) iterations of a tile form one pipeline.  Densities
// stream through a double-buffered LDS ring one K chunk (S slices x 64
// columns x 32) at a time: the chunk for the next iteration -- which may
// belong to the next step -- is fetched into registers before the current
// chunk's MFMAs are issued and written to the other ring slot afterwards.
// Table fragments (MFMA fragment order, 1 KiB per wave load) for the next
// chunk are issued right after their last use in the current one.
// Slice products (i, j), i + j < S, accumulate into set min(i + j, 2):
// sets 0 and 1 are exact integer sums; all three are added in FP64, scaled
// by the row/column exponents and the pair scaling, at the end of a step
// into the FP64 output block, which lives in LDS (each lane owns its own
// words there, in MFMA output order) so that registers hold only the
// accumulator sets and operands.
//
// LDS: ring [2][S][64 columns][32] bf16, 16-byte vector v of column c stored
// at v ^ ((c >> 1) & 3) (conflict-free b128 reads without padding); output
// block [MP/16][4][4][64] doubles; column descriptors [2][64] x 3.
// static LDS of the BF16 and FP64 tile kernels (the phase word), rounded up
constexpr size_t kStaticLds = 16;
inline size_t split_lds_bytes(int S, int MP) {
  return (size_t)2 * S * kCols * 32 * 2 + (size_t)(MP / 16) * 1024 * 8 + 6 * kCols * 8;
}

template <int S, int NRT>
__global__ void __launch_bounds__(512) split_bf16_kernel(const TileParams p) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int RING = S * kCols * 32;  // bf16 per ring slot
  __bf16* sX = reinterpret_cast<__bf16*>(smem_raw);             // [2][S][64][32]
  double* sY = reinterpret_cast<double*>(sX + 2 * RING);         // [MP/16][4][4][64]
  double* sCol = sY + (size_t)(p.MP / 16) * 1024;                // [2][64]
  double* sDsp = sCol + 2 * kCols;                               // [2][64]
  int64_t* sBase = reinterpret_cast<int64_t*>(sDsp + 2 * kCols);  // [2][64]

  const int tile = xcd_tile(blockIdx.x, p.n_tiles);
  if (tile >= p.n_tiles) return;  // uniform over the workgroup
  // NRT = 2: wave w owns row tiles w and w + 8 (blocks of up to 256 rows);
  // NRT = 1: wave w owns row tile w (up to 128 rows), and a wave without
  // one skips the MFMA block as a whole (wave-uniform branch), so small
  // blocks issue no MFMAs on empty row-tile slots.
  const int wave = NRT == 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x) / 64 : threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int l16 = lane % 16, lg = lane / 16;
  const int ncols = p.B * p.Rc;
  const int nrt = NRT == 2 ? (p.MP / 16 - wave + 7) / 8 : (wave < p.MP / 16 ? 1 : 0);
  const int nkc = p.KP / 32;
  const int xcd = blockIdx.x % 8;
  const int st0 = p.tile_step_start[tile], st1 = p.tile_step_start[tile + 1];
  const int nst = st1 - st0;
  const int rot = phase_start(p, st0, st1, xcd);
  auto step_of = [&](int k) { int st = st0 + rot + k; return st >= st1 ? st - nst : st; };

  // column descriptors of a step into slot b: scale, displacement, base
  auto load_cols = [&](int st, int slot) {
    if (threadIdx.x < kCols) {
      const int c = threadIdx.x;
      double cs = 0.0, ds = 0.0;
      int64_t base = -1;
      if (c < ncols) {
        const int b = c / p.Rc, r = c % p.Rc;
        const int sb = p.step_src[st * p.B + b];
        if (sb >= 0) {
          const int64_t ci = (int64_t)sb * p.R + p.r0 + r;
          cs = p.colscale[ci] * p.step_alpha[st * p.B + b];
          ds = p.step_disp[st * p.B + b] * p.nmdot[ci];
          base = ((int64_t)sb * S * p.R + p.r0 + r) * p.KP;
        }
      }
      sCol[slot * kCols + c] = cs;
      sDsp[slot * kCols + c] = ds;
      sBase[slot * kCols + c] = base;
    }
  };
  // chunk kc of the density image of the step whose columns are in slot cs:
  // S x 64 columns x 4 vectors = 256 S vectors over 512 threads
  constexpr int NV = (S + 1) / 2;
  const int64_t jstride = (int64_t)p.R * p.KP;
  auto fetch_chunk = [&](int cs, int kc, uint4 (&buf)[NV]) {
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const int u = threadIdx.x + 512 * m;
      const int j = u / 256, col = (u % 256) / 4, v = u % 4;
      const int64_t base = j < S ? sBase[cs * kCols + col] : -1;
      buf[m] = base >= 0 ? *reinterpret_cast<const uint4*>(p.X + base + j * jstride + kc * 32 + v * 8)
                         : make_uint4(0, 0, 0, 0);
    }
  };
  auto store_chunk = [&](int ring, const uint4 (&buf)[NV]) {
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const int u = threadIdx.x + 512 * m;
      const int j = u / 256, col = (u % 256) / 4, v = u % 4;
      if (j < S)
        *reinterpret_cast<uint4*>(sX + ring * RING + ((size_t)j * kCols + col) * 32 +
                                  ((v ^ ((col >> 1) & 3)) * 8)) = buf[m];
    }
  };
  // this lane's word of output element (row tile rt16, column tile ct, reg e)
  auto yslot = [&](int rt16, int ct, int e) { return sY + ((rt16 * 4 + ct) * 4 + e) * 64 + lane; };

#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
    if (rt < nrt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int e = 0; e < 4; ++e) *yslot(wave + 8 * rt, ct, e) = 0.0;
  // sets 0, 1: exact diagonal sums; 2 (and 3 when NRT = 1): the tail, split
  // over two sets for one row tile per wave so that consecutive tail products
  // into one accumulator stay 8 MFMAs apart
  constexpr int NSET = NRT == 1 ? 4 : 3;
  floatx4 acc[NSET][NRT][4];
#pragma unroll
  for (int g = 0; g < NSET; ++g)
#pragma unroll
    for (int a = 0; a < NRT; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[g][a][b] = floatx4{0.f, 0.f, 0.f, 0.f};

  if (nst > 0) {
    // prologue: descriptors of the first step, its chunk 0 into ring 0
    load_cols(step_of(0), 0);
    __syncthreads();
    uint4 buf[NV];
    fetch_chunk(0, 0, buf);
    store_chunk(0, buf);
  }
  const int64_t frag_tile = (int64_t)nkc * S * 512;  // bf16 per (block, row tile)
  auto arow = [&](int blk, int rt) {
    const int rt16 = rt < nrt ? wave + 8 * rt : 0;
    return p.A + ((int64_t)blk * (p.MP / 16) + rt16) * frag_tile + lane * 8;
  };
  bf16x8 af[S][NRT];
  if (nst > 0) {
    const int blk = p.step_block[step_of(0)];
    if (threadIdx.x == 0) __builtin_nontemporal_store(blk, p.phase + xcd);
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
      const __bf16* ar = arow(blk, rt);
#pragma unroll
      for (int i = 0; i < S; ++i)
        if (rt < nrt) af[i][rt] = *reinterpret_cast<const bf16x8*>(ar + i * 512);
    }
  }
  __syncthreads();

  const int total = nst * nkc;
  // With two or more chunks per step the next step's column descriptors
  // are written one chunk early (chunk nkc - 2), so its first chunk is
  // fetched before this step's last MFMAs like any other chunk; with one
  // chunk per step they are written in that chunk and the fetch waits for
  // the barrier.
  const bool early = nkc >= 2;
  for (int it = 0; it < total; ++it) {
    const int k = it / nkc, kc = it % nkc;
    const int ring = it & 1, cslot = k & 1;
    const int st = step_of(k);
    const int blk = p.step_block[st];
    // where the next iteration's data comes from
    const bool has_next = it + 1 < total;
    const bool next_step = kc + 1 == nkc;
    const int nkc_ = next_step ? 0 : kc + 1;
    int nblk = blk;
    if (has_next && next_step) nblk = p.step_block[step_of(k + 1)];
    // slot cslot ^ 1 held step k - 1, whose last use (its step end) is done
    if (k + 1 < nst && (early ? kc + 2 == nkc : next_step))
      load_cols(step_of(k + 1), cslot ^ 1);  // read after the barriers below
    uint4 xnext[NV];
    if (has_next && (!next_step || early)) fetch_chunk(next_step ? cslot ^ 1 : cslot, nkc_, xnext);

    // MFMAs of this chunk.  Two orders, chosen by measurement per kernel:
    // - one row tile per wave (small blocks): table slice by table slice.
    //   Slot i meets density slices 0 .. S-1-i and is then refilled for the
    //   next chunk, so every refill has a whole chunk to arrive; density
    //   slice 0 stays in registers, the others are re-read from LDS per use
    //   (2.2x the LDS reads at S = 5; 12 % faster on 72-row complex blocks);
    // - two row tiles per wave: density slice by density slice, each read
    //   from LDS once; slot S-1-j is refilled after its last partner j (the
    //   table-slice order measured 5 % slower on 224-row blocks, where the
    //   extra LDS reads cost more than the earlier refills gain).
    // Column tiles are innermost in both: consecutive products into one
    // accumulator are 4 NRT MFMAs apart (closer dependent MFMAs stall).
    const __bf16* rX = sX + ring * RING;
    auto read_x = [&](int j, bf16x8 (&xf)[4]) {
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const int col = 16 * ct + l16;
        xf[ct] = *reinterpret_cast<const bf16x8*>(
            rX + ((size_t)j * kCols + col) * 32 + ((lg ^ ((col >> 1) & 3)) * 8));
      }
    };
    if constexpr (NRT == 1) {
      if (nrt > 0) {
        bf16x8 x0[4];
        read_x(0, x0);
        static_for<0, S>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          // keeps the re-reads below: without it the compiler holds every
          // density fragment of the chunk in registers
          asm volatile("" ::: "memory");
          static_for<0, S - i>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            // the tail alternates sets by density slice
            constexpr int g = (i + j) < 2 ? (i + j) : 2 + (j & 1);
            bf16x8 xt[4];
            if constexpr (j > 0) read_x(j, xt);
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt)
              if (rt < nrt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
                  acc[g][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                      af[i][rt], j == 0 ? x0[ct] : xt[ct], acc[g][rt][ct], 0, 0, 0);
          });
          if (has_next) {
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt)
              if (rt < nrt)
                af[i][rt] = *reinterpret_cast<const bf16x8*>(arow(nblk, rt) + ((int64_t)nkc_ * S + i) * 512);
          }
        });
      }
    } else {
      static_for<0, S>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        bf16x8 xf[4];
        read_x(j, xf);
        static_for<0, S - j>([&](auto ic) {
          constexpr int i = S - j - 1 - decltype(ic)::value;
          constexpr int g = (i + j) < 2 ? (i + j) : 2;
#pragma unroll
          for (int rt = 0; rt < NRT; ++rt)
            if (rt < nrt)
#pragma unroll
              for (int ct = 0; ct < 4; ++ct)
                acc[g][rt][ct] =
                    __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][rt], xf[ct], acc[g][rt][ct], 0, 0, 0);
        });
        // slot i = S-1-j had its last partner: refill it for the next chunk
        if (has_next) {
          constexpr int i = S - 1 - j;
#pragma unroll
          for (int rt = 0; rt < NRT; ++rt)
            if (rt < nrt)
              af[i][rt] = *reinterpret_cast<const bf16x8*>(arow(nblk, rt) + ((int64_t)nkc_ * S + i) * 512);
        }
      });
    }

    if (next_step) {
      // step done: add the three sets, scaled, to the FP64 block
      const double* rs = p.rowscale + (int64_t)blk * p.MP;
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) {
        if (rt >= nrt) continue;
        const int rt16 = wave + 8 * rt;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          const int c = 16 * ct + l16;
          const double cs = sCol[cslot * kCols + c], ds = sDsp[cslot * kCols + c];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double rsc = rs[16 * rt16 + 4 * lg + e];
            double v = ((double)acc[0][rt][ct][e] + (double)acc[1][rt][ct][e]) +
                       (double)acc[2][rt][ct][e];
            if constexpr (NSET == 4) v += (double)acc[3][rt][ct][e];
            double* ys = yslot(rt16, ct, e);
            *ys = fma(v * rsc, cs, *ys + ds);
          }
        }
      }
#pragma unroll
      for (int g = 0; g < NSET; ++g)
#pragma unroll
        for (int a = 0; a < NRT; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[g][a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
      if (has_next && threadIdx.x == 0) __builtin_nontemporal_store(nblk, p.phase + xcd);
    }
    // ring[ring ^ 1] was last read in the previous iteration, which ended
    // with a barrier; only the late descriptor path (one chunk per step)
    // needs one before the fetch reads them
    if (!early) __syncthreads();
    if (has_next) {
      if (next_step && !early) fetch_chunk(cslot ^ 1, 0, xnext);
      store_chunk(ring ^ 1, xnext);
    }
    __syncthreads();  // next chunk (and next descriptors) visible
  }
  // epilogue: write targets of the tile's boxes
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt) {
    if (rt >= nrt) continue;
    const int rt16 = wave + 8 * rt;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * rt16 + 4 * lg + e;
      const int grow = (tile % p.n_rg) * p.rg_size + row;  // row of the whole block
      if (row >= p.rg_size || grow >= p.n) continue;
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const int c = 16 * ct + l16;
        if (c >= ncols) continue;
        const int b = c / p.Rc, r = c % p.Rc;
        const int tb = p.tile_boxes[tile * p.B + b];
        if (tb < 0) continue;
        p.y[(int64_t)(p.r0 + r) * p.ldy + out_index(p, tb, grow)] = *yslot(rt16, ct, e);
      }
    }
  }
}

// Dense FP64 form, software-pipelined like the split kernel: 8 waves, wave
// w owns row tiles w and w + 8 (NRT = 2) or w (NRT = 1) and all 64 columns;
// FP64 accumulation in the MFMA accumulators (v_mfma_f64_16x16x4f64; lane l,
// register r of a 16 x 16 output tile holds row l/16 + 4 r, column l % 16).
// Densities stream through a double-buffered LDS ring one K chunk of 32 at
// a time, scaled by the pair scaling as they are stored; the column pitch is
// 33 doubles, so the 16 columns of a b64 read pass start 8 bytes apart mod
// 128 (conflict free).  Table fragments (fragment-ordered FP64 blocks, 512 B
// per wave load) for the next chunk are double-buffered in registers and
// loaded at the top of the chunk.  Each step ends by adding the column's
// 2D log displacement.
constexpr int kDPitch = 34;

template <int NRT>
__global__ void __launch_bounds__(512) dense_f64_kernel(const TileParams p) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* sX = reinterpret_cast<double*>(smem_raw);  // [2][64][kDPitch]
  double* sDsp = sX + 2 * kCols * kDPitch;           // [2][64]

  const int tile = xcd_tile(blockIdx.x, p.n_tiles);
  if (tile >= p.n_tiles) return;
  const int wave = NRT == 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x) / 64 : threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int l16 = lane % 16, lg = lane / 16;
  const int ncols = p.B * p.Rc;
  const int nrt = NRT == 2 ? (p.MP / 16 - wave + 7) / 8 : (wave < p.MP / 16 ? 1 : 0);
  const int nkc = p.KP / 32;
  const int xcd = blockIdx.x % 8;
  const int st0 = p.tile_step_start[tile], st1 = p.tile_step_start[tile + 1];
  const int nst = st1 - st0;
  const int rot = phase_start(p, st0, st1, xcd);
  auto step_of = [&](int k) { int st = st0 + rot + k; return st >= st1 ? st - nst : st; };

  // step displacements per column (2D log term), read at the step's end
  auto load_disp = [&](int st, int slot) {
    if (threadIdx.x < kCols) {
      const int c = threadIdx.x;
      double ds = 0.0;
      if (c < ncols) {
        const int b = c / p.Rc, r = c % p.Rc;
        const int sb = p.step_src[st * p.B + b];
        if (sb >= 0) ds = p.step_disp[st * p.B + b] * p.nmdot[(int64_t)sb * p.R + p.r0 + r];
      }
      sDsp[slot * kCols + c] = ds;
    }
  };
  // A thread moves 2 of the 1024 16-byte vectors of a chunk (64 columns x
  // 32 doubles).  It keeps its columns' density row offset and pair scaling
  // for the current step and, loaded one step ahead, for the next step in
  // registers, so a chunk fetch -- the next step's first one included --
  // never waits on a schedule load and needs no barrier.
  auto col_desc = [&](int st, int64_t (&off)[2], double (&al)[2]) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int col = (threadIdx.x + 512 * m) / 16;
      off[m] = -1;
      al[m] = 0.0;
      if (col < ncols) {
        const int b = col / p.Rc, r = col % p.Rc;
        const int sb = p.step_src[st * p.B + b];
        if (sb >= 0) {
          off[m] = ((int64_t)sb * p.R + p.r0 + r) * p.KP;
          al[m] = p.step_alpha[st * p.B + b];
        }
      }
    }
  };
  auto fetch_chunk = [&](const int64_t (&off)[2], int kc, uint4 (&buf)[2]) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int v = (threadIdx.x + 512 * m) % 16;
      buf[m] = off[m] >= 0 ? *reinterpret_cast<const uint4*>(p.XF + off[m] + kc * 32 + 2 * v)
                           : make_uint4(0, 0, 0, 0);
    }
  };
  auto store_chunk = [&](int ring, const uint4 (&buf)[2], const double (&al)[2]) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int u = threadIdx.x + 512 * m;
      const int col = u / 16, v = u % 16;
      const double2 d = *reinterpret_cast<const double2*>(&buf[m]);
      double* dst = sX + ((size_t)ring * kCols + col) * kDPitch + 2 * v;
      dst[0] = d.x * al[m];
      dst[1] = d.y * al[m];
    }
  };

  doublex4 yacc[NRT][4];
#pragma unroll
  for (int a = 0; a < NRT; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) yacc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};

  // fragment of (block, row tile slot rt, k group k4) for this lane; slots
  // past the wave's count read row tile 0 and are not used
  auto frag = [&](int blk, int rt, int k4) {
    const int rt16 = rt < nrt ? wave + 8 * rt : 0;
    return p.MF + (((int64_t)blk * (p.MP / 16) + rt16) * (p.KP / 4) + k4) * 64 + lane;
  };
  double ac[NRT][8], an[NRT][8];  // current / next chunk's table fragments
  int64_t offC[2] = {-1, -1}, offN[2] = {-1, -1};
  double alC[2] = {0.0, 0.0}, alN[2] = {0.0, 0.0};
  if (nst > 0) {
    load_disp(step_of(0), 0);
    col_desc(step_of(0), offC, alC);
    if (nst > 1) col_desc(step_of(1), offN, alN);
    uint4 buf[2];
    fetch_chunk(offC, 0, buf);
    store_chunk(0, buf, alC);
    const int blk = p.step_block[step_of(0)];
    if (threadIdx.x == 0) __builtin_nontemporal_store(blk, p.phase + xcd);
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
      for (int u = 0; u < 8; ++u) ac[rt][u] = *frag(blk, rt, u);
  }
  __syncthreads();

  const int total = nst * nkc;
  for (int it = 0; it < total; ++it) {
    const int k = it / nkc, kc = it % nkc;
    const int ring = it & 1, cslot = k & 1;
    const int blk = p.step_block[step_of(k)];
    const bool has_next = it + 1 < total;
    const bool next_step = kc + 1 == nkc;
    const int nkc_ = next_step ? 0 : kc + 1;
    int nblk = blk;
    if (has_next && next_step) {
      nblk = p.step_block[step_of(k + 1)];
      load_disp(step_of(k + 1), cslot ^ 1);  // read at the end of step k + 1
    }
    if (kc == 0 && k > 0) {  // step k begins: its descriptors, then load step k+1's
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        offC[m] = offN[m];
        alC[m] = alN[m];
      }
      if (k + 1 < nst) col_desc(step_of(k + 1), offN, alN);
    }
    uint4 xnext[2];
    if (has_next) fetch_chunk(next_step ? offN : offC, nkc_, xnext);
    if (has_next) {
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int u = 0; u < 8; ++u) an[rt][u] = *frag(nblk, rt, nkc_ * 8 + u);
    }
    const double* rX = sX + (size_t)ring * kCols * kDPitch;
    if (NRT == 2 || nrt > 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        double bx[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) bx[ct] = rX[(16 * ct + l16) * kDPitch + 4 * u + lg];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
          if (rt < nrt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
              yacc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[rt][u], bx[ct], yacc[rt][ct], 0, 0, 0);
      }
    }
    if (has_next) {
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int u = 0; u < 8; ++u) ac[rt][u] = an[rt][u];
    }
    if (next_step) {
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const double ds = sDsp[cslot * kCols + 16 * ct + l16];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
          for (int e = 0; e < 4; ++e) yacc[rt][ct][e] += ds;
      }
      if (has_next && threadIdx.x == 0) __builtin_nontemporal_store(nblk, p.phase + xcd);
    }
    // the other ring slot was last read in the previous iteration, which
    // every wave has left: write the next chunk, then one barrier
    if (has_next) store_chunk(ring ^ 1, xnext, next_step ? alN : alC);
    __syncthreads();
  }
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt) {
    if (rt >= nrt) continue;
    const int rt16 = wave + 8 * rt;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * rt16 + lg + 4 * e;
      const int grow = (tile % p.n_rg) * p.rg_size + row;  // row of the whole block
      if (row >= p.rg_size || grow >= p.n) continue;
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const int c = 16 * ct + l16;
        if (c >= ncols) continue;
        const int b = c / p.Rc, r = c % p.Rc;
        const int tb = p.tile_boxes[tile * p.B + b];
        if (tb < 0) continue;
        p.y[(int64_t)(p.r0 + r) * p.ldy + out_index(p, tb, grow)] = yacc[rt][ct][e];
      }
    }
  }
}

}  // namespace vl1

using namespace vl1;

// ------------------------------------------------------------------- plan
struct TileSchedule {
  int B = 0, Rc = 0, n_tiles = 0, max_steps = 0;
  DevBuf tile_step_start, tile_boxes, step_block, step_src, step_alpha, step_disp;
  int64_t n_steps = 0;
};

// The dynamic-LDS limit of a tile kernel is raised to the full 160 KiB (less
// the kernel's static LDS, the phase word's) once per device (bit d of
// `done`), so later launches -- including ones recorded into a HIP graph --
// make no attribute calls.
static hipError_t allow_full_lds(const void* fn, std::atomic<unsigned>& done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const unsigned bit = 1u << (dev & 31);
  if (done.load(std::memory_order_acquire) & bit) return hipSuccess;
  hipFuncAttributes fa{};
  e = hipFuncGetAttributes(&fa, fn);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                          160 * 1024 - (int)fa.sharedSizeBytes);
  if (e == hipSuccess) done.fetch_or(bit, std::memory_order_acq_rel);
  return e;
}

template <int S, int NRT>
static hipError_t launch_split_t(dim3 grid, size_t lds, hipStream_t s, const TileParams& prm) {
  static std::atomic<unsigned> attr_done{0};
  hipError_t e = allow_full_lds((const void*)split_bf16_kernel<S, NRT>, attr_done);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((split_bf16_kernel<S, NRT>), grid, dim3(512), lds, s, prm);
  return hipGetLastError();
}

static hipError_t launch_split(int S, bool one, dim3 grid, size_t lds, hipStream_t s,
                               const TileParams& prm) {
  switch (S * 2 + (one ? 1 : 0)) {
    case 2: return launch_split_t<1, 2>(grid, lds, s, prm);
    case 3: return launch_split_t<1, 1>(grid, lds, s, prm);
    case 4: return launch_split_t<2, 2>(grid, lds, s, prm);
    case 5: return launch_split_t<2, 1>(grid, lds, s, prm);
    case 6: return launch_split_t<3, 2>(grid, lds, s, prm);
    case 7: return launch_split_t<3, 1>(grid, lds, s, prm);
    case 8: return launch_split_t<4, 2>(grid, lds, s, prm);
    case 9: return launch_split_t<4, 1>(grid, lds, s, prm);
    case 10: return launch_split_t<5, 2>(grid, lds, s, prm);
    case 11: return launch_split_t<5, 1>(grid, lds, s, prm);
    case 12: return launch_split_t<6, 2>(grid, lds, s, prm);
    default: return launch_split_t<6, 1>(grid, lds, s, prm);
  }
}

template <int S>
static hipError_t launch_i8_t(dim3 grid, dim3 block, size_t lds, hipStream_t s, const TileParams& prm) {
  static std::atomic<unsigned> attr_done{0};
  hipError_t e = allow_full_lds((const void*)split_i8_kernel<S>, attr_done);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((split_i8_kernel<S>), grid, block, lds, s, prm);
  return hipGetLastError();
}

static hipError_t launch_i8(int S, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                            const TileParams& prm) {
  switch (S) {
    case 1: return launch_i8_t<1>(grid, block, lds, s, prm);
    case 2: return launch_i8_t<2>(grid, block, lds, s, prm);
    case 3: return launch_i8_t<3>(grid, block, lds, s, prm);
    case 4: return launch_i8_t<4>(grid, block, lds, s, prm);
    case 5: return launch_i8_t<5>(grid, block, lds, s, prm);
    default: return launch_i8_t<6>(grid, block, lds, s, prm);
  }
}

struct vl1_plan {
  const vl1_payload* payload = nullptr;
  int dim = 0, n = 0, MP = 0, KP = 0;
  int cplx = 0, rows = 0;  // complex data; block rows in use (n, or 2n if complex)
  int nboxes = 0, n_tgt = 0;
  int64_t n_points = 0, n_pairs = 0;
  int log_disp = 0;
  // host pair data
  std::vector<int32_t> tgt_boxes, pair_tgt, pair_src, pair_block;
  std::vector<uint64_t> tgt_morton;
  std::vector<double> pair_alpha, pair_disp;
  int n_blocks = 0;  // plan blocks: n_keys (case, level) blocks x n_rg row groups
  int n_keys = 0, n_rg = 1, rg_size = 0;
  bool fused_split = false;  // Helmholtz split blocks (table part + remainder)
  std::vector<int32_t> blk_case, blk_lev;
  std::vector<double> blk_alpha, blk_disp, blk_h;  // Helmholtz split blocks only
  // device
  DevBuf mode_point, box_starts, M, MF, rowscale;
  // split slices of the blocks, one set per slice count used (a graph
  // captured at one count keeps reading its own set)
  std::map<int, DevBuf*> A_by_slices;
  // int8 split form: digit slices per slice count, row scales [n_rg][MP],
  // and per source box the pair scaling alpha and the 2D log-term factor
  std::map<int, DevBuf*> A8_by_slices;
  DevBuf rho, box_alpha, box_disp;
  // streaming column maxima (no 2D log term): per point the index of its
  // source box's scaling in alpha_lut (255: no list-1 source)
  DevBuf pt_class, alpha_lut;
  bool stream_colmax = false;
  std::map<int, TileSchedule*> schedules;  // by 2 Rc + (Morton-only order)
  std::map<int, int> tuned;                 // variant by (form, slices, Rc), from vl1_plan_tune
  int forced_variant = -1;                  // candidate under test in vl1_plan_tune
  // staging buffers (grown on demand).  A buffer that is outgrown is
  // retired, not freed, until the plan is destroyed: a HIP graph captured
  // from an earlier evaluation holds its address.
  DevBuf X, colscale, nmdot, XF, X8, sigma, est_part;
  std::vector<DevBuf*> retired;
  DevBuf phase;  // int32 [8]
  int staged_form = -1, staged_rhs = 0, staged_slices = -1;
  ~vl1_plan() {
    for (auto& kv : schedules) delete kv.second;
    for (auto& kv : A_by_slices) delete kv.second;
    for (auto& kv : A8_by_slices) delete kv.second;
    for (DevBuf* b : retired) delete b;
  }
  // Ensure `b` holds at least `bytes`; a smaller buffer in use moves to
  // `retired` first.
  hipError_t grow(DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes) return hipSuccess;
    if (b.ptr) {
      auto old = new DevBuf();
      old->ptr = b.ptr;
      old->bytes = b.bytes;
      b.ptr = nullptr;
      b.bytes = 0;
      retired.push_back(old);
    }
    return b.alloc(bytes);
  }
  const DevBuf* A(int S) const {
    auto it = A_by_slices.find(S);
    return it == A_by_slices.end() ? nullptr : it->second;
  }
  const DevBuf* A8(int S) const {
    auto it = A8_by_slices.find(S);
    return it == A8_by_slices.end() ? nullptr : it->second;
  }
  // slices the form needs (null: not built)
  const DevBuf* slices(int form, int S) const {
    return form == VL1_FORM_SPLIT_I8 ? A8(S) : (form == VL1_FORM_SPLIT_BF16 ? A(S) : nullptr);
  }
};

static int build_schedule(vl1_plan* pl, int Rc, bool morton_only, TileSchedule** out,
                          hipStream_t s) {
  const int key = Rc * 2 + (morton_only ? 1 : 0);
  auto it = pl->schedules.find(key);
  if (it != pl->schedules.end()) {
    *out = it->second;
    return VL1_OK;
  }
  const int B = std::max(1, kCols / Rc);
  // per target: sorted list of (block, pair)
  std::vector<std::vector<std::pair<int, int>>> per_tgt(pl->n_tgt);
  for (int64_t k = 0; k < pl->n_pairs; ++k) per_tgt[pl->pair_tgt[k]].push_back({pl->pair_block[k], (int)k});
  for (auto& v : per_tgt) std::sort(v.begin(), v.end());
  // Tiles follow the Morton order of the target boxes: concurrently running
  // tiles then share most source boxes (L2 reuse of their density slices).
  // With several boxes per tile, boxes are first grouped by their block
  // signature (the sorted list of their (case, level) blocks), so the boxes
  // of a tile need the same blocks and every step fills all their columns
  // (on adaptive trees Morton neighbours often differ in level or case).
  std::vector<uint64_t> sig(pl->n_tgt, 0);
  if (B > 1 && !morton_only) {
    for (int i = 0; i < pl->n_tgt; ++i) {
      uint64_t h = 1469598103934665603ull;  // FNV-1a over the block ids
      for (auto& pr : per_tgt[i]) {
        h ^= (uint64_t)(uint32_t)pr.first;
        h *= 1099511628211ull;
      }
      sig[i] = h;
    }
  }
  std::vector<int> order(pl->n_tgt);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (sig[a] != sig[b]) return sig[a] < sig[b];
    return pl->tgt_morton[a] < pl->tgt_morton[b];
  });
  auto sch = new TileSchedule();
  sch->B = B;
  sch->Rc = Rc;
  // one tile per (group of B target boxes, row group); the row groups of a
  // box group are consecutive tiles (same XCD: they share density images)
  const int n_box_tiles = (pl->n_tgt + B - 1) / B;
  const int G = pl->n_rg;
  sch->n_tiles = n_box_tiles * G;
  std::vector<int32_t> tss(sch->n_tiles + 1, 0), tboxes((size_t)sch->n_tiles * B, -1);
  std::vector<int32_t> sblk, ssrc;
  std::vector<double> salpha, sdisp;
  for (int bt = 0; bt < n_box_tiles; ++bt) {
    std::vector<int> blocks;  // (case, level) keys
    std::vector<int32_t> boxes(B, -1);
    for (int b = 0; b < B; ++b) {
      const int idx = bt * B + b;
      if (idx >= pl->n_tgt) break;
      const int tg = order[idx];
      boxes[b] = pl->tgt_boxes[tg];
      for (auto& pr : per_tgt[tg]) blocks.push_back(pr.first);
    }
    std::sort(blocks.begin(), blocks.end());
    blocks.erase(std::unique(blocks.begin(), blocks.end()), blocks.end());
    std::vector<int32_t> src_l;
    std::vector<double> al_l, ds_l;
    for (int blk : blocks) {
      for (int b = 0; b < B; ++b) {
        const int idx = bt * B + b;
        int src = -1;
        double al = 0.0, ds = 0.0;
        if (idx < pl->n_tgt) {
          const auto& v = per_tgt[order[idx]];
          auto f = std::lower_bound(v.begin(), v.end(), std::make_pair(blk, -1));
          if (f != v.end() && f->first == blk) {
            src = pl->pair_src[f->second];
            al = pl->pair_alpha[f->second];
            ds = pl->pair_disp[f->second];
          }
        }
        src_l.push_back(src);
        al_l.push_back(al);
        ds_l.push_back(ds);
      }
    }
    for (int g = 0; g < G; ++g) {
      const int t = bt * G + g;
      for (int b = 0; b < B; ++b) tboxes[(size_t)t * B + b] = boxes[b];
      for (int blk : blocks) sblk.push_back(blk * G + g);  // plan block of this row group
      ssrc.insert(ssrc.end(), src_l.begin(), src_l.end());
      salpha.insert(salpha.end(), al_l.begin(), al_l.end());
      sdisp.insert(sdisp.end(), ds_l.begin(), ds_l.end());
      tss[t + 1] = (int32_t)sblk.size();
      sch->max_steps = std::max(sch->max_steps, tss[t + 1] - tss[t]);
    }
  }
  sch->n_steps = (int64_t)sblk.size();
  hipError_t e = upload(sch->tile_step_start, tss.data(), tss.size() * 4, s);
  if (e == hipSuccess) e = upload(sch->tile_boxes, tboxes.data(), tboxes.size() * 4, s);
  if (e == hipSuccess) e = upload(sch->step_block, sblk.data(), sblk.size() * 4, s);
  if (e == hipSuccess) e = upload(sch->step_src, ssrc.data(), ssrc.size() * 4, s);
  if (e == hipSuccess) e = upload(sch->step_alpha, salpha.data(), salpha.size() * 8, s);
  if (e == hipSuccess) e = upload(sch->step_disp, sdisp.data(), sdisp.size() * 8, s);
  if (e != hipSuccess) {
    delete sch;
    return fail(VL1_E_HIP, std::string("plan schedule upload: ") + hipGetErrorString(e));
  }
  pl->schedules[key] = sch;
  *out = sch;
  return VL1_OK;
}

static int ensure_A(vl1_plan* pl, int S, hipStream_t s) {
  if (pl->A(S)) return VL1_OK;
  auto A = new DevBuf();
  hipError_t e = A->alloc((size_t)pl->n_blocks * S * pl->MP * pl->KP * sizeof(__bf16));
  if (e == hipSuccess && pl->rowscale.bytes == 0)
    e = pl->rowscale.alloc((size_t)pl->n_blocks * pl->MP * sizeof(double));
  if (e != hipSuccess) {
    delete A;
    return fail(VL1_E_HIP, std::string("plan A alloc: ") + hipGetErrorString(e));
  }
  const int rows = pl->n_blocks * pl->MP;
  const int wpb = 4;
  // the row scales do not depend on S: rewriting them leaves earlier slice
  // sets (and graphs reading them) valid
  hipLaunchKernelGGL(split_rows_kernel, dim3((rows + wpb - 1) / wpb), dim3(64 * wpb), 0, s, pl->MP,
                     pl->KP, pl->n_blocks, S, pl->M.as<double>(), A->as<__bf16>(),
                     pl->rowscale.as<double>());
  e = hipGetLastError();
  if (e != hipSuccess) {
    delete A;
    return fail(VL1_E_HIP, std::string("split_rows: ") + hipGetErrorString(e));
  }
  pl->A_by_slices[S] = A;
  return VL1_OK;
}

// int8 digit slices for S (and, once, the row scales).
static int ensure_A8(vl1_plan* pl, int S, hipStream_t s) {
  if (pl->A8(S)) return VL1_OK;
  hipError_t e = hipSuccess;
  const int G = pl->n_rg;
  if (pl->rho.bytes == 0) {
    e = pl->rho.alloc((size_t)G * pl->MP * sizeof(double));
    if (e == hipSuccess) {
      const int rows = G * pl->MP, wpb = 4;
      hipLaunchKernelGGL(i8_rowmax_kernel, dim3((rows + wpb - 1) / wpb), dim3(64 * wpb), 0, s,
                         pl->MP, pl->KP, pl->n_keys, G, pl->M.as<double>(), pl->rho.as<double>());
      e = hipGetLastError();
    }
    if (e != hipSuccess) {
      pl->rho.release();
      return fail(VL1_E_HIP, std::string("plan row scales: ") + hipGetErrorString(e));
    }
  }
  auto A = new DevBuf();
  e = A->alloc((size_t)pl->n_blocks * S * pl->MP * pl->KP);
  if (e == hipSuccess) {
    const int64_t total = (int64_t)pl->n_blocks * pl->MP * (pl->KP / 16);
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    switch (S) {
#define VL1_I8_ROWS(SV)                                                                          \
  case SV:                                                                                       \
    hipLaunchKernelGGL(i8_split_rows_kernel<SV>, dim3(grid), dim3(256), 0, s, pl->MP, pl->KP,   \
                       pl->n_blocks, G, pl->M.as<double>(), pl->rho.as<double>(),                 \
                       A->as<int8_t>());                                                          \
    break;
      VL1_I8_ROWS(1) VL1_I8_ROWS(2) VL1_I8_ROWS(3) VL1_I8_ROWS(4) VL1_I8_ROWS(5) VL1_I8_ROWS(6)
#undef VL1_I8_ROWS
    }
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    delete A;
    return fail(VL1_E_HIP, std::string("plan int8 slices: ") + hipGetErrorString(e));
  }
  pl->A8_by_slices[S] = A;
  return VL1_OK;
}

static int plan_create(vl1_plan** out, const vl1_payload* pay, const vl1_plan_desc* d,
                       const vl1_helmholtz_split_blocks* sp, void* stream) {
  clear_error();
  if (!out || !pay || !d) return fail(VL1_E_ARG, "vl1_plan_create: null argument");
  *out = nullptr;
  if (d->dim != pay->dim) return fail(VL1_E_ARG, "vl1_plan_create: dimension mismatch");
  if (d->n_boxes < 1 || d->n_tgt_boxes < 0 || d->n_points < 1)
    return fail(VL1_E_ARG, "vl1_plan_create: bad sizes");
  if (!d->box_centers || !d->box_levels || !d->box_starts || !d->box_counts || !d->target_boxes ||
      !d->nbr_starts || !d->nbr_lists || !d->local_ids)
    return fail(VL1_E_ARG, "vl1_plan_create: null host array");
  if (d->scaling_mode < 0 || d->scaling_mode > 3)
    return fail(VL1_E_ARG, "vl1_plan_create: bad scaling mode");
  bool custom_disp = false;
  if (d->scaling_mode == VL1_SCALING_CUSTOM) {
    if (!d->level_scaling || !d->level_disp)
      return fail(VL1_E_ARG, "vl1_plan_create: custom scaling needs level_scaling and level_disp");
    for (int l = 0; l < VL1_MAX_LEVELS; ++l) {
      if (!std::isfinite(d->level_scaling[l]) || !std::isfinite(d->level_disp[l]))
        return fail(VL1_E_ARG, "vl1_plan_create: custom scaling values must be finite");
      custom_disp = custom_disp || d->level_disp[l] != 0.0;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  const int n = pay->n, dim = pay->dim;
  // Helmholtz split: complex blocks over a real Laplace payload, the
  // canonical scaling and the 2D log term folded into per-(case, level) blocks
  const bool split = sp != nullptr;
  if (split) {
    if (sp->deriv_kind < VL1_SPLIT_DERIV_NONE || sp->deriv_kind > VL1_SPLIT_DERIV_SOURCE ||
        (sp->deriv_kind != VL1_SPLIT_DERIV_NONE && (sp->deriv_axis < 0 || sp->deriv_axis >= dim)))
      return fail(VL1_E_ARG, "Helmholtz split plans: bad derivative wrapper");
    // a derivative wrapper takes the wrapped (signed) Laplace table
    if (pay->is_complex || (pay->sign_mask && sp->deriv_kind == VL1_SPLIT_DERIV_NONE))
      return fail(VL1_E_ARG, "Helmholtz split plans take a real scalar (Laplace) table payload");
    if (pay->dense_kind)
      return fail(VL1_E_UNSUPPORTED,
                  "Helmholtz split plans take an arithmetic-orbit payload, not the dense kind");
    // the canonical single-table scaling, or per-level tables (no scaling,
    // no log term) at split order 1 (the term tables are single-table)
    if (d->scaling_mode != VL1_SCALING_CANONICAL &&
        !(d->scaling_mode == VL1_SCALING_PER_LEVEL && sp->n_terms == 0))
      return fail(VL1_E_ARG,
                  "Helmholtz split plans use the canonical single-table scaling, or per-level "
                  "tables at split order 1");
    if (sp->series_nmax < 1 || sp->series_nmax > kSplitMaxTerms)
      return fail(VL1_E_ARG, "Helmholtz split plans: series_nmax must be 1..96");
    if (!sp->unit_nodes || !sp->unit_weights || !sp->case_vecs)
      return fail(VL1_E_ARG, "Helmholtz split plans: null host array");
    const int order = sp->split_order < 1 ? 1 : sp->split_order;
    if (2 * (order - 1) > kSplitMaxTerms)
      return fail(VL1_E_ARG, "Helmholtz split plans: split_order must be 1..49");
    if (sp->n_terms != order - 1)
      return fail(VL1_E_ARG, "Helmholtz split plans: n_terms must be split_order - 1");
    if (sp->n_terms > kSplitMaxBlockTerms)
      return fail(VL1_E_UNSUPPORTED, "Helmholtz split plans: at most " +
                                         std::to_string(kSplitMaxBlockTerms) +
                                         " tabulated terms (split order 17)");
    if (sp->n_terms > 0 && (!sp->term_tables || !sp->term_powers || !sp->term_logs ||
                            !sp->term_coefs || sp->n_term_entries != pay->n_entries))
      return fail(VL1_E_ARG,
                  "Helmholtz split plans: term tables must be given in the payload's compact "
                  "layout (n_term_entries = its entry count)");
    for (int k = 0; k < sp->n_terms; ++k)
      if (sp->term_powers[k] < 0 || sp->term_powers[k] > 2 * kSplitMaxTerms ||
          (sp->term_logs[k] != 0 && sp->term_logs[k] != 1) || (sp->term_logs[k] && dim != 2))
        return fail(VL1_E_ARG, "Helmholtz split plans: bad term power or log flag");
    if (sp->smooth_order < 0 || sp->smooth_order > 64 ||
        (sp->smooth_order > 0 && (!sp->smooth_nodes || !sp->smooth_weights || !sp->smooth_interp)))
      return fail(VL1_E_ARG, "Helmholtz split plans: smooth_order must be 0..64 with its nodes, "
                             "weights and interpolation matrix");
  }
  auto pl = new vl1_plan();
  pl->payload = pay;
  pl->dim = dim;
  pl->n = n;
  pl->fused_split = split;
  pl->cplx = (pay->is_complex || split) ? 1 : 0;
  pl->rows = pl->cplx ? 2 * n : n;
  // tiles take at most kMaxTileRows block rows (the split form's LDS budget
  // at 5 slices): larger blocks are cut into row groups, one tile per group
  pl->n_rg = (pl->rows + kMaxTileRows - 1) / kMaxTileRows;
  pl->rg_size = (pl->rows + pl->n_rg - 1) / pl->n_rg;
  pl->MP = round_up(pl->rg_size, 32);
  pl->KP = round_up(pl->rows, 32);
  pl->nboxes = d->n_boxes;
  pl->n_tgt = d->n_tgt_boxes;
  pl->n_points = d->n_points;
  const bool want_log =
      (d->scaling_mode == VL1_SCALING_CANONICAL && d->log_displacement) || custom_disp;
  pl->log_disp = (want_log && !split) ? 1 : 0;  // split blocks carry the term themselves
  if (pl->log_disp && pay->has_missing) {
    // the blocks add the log term for every (s, t); the reference skips it
    // where an entry is missing
    delete pl;
    return fail(VL1_E_UNSUPPORTED,
                "vl1_plan_create: a dense payload with missing entries and the 2D log "
                "displacement evaluates through vl1_near_field_f64");
  }
  if (pl->cplx && pl->log_disp) {
    delete pl;
    return fail(VL1_E_UNSUPPORTED, "vl1_plan_create: the 2D log displacement applies to real tables only");
  }

  // CSR and box ids are host arrays from the caller: range-check them
  // before anything is indexed with them
  if (d->nbr_starts[0] != 0) {
    delete pl;
    return fail(VL1_E_ARG, "vl1_plan_create: neighbor source starts must begin at 0");
  }
  for (int i = 0; i < d->n_tgt_boxes; ++i) {
    const int tb = d->target_boxes[i];
    if (tb < 0 || tb >= d->n_boxes) {
      delete pl;
      return fail(VL1_E_ARG, "vl1_plan_create: target box id " + std::to_string(tb) +
                                 " out of range [0, " + std::to_string(d->n_boxes) + ")");
    }
    if (d->nbr_starts[i + 1] < d->nbr_starts[i]) {
      delete pl;
      return fail(VL1_E_ARG, "vl1_plan_create: neighbor source starts must be non-decreasing");
    }
  }

  // layout: every listed box holds exactly n particles whose local ids are a
  // permutation of 0..n-1 (expansion_wrangler_fpnd.py:427-463)
  std::vector<int32_t> mode_point((size_t)d->n_boxes * n, -1);
  std::vector<char> used(d->n_boxes, 0);
  for (int i = 0; i < d->n_tgt_boxes; ++i) used[d->target_boxes[i]] = 1;
  for (int64_t k = 0; k < d->nbr_starts[d->n_tgt_boxes]; ++k) {
    const int b = d->nbr_lists[k];
    if (b < 0 || b >= d->n_boxes) {
      delete pl;
      return fail(VL1_E_ARG, "vl1_plan_create: source box id out of range");
    }
    used[b] = 1;
  }
  for (int b = 0; b < d->n_boxes; ++b) {
    if (!used[b]) continue;
    if (d->box_levels[b] > 30) {
      delete pl;
      return fail(VL1_E_ARG, "vl1_plan_create: box level above 30");
    }
    if (d->box_counts[b] != n) {
      delete pl;
      return fail(VL1_E_LAYOUT, "table-based near-field evaluation requires exactly " +
                                    std::to_string(n) + " quadrature points per active box");
    }
    for (int k = 0; k < n; ++k) {
      const int64_t pt = (int64_t)d->box_starts[b] + k;
      if (pt < 0 || pt >= d->n_points) {
        delete pl;
        return fail(VL1_E_LAYOUT, "box particle range outside the point array");
      }
      const int m = d->local_ids[pt];
      if (m < 0 || m >= n || mode_point[(size_t)b * n + m] >= 0) {
        delete pl;
        return fail(VL1_E_LAYOUT, "box-local node ids must be a permutation of 0..q^d-1");
      }
      if (split && want_log && m != k) {
        // the log term pairs nmlz[sid] with the sid-th particle: folded into
        // a block it needs particle k to carry mode k
        delete pl;
        return fail(VL1_E_UNSUPPORTED,
                    "2D Helmholtz split plans need box-local node ids in mode order");
      }
      mode_point[(size_t)b * n + m] = (int32_t)pt;
    }
  }
  for (auto& v : mode_point)
    if (v < 0) v = 0;  // unused boxes: never read

  // pairs: case id by the reference's float encoding, level, scaling
  const int64_t n_pairs = d->nbr_starts[d->n_tgt_boxes];
  pl->n_pairs = n_pairs;
  std::map<std::pair<int, int>, int> blk_of;
  const double H = d->table_root_extent;
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < d->n_tgt_boxes; ++i)
      for (int a = 0; a < dim; ++a) {
        const double c = d->box_centers[(int64_t)a * d->n_boxes + d->target_boxes[i]];
        lo[a] = std::min(lo[a], c);
        hi[a] = std::max(hi[a], c);
      }
    pl->tgt_morton.resize(d->n_tgt_boxes);
    for (int i = 0; i < d->n_tgt_boxes; ++i) {
      uint64_t code = 0;
      uint32_t u[3] = {0, 0, 0};
      for (int a = 0; a < dim; ++a) {
        const double c = d->box_centers[(int64_t)a * d->n_boxes + d->target_boxes[i]];
        const double w = hi[a] > lo[a] ? (c - lo[a]) / (hi[a] - lo[a]) : 0.0;
        u[a] = (uint32_t)std::min(1048575.0, std::floor(w * 1048575.0));
      }
      for (int bit = 19; bit >= 0; --bit)
        for (int a = 0; a < dim; ++a) code = (code << 1) | ((u[a] >> bit) & 1u);
      pl->tgt_morton[i] = code;
    }
  }
  for (int i = 0; i < d->n_tgt_boxes; ++i) {
    const int tb = d->target_boxes[i];
    pl->tgt_boxes.push_back(tb);
    for (int64_t k = d->nbr_starts[i]; k < d->nbr_starts[i + 1]; ++k) {
      const int sb = d->nbr_lists[k];
      const int lvl = d->box_levels[sb];
      const double sext = d->root_extent * (1.0 / (double)(1 << lvl));
      double code = 0.0;
      for (int a = 0; a < dim; ++a)
        code = code * 13.0 + ((d->box_centers[(int64_t)a * d->n_boxes + tb] -
                               d->box_centers[(int64_t)a * d->n_boxes + sb]) /
                                  sext * 4.0 +
                              6.0 + d->case_encoding_bias);
      const int64_t vid = (int64_t)(code + 0.5);
      const int cid = (vid >= 0 && vid < (int64_t)pay->case_indices.size()) ? (int)pay->case_indices[vid] : -1;
      if (cid < 0) {
        delete pl;
        return fail(VL1_E_COVERAGE, "list-1 case encoding produced an invalid case id (target box " +
                                   std::to_string(tb) + ", source box " + std::to_string(sb) + ")");
      }
      int lev = 0;
      double alpha = 1.0, disp = 0.0;
      if (d->scaling_mode == VL1_SCALING_CANONICAL) {
        double num = 1.0, den = 1.0;
        for (int p = 0; p < d->scaling_power; ++p) {
          num *= sext;
          den *= H;
        }
        alpha = num / den;
        if (pl->log_disp) disp = -0.5 / M_PI * alpha * log(sext / H);
      } else if (d->scaling_mode == VL1_SCALING_CUSTOM) {
        alpha = d->level_scaling[std::min(lvl, VL1_MAX_LEVELS - 1)];
        disp = d->level_disp[std::min(lvl, VL1_MAX_LEVELS - 1)];
      } else if (d->scaling_mode == VL1_SCALING_PER_LEVEL) {
        const int rel = lvl - d->table_starting_level;
        lev = lvl < d->table_starting_level ? 0 : (rel < pay->n_tables ? rel : pay->n_tables - 1);
      } else if (lvl != d->table_starting_level) {
        delete pl;
        return fail(VL1_E_LEVEL, "Single near-field table level mismatch: source level " +
                                     std::to_string(lvl) + " but table_starting_level is " +
                                     std::to_string(d->table_starting_level));
      }
      // split blocks depend on the source level itself (R is not scale free)
      auto key = std::make_pair(cid, split ? lvl : lev);
      auto f = blk_of.find(key);
      int blk;
      if (f == blk_of.end()) {
        blk = (int)pl->blk_case.size();
        blk_of[key] = blk;
        pl->blk_case.push_back(cid);
        pl->blk_lev.push_back(lev);
        if (split) {
          pl->blk_alpha.push_back(alpha);
          pl->blk_disp.push_back(want_log ? -0.5 / M_PI * alpha * log(sext / H) : 0.0);
          pl->blk_h.push_back(sext);
        }
      } else {
        blk = f->second;
      }
      pl->pair_tgt.push_back(i);
      pl->pair_src.push_back(sb);
      pl->pair_block.push_back(blk);
      pl->pair_alpha.push_back(split ? 1.0 : alpha);
      pl->pair_disp.push_back(split ? 0.0 : disp);
    }
  }
  pl->n_keys = (int)pl->blk_case.size();
  pl->n_blocks = pl->n_keys * pl->n_rg;
  // per source box: the pair scaling and the 2D log-term factor depend on
  // the source box alone (its level); the int8 form folds alpha into the
  // box's density digits
  std::vector<double> balpha(d->n_boxes, 0.0), bdisp(d->n_boxes, 0.0);
  for (int64_t k = 0; k < n_pairs; ++k) {
    // a zero scaling marks a box that is no source in the int8 staging
    if (!(std::isfinite(pl->pair_alpha[k]) && pl->pair_alpha[k] != 0.0)) {
      const double a = pl->pair_alpha[k];
      delete pl;
      return fail(VL1_E_ARG, "vl1_plan_create: source box scaling " + std::to_string(a) +
                                 " must be finite and nonzero (root extents)");
    }
    balpha[pl->pair_src[k]] = pl->pair_alpha[k];
    bdisp[pl->pair_src[k]] = pl->pair_disp[k];
  }

  // per point the class of its source box's scaling, for the streaming
  // column maxima: distinct scalings (bitwise) get classes 0..254; more of
  // them, points shared by two source boxes or the 2D log term (whose dots
  // the per-box pass computes) keep the per-box pass
  std::vector<uint8_t> pcls;
  std::vector<double> lut(256, 0.0);
  bool stream_ok = !pl->log_disp;
  if (stream_ok) {
    pcls.assign((size_t)d->n_points, 255);
    int n_lut = 0;
    for (int b = 0; b < d->n_boxes && stream_ok; ++b) {
      if (balpha[b] == 0.0) continue;  // no list-1 source
      int idx = -1;
      for (int k = 0; k < n_lut; ++k)
        if (std::memcmp(&lut[k], &balpha[b], sizeof(double)) == 0) idx = k;
      if (idx < 0) {
        if (n_lut == 255) {
          stream_ok = false;
          break;
        }
        idx = n_lut;
        lut[n_lut++] = balpha[b];
      }
      for (int k = 0; k < n; ++k) {
        const int64_t pt = (int64_t)d->box_starts[b] + k;  // in range: checked above
        if (pcls[pt] != 255) {
          stream_ok = false;
          break;
        }
        pcls[pt] = (uint8_t)idx;
      }
    }
  }

  hipError_t e = upload(pl->mode_point, mode_point.data(), mode_point.size() * 4, s);
  if (e == hipSuccess && stream_ok) {
    e = upload(pl->pt_class, pcls.data(), pcls.size(), s);
    if (e == hipSuccess) e = upload(pl->alpha_lut, lut.data(), lut.size() * 8, s);
    pl->stream_colmax = e == hipSuccess;
  }
  if (e == hipSuccess) {
    const int32_t zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    e = upload(pl->phase, zeros, sizeof(zeros), s);
  }
  std::vector<int32_t> bstarts(d->box_starts, d->box_starts + d->n_boxes);
  if (e == hipSuccess) e = upload(pl->box_starts, bstarts.data(), bstarts.size() * 4, s);
  if (e == hipSuccess) e = upload(pl->box_alpha, balpha.data(), balpha.size() * 8, s);
  if (e == hipSuccess) e = upload(pl->box_disp, bdisp.data(), bdisp.size() * 8, s);
  DevBuf dcase, dlev, dperm, dsign, dgroup;
  if (e == hipSuccess && pl->n_blocks > 0) {
    e = upload(dcase, pl->blk_case.data(), pl->blk_case.size() * 4, s);
    if (e == hipSuccess) e = upload(dlev, pl->blk_lev.data(), pl->blk_lev.size() * 4, s);
    if (e == hipSuccess) e = upload(dperm, pay->perm.data(), pay->perm.size(), s);
    if (e == hipSuccess) e = upload(dsign, pay->sign.data(), pay->sign.size(), s);
    if (e == hipSuccess) e = upload(dgroup, pay->group.data(), pay->group.size(), s);
    if (e == hipSuccess) e = pl->M.alloc((size_t)pl->n_blocks * pl->MP * pl->KP * sizeof(double));
    if (e == hipSuccess && split) {
      std::vector<int32_t> vecs((size_t)pay->n_cases * dim);
      for (size_t k = 0; k < vecs.size(); ++k) vecs[k] = sp->case_vecs[k];
      const bool deriv = sp->deriv_kind != VL1_SPLIT_DERIV_NONE;
      const std::complex<double> kk(sp->k_re, sp->k_im);
      const int order = sp->split_order < 1 ? 1 : sp->split_order;
      std::vector<double> coef = deriv ? split_derivative_coefficients(dim, kk, sp->series_nmax, order)
                                       : split_coefficients(dim, kk, sp->series_nmax, order);
      if (coef.empty()) coef.assign(2, 0.0);  // 3D derivative with one term: R' = 0
      DevBuf dal, ddi, dh, dvec, dun, duw, dco, dext;
      e = upload(dal, pl->blk_alpha.data(), pl->blk_alpha.size() * 8, s);
      if (e == hipSuccess) e = upload(ddi, pl->blk_disp.data(), pl->blk_disp.size() * 8, s);
      if (e == hipSuccess) e = upload(dh, pl->blk_h.data(), pl->blk_h.size() * 8, s);
      if (e == hipSuccess) e = upload(dvec, vecs.data(), vecs.size() * 4, s);
      if (e == hipSuccess) e = upload(dun, sp->unit_nodes, (size_t)pay->q * 8, s);
      if (e == hipSuccess) e = upload(duw, sp->unit_weights, (size_t)pay->q * 8, s);
      if (e == hipSuccess) e = upload(dco, coef.data(), coef.size() * 8, s);
      if (e == hipSuccess && (order > 1 || sp->smooth_order > 0)) {
        // per (case, level) block -- n_keys of them, before the cut into
        // row groups -- and (target, source) the remainder of order m, the
        // terms' tables and the 2D log parts, complex
        DevBuf dtt, dtp, dtl, dtc, dsn, dsw, dsi;
        const int nt = sp->n_terms, ms = sp->smooth_order;
        e = dext.alloc((size_t)pl->n_keys * n * n * sizeof(double2));
        if (e == hipSuccess && nt > 0) {
          e = upload(dtt, sp->term_tables, (size_t)nt * pay->n_entries * 8, s);
          if (e == hipSuccess) e = upload(dtp, sp->term_powers, (size_t)nt * 4, s);
          if (e == hipSuccess) e = upload(dtl, sp->term_logs, (size_t)nt * 4, s);
          if (e == hipSuccess) e = upload(dtc, sp->term_coefs, (size_t)nt * 16, s);
        }
        if (e == hipSuccess && ms > 0) {
          e = upload(dsn, sp->smooth_nodes, (size_t)ms * 8, s);
          if (e == hipSuccess) e = upload(dsw, sp->smooth_weights, (size_t)ms * 8, s);
          if (e == hipSuccess) e = upload(dsi, sp->smooth_interp, (size_t)ms * pay->q * 8, s);
        }
        if (e == hipSuccess) {
          const int64_t total = (int64_t)pl->n_keys * n * n;
          const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
          auto kern = dim == 2 ? split_extra_kernel<2> : split_extra_kernel<3>;
          hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, pay->q, n, pl->n_keys,
                             dcase.as<int32_t>(), dh.as<double>(), H, dvec.as<int32_t>(),
                             dperm.as<uint8_t>(), dsign.as<int8_t>(), dgroup.as<uint8_t>(),
                             pay->ranks.as<uint16_t>(), pay->offs.as<int32_t>(), pay->n_entries,
                             dun.as<double>(), duw.as<double>(), dco.as<double>(),
                             sp->series_nmax, ms, dsn.as<double>(), dsw.as<double>(),
                             dsi.as<double>(), nt, dtt.as<double>(), dtp.as<int32_t>(),
                             dtl.as<int32_t>(), dtc.as<double>(), deriv ? sp->deriv_axis : -1,
                             sp->deriv_kind == VL1_SPLIT_DERIV_SOURCE ? -1.0 : 1.0,
                             pay->sign_mask, dext.as<double2>());
          e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);  // before the uploads are freed
      }
      if (e == hipSuccess) {
        const int64_t total = (int64_t)pl->n_blocks * pl->MP * pl->KP;
        const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(expand_split_blocks_kernel, dim3(grid), dim3(256), 0, s, dim, pay->q, n,
                           pay->sign_mask, deriv ? sp->deriv_axis : -1,
                           sp->deriv_kind == VL1_SPLIT_DERIV_SOURCE ? -1.0 : 1.0, pl->MP, pl->KP,
                           pl->n_blocks, pl->n_rg, pl->rg_size, dcase.as<int32_t>(),
                           dlev.as<int32_t>(), pay->n_entries, dal.as<double>(),
                           ddi.as<double>(), dh.as<double>(), dvec.as<int32_t>(),
                           dperm.as<uint8_t>(), dsign.as<int8_t>(), dgroup.as<uint8_t>(),
                           pay->ranks.as<uint16_t>(), pay->offs.as<int32_t>(),
                           pay->table.as<double>(), pay->nmlz.as<double>(), dun.as<double>(),
                           duw.as<double>(), dco.as<double>(), sp->series_nmax,
                           dext.ptr ? dext.as<double2>() : nullptr, pl->M.as<double>());
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(s);  // before the uploads are freed
    } else if (e == hipSuccess && pay->dense_kind) {
      const int64_t total = (int64_t)pl->n_blocks * pl->MP * pl->KP;
      const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
      hipLaunchKernelGGL(expand_dense_blocks_kernel, dim3(grid), dim3(256), 0, s, n, pay->n_cases,
                         pl->cplx, pl->MP, pl->KP, pl->n_blocks, pl->n_rg, pl->rg_size,
                         dcase.as<int32_t>(), dlev.as<int32_t>(), pay->dense.as<double>(),
                         pl->M.as<double>());
      e = hipGetLastError();
    } else if (e == hipSuccess) {
      const int64_t total = (int64_t)pl->n_blocks * pl->MP * pl->KP;
      const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
      hipLaunchKernelGGL(expand_blocks_kernel, dim3(grid), dim3(256), 0, s, dim, pay->q, n, pl->cplx,
                         pay->sign_mask, pl->MP, pl->KP, pl->n_blocks, pl->n_rg, pl->rg_size,
                         dcase.as<int32_t>(), dlev.as<int32_t>(),
                         dperm.as<uint8_t>(), dsign.as<int8_t>(), dgroup.as<uint8_t>(),
                         pay->ranks.as<uint16_t>(), pay->offs.as<int32_t>(), pay->table.as<double>(),
                         pay->n_entries, pl->M.as<double>());
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = pl->MF.alloc(pl->M.bytes);
    if (e == hipSuccess) {
      const int64_t total = (int64_t)pl->n_blocks * pl->MP * pl->KP;
      const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
      hipLaunchKernelGGL(frag_f64_kernel, dim3(grid), dim3(256), 0, s, pl->MP, pl->KP, pl->n_blocks,
                         pl->M.as<double>(), pl->MF.as<double>());
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    delete pl;
    return fail(VL1_E_HIP, std::string("vl1_plan_create: ") + hipGetErrorString(e));
  }
  *out = pl;
  return VL1_OK;
}

extern "C" {

int vl1_plan_create(vl1_plan** out, const vl1_payload* pay, const vl1_plan_desc* d, void* stream) {
  return plan_create(out, pay, d, nullptr, stream);
}

int vl1_plan_create_helmholtz_split(vl1_plan** out, const vl1_payload* laplace,
                                    const vl1_plan_desc* d, const vl1_helmholtz_split_blocks* sp,
                                    void* stream) {
  if (!sp) {
    clear_error();
    return fail(VL1_E_ARG, "vl1_plan_create_helmholtz_split: null split description");
  }
  return plan_create(out, laplace, d, sp, stream);
}

int vl1_plan_destroy(vl1_plan* plan) {
  delete plan;
  return VL1_OK;
}

// Tile-kernel variants (bit set): kVarOneRow = one 16-row tile per wave
// (split form, blocks of at most 128 rows); kVarMorton = tiles in plain Morton
// order instead of grouped by block signature (only matters with several
// boxes per tile); kVarNoPhase = no XCD phase alignment of the step order.
constexpr int kVarOneRow = 1, kVarMorton = 2, kVarNoPhase = 4;

static int tune_key(int form, int S, int Rc) { return (form * 8 + S) * 1024 + Rc; }

static int default_variant(const vl1_plan* pl, int form) {
  (void)form;
  return pl->MP <= 128 ? kVarOneRow : 0;
}

static int plan_variant(const vl1_plan* pl, int form, int S, int Rc) {
  if (pl->forced_variant >= 0) return pl->forced_variant;
  auto it = pl->tuned.find(tune_key(form, S, Rc));
  return it != pl->tuned.end() ? it->second : default_variant(pl, form);
}

// Argument checks shared by every evaluation entry point; *S_out receives
// the slice count (0 for the FP64 form).
static int plan_check(const vl1_plan* pl, int form, int n_rhs, const vl1_split_opts* opts,
                      int* S_out) {
  if (form != VL1_FORM_F64 && form != VL1_FORM_SPLIT_BF16 && form != VL1_FORM_SPLIT_I8)
    return fail(VL1_E_ARG, "vl1_plan: unknown form");
  if (n_rhs < 1) return fail(VL1_E_ARG, "vl1_plan: n_rhs must be >= 1");
  if (n_rhs > kCols && n_rhs % kCols != 0)
    return fail(VL1_E_UNSUPPORTED, "vl1_plan: RHS count must be <= 64 or a multiple of 64");
  // one staging column per (box, RHS); the kernels index columns with int
  if ((int64_t)pl->nboxes * n_rhs > (int64_t)INT32_MAX)
    return fail(VL1_E_UNSUPPORTED, "vl1_plan: boxes x RHS exceeds 2^31 - 1 columns");
  int S = 0;
  if (form == VL1_FORM_SPLIT_BF16) {
    S = opts ? opts->n_slices : 5;
    if (S < 1 || S > kMaxSlices) return fail(VL1_E_ARG, "vl1_plan: n_slices must be 1..6");
    // fused split blocks hold the near-singular table part and the smooth
    // remainder in one row; beyond 256 columns (3D q >= 6, complex) the
    // split form's error, relative to each row's largest entry, exceeds
    // 1e-10 of the result at the targets: the FP64 form evaluates these
    if (pl->fused_split && pl->KP > 256)
      return fail(VL1_E_UNSUPPORTED,
                  "vl1_plan: fused Helmholtz split blocks beyond 256 columns take the FP64 form");
    if (split_lds_bytes(S, pl->MP) + kStaticLds > 160 * 1024)
      return fail(VL1_E_UNSUPPORTED, "vl1_plan: too many slices for the box size (LDS)");
  } else if (form == VL1_FORM_SPLIT_I8) {
    S = opts ? opts->n_slices : 6;
    if (S < 1 || S > kI8MaxSlices) return fail(VL1_E_ARG, "vl1_plan: n_slices must be 1..6");
    const int rows = pl->cplx ? 2 * pl->n : pl->n;
    if (rows / 2 + 2 + rows > kI8ColsLds)  // the staging's offsets and one column in LDS
      return fail(VL1_E_UNSUPPORTED, "vl1_plan: int8 split form: more than " +
                                         std::to_string(2 * (kI8ColsLds - 2) / 3) +
                                         " real rows per box");
    if (pl->fused_split && pl->KP > 256)
      return fail(VL1_E_UNSUPPORTED,
                  "vl1_plan: fused Helmholtz split blocks beyond 256 columns take the FP64 form");
  }
  *S_out = S;
  return VL1_OK;
}

// Ring slots of the int8 tile kernel (at least 3: the consumers read the
// next chunk's block id one barrier after the producer wrote it).  A
// workgroup of at most 4 waves (blocks of at most 96 rows) gets a ring small
// enough for two workgroups per CU, so each SIMD runs two waves.
static int i8_ring_slots(int S, int waves) {
  const size_t budget = waves <= 4 ? 80 * 1024 : 160 * 1024;
  int ns = kI8Ring;
  while (ns > 3 && i8_lds_bytes(S, ns) > budget) --ns;
  return ns;
}

// The int8 form sums a tile's digit products exactly in int32: the largest
// diagonal (S products per K column, |d| <= 127) over all K columns of all
// the tile's steps must stay below 2^31.
static int i8_check_range(const vl1_plan* pl, const TileSchedule* sch, int S) {
  const double worst = (double)sch->max_steps * pl->KP * S * 127.0 * 127.0;
  if (worst >= 2147483647.0)
    return fail(VL1_E_UNSUPPORTED, "vl1_plan: int8 split form: " + std::to_string(sch->max_steps) +
                                       " steps per tile could overflow the exact int32 sums; "
                                       "use the BF16 split or FP64 form");
  return VL1_OK;
}

// Checks, then builds (allocating) the slice set and tile schedule this
// evaluation needs if the plan does not hold them yet.
static int plan_prepare(vl1_plan* pl, int form, int n_rhs, const vl1_split_opts* opts,
                        hipStream_t s, int* S_out, TileSchedule** sch_out) {
  int S = 0;
  int rc = plan_check(pl, form, n_rhs, opts, &S);
  if (rc != VL1_OK) return rc;
  if (form == VL1_FORM_SPLIT_BF16) {
    rc = ensure_A(pl, S, s);
    if (rc != VL1_OK) return rc;
  } else if (form == VL1_FORM_SPLIT_I8) {
    rc = ensure_A8(pl, S, s);
    if (rc != VL1_OK) return rc;
  }
  *S_out = S;
  const int Rc = std::min(n_rhs, kCols);
  rc = build_schedule(pl, Rc, (plan_variant(pl, form, S, Rc) & kVarMorton) != 0, sch_out, s);
  if (rc == VL1_OK && form == VL1_FORM_SPLIT_I8) rc = i8_check_range(pl, *sch_out, S);
  return rc;
}

// Per-evaluation staging: density slices, column scales and the 2D log-term
// dots (split form), or the gathered FP64 densities and dots (FP64 form).
struct Staging {
  __bf16* X = nullptr;
  double* colscale = nullptr;
  double* nmdot = nullptr;
  double* XF = nullptr;
  int8_t* X8 = nullptr;      // int8 form: density digits
  double* sigma = nullptr;   // int8 form: column scale per RHS
  double* est_part = nullptr;  // int8 form: error-estimate partials per (tile, column)
};

// Bytes of staging for (form, S, R) laid out in one region, each part
// 256-byte aligned: split X | colscale | nmdot, FP64 XF | nmdot.
static size_t staging_layout(const vl1_plan* pl, int form, int S, int R, int n_tiles, char* base,
                             Staging* st) {
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t cols = (size_t)pl->nboxes * R;
  size_t off = 0;
  Staging t;
  if (form == VL1_FORM_SPLIT_I8) {
    t.X8 = reinterpret_cast<int8_t*>(base + off);
    off += al((size_t)pl->nboxes * S * R * pl->KP);
    t.colscale = reinterpret_cast<double*>(base + off);  // per-column maxima
    off += al(cols * 8);
    t.sigma = reinterpret_cast<double*>(base + off);
    off += al((size_t)R * 8);
    t.est_part = reinterpret_cast<double*>(base + off);
    off += al((size_t)n_tiles * 64 * 4 * 8);
  } else if (form == VL1_FORM_SPLIT_BF16) {
    t.X = reinterpret_cast<__bf16*>(base + off);
    off += al((size_t)pl->nboxes * S * R * pl->KP * 2);
    t.colscale = reinterpret_cast<double*>(base + off);
    off += al(cols * 8);
  } else {
    t.XF = reinterpret_cast<double*>(base + off);
    off += al((size_t)pl->nboxes * R * pl->KP * 8);
  }
  t.nmdot = reinterpret_cast<double*>(base + off);
  off += al(cols * 8);
  if (st) *st = t;
  return off;
}

// Launch the staging kernel of `form` into `st`.
static int stage_launch(vl1_plan* pl, int form, int S, const double* f, int64_t ldf, int R,
                        const Staging& st, hipStream_t s) {
  const int ncolumns = pl->nboxes * R;
  const int wpb = 4;
  const double* nmlz = pl->payload->nmlz.as<double>();  // level-0 normalisers
  if (form == VL1_FORM_SPLIT_I8) {
    // column scale per RHS: the largest |f alpha| over all boxes (per-column
    // maxima, then one reduction per RHS)
    const int rows = pl->cplx ? 2 * pl->n : pl->n;
    int n_partial = pl->nboxes;  // maxima to fold: per box, or per value chunk
    const int64_t n_values = pl->n_points * (pl->cplx ? 2 : 1);
    const int64_t n_chunks = (n_values + kI8StreamChunk - 1) / kI8StreamChunk;
    if (pl->stream_colmax && n_chunks <= pl->nboxes) {
      // flat streams of the density rows (partials per 2048-value chunk fit
      // the per-box maxima's space)
      n_partial = (int)n_chunks;
      hipLaunchKernelGGL(i8_colmax_stream_kernel,
                         dim3((unsigned)n_chunks, (unsigned)((R + kI8StreamRows - 1) / kI8StreamRows)),
                         dim3(256), 0, s, n_values, pl->cplx ? 1 : 0, R, f, ldf,
                         pl->pt_class.as<uint8_t>(), pl->alpha_lut.as<double>(), st.colscale);
    } else if (rows <= 128) {
      // small boxes: a lane per (box, RHS) (2D Helmholtz forest: 2.37 -> 1.3 ms)
      const int64_t cm_lanes = ((int64_t)pl->nboxes + 63) / 64 * 64 * R;
      hipLaunchKernelGGL(i8_colmax_lanes_kernel, dim3((unsigned)((cm_lanes + 255) / 256)),
                         dim3(256), 0, s, pl->n, pl->cplx, pl->nboxes, R, f, ldf,
                         pl->box_starts.as<int32_t>(), pl->box_alpha.as<double>(), nmlz,
                         pl->log_disp, st.colscale, st.nmdot);
    } else {
      // large boxes: four boxes per wavefront, whole-wave rows
      const int64_t cm_waves = (int64_t)((pl->nboxes + 3) / 4) * R;
      hipLaunchKernelGGL(i8_colmax_kernel<4>, dim3((unsigned)((cm_waves + wpb - 1) / wpb)),
                         dim3(64 * wpb), 0, s, pl->n, pl->cplx, pl->nboxes, R, f, ldf,
                         pl->box_starts.as<int32_t>(), pl->box_alpha.as<double>(), nmlz,
                         pl->log_disp, st.colscale, st.nmdot);
    }
    // maxima over the boxes in two stages (G partial maxima per RHS, then one)
    const int G = std::max(1, std::min(256, n_partial / 1024));
    if (G > 1)
      hipLaunchKernelGGL(i8_colmax_fold_kernel, dim3(G, R), dim3(256), 0, s, n_partial, R,
                         st.colscale);
    hipLaunchKernelGGL(i8_sigma_kernel, dim3(R), dim3(256), 0, s, G > 1 ? G : n_partial, R,
                       st.colscale, st.sigma);
    const int RR = i8_cols_round(rows, R);
    const int tasks = RR * (pl->KP / 32) * 2;  // 16 K values of one column each
    const int threads = tasks > 128 ? 256 : (tasks > 64 ? 128 : 64);
    const size_t lds = sizeof(double) * ((size_t)(rows + 1) / 2 + (size_t)RR * (rows + 1));
    const dim3 g(pl->nboxes), b(threads);
    switch (S) {
#define VL1_I8_COLS(SV)                                                                          \
  case SV:                                                                                       \
    hipLaunchKernelGGL(i8_split_cols_kernel<SV>, g, b, lds, s, pl->n, pl->cplx, pl->KP,         \
                       pl->nboxes, R, RR, f, ldf, pl->mode_point.as<int32_t>(),                   \
                       pl->box_starts.as<int32_t>(), pl->box_alpha.as<double>(), st.sigma,        \
                       st.X8);                                                                    \
    break;
      VL1_I8_COLS(1) VL1_I8_COLS(2) VL1_I8_COLS(3) VL1_I8_COLS(4) VL1_I8_COLS(5) VL1_I8_COLS(6)
#undef VL1_I8_COLS
    }
  } else if (form == VL1_FORM_SPLIT_BF16) {
    // two columns per wavefront, kColBlock / 32 columns per block
    hipLaunchKernelGGL(split_cols_kernel, dim3((ncolumns + kColBlock / 32 - 1) / (kColBlock / 32)),
                       dim3(kColBlock), 0, s, pl->n, pl->cplx, pl->KP, pl->nboxes, R, S, f, ldf,
                       pl->mode_point.as<int32_t>(), pl->box_starts.as<int32_t>(), nmlz,
                       pl->log_disp, st.X, st.colscale, st.nmdot);
  } else {
    hipLaunchKernelGGL(gather_cols_f64_kernel, dim3((ncolumns + wpb - 1) / wpb), dim3(64 * wpb), 0, s,
                       pl->n, pl->cplx, pl->KP, pl->nboxes, R, f, ldf, pl->mode_point.as<int32_t>(),
                       pl->box_starts.as<int32_t>(), nmlz, pl->log_disp, st.XF, st.nmdot);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("staging launch: ") + hipGetErrorString(e));
  return VL1_OK;
}

// Launch the tile kernels of `form` reading the staging `st`.
// error_est: optional device [R][2] (rel Linf, rel L2 estimates per RHS).
static int contract_launch(vl1_plan* pl, int form, int S, TileSchedule* sch, int R, double* y,
                           int64_t ldy, const Staging& st, hipStream_t s,
                           double* error_est = nullptr) {
  const int Rc = std::min(R, kCols);
  hipError_t e;
  TileParams prm{};
  prm.n = pl->rows;
  prm.n_rg = pl->n_rg;
  prm.rg_size = pl->rg_size;
  prm.nm = pl->n;
  prm.cplx = pl->cplx;
  prm.MP = pl->MP;
  prm.KP = pl->KP;
  prm.S = S;
  prm.B = sch->B;
  prm.Rc = Rc;
  prm.R = R;
  prm.n_tiles = sch->n_tiles;
  prm.tile_step_start = sch->tile_step_start.as<int32_t>();
  prm.tile_boxes = sch->tile_boxes.as<int32_t>();
  prm.step_block = sch->step_block.as<int32_t>();
  prm.step_src = sch->step_src.as<int32_t>();
  prm.step_alpha = sch->step_alpha.as<double>();
  prm.step_disp = sch->step_disp.as<double>();
  prm.mode_point = pl->mode_point.as<int32_t>();
  prm.A = form == VL1_FORM_SPLIT_BF16 ? pl->A(S)->as<__bf16>() : nullptr;
  prm.A8 = form == VL1_FORM_SPLIT_I8 ? pl->A8(S)->as<int8_t>() : nullptr;
  prm.X8 = st.X8;
  prm.rho = pl->rho.as<double>();
  prm.sigma = st.sigma;
  prm.box_disp = pl->box_disp.as<double>();
  prm.est_part = error_est ? st.est_part : nullptr;
  prm.nslot = i8_ring_slots(S, pl->MP / 32 + 1);
  prm.log_disp = pl->log_disp;
  prm.rowscale = pl->rowscale.as<double>();
  prm.X = st.X;
  prm.colscale = st.colscale;
  prm.M = pl->M.as<double>();
  prm.MF = pl->MF.as<double>();
  prm.XF = st.XF;
  prm.nmdot = st.nmdot;
  prm.y = y;
  prm.ldy = ldy;
  prm.phase = pl->phase.as<int32_t>();
  const int variant = plan_variant(pl, form, S, Rc);
  prm.use_phase = (variant & kVarNoPhase) ? 0 : 1;
  const int grid = ((sch->n_tiles + 7) / 8) * 8;
  for (int r0 = 0; r0 < R; r0 += Rc) {
    prm.r0 = r0;
    if (form == VL1_FORM_SPLIT_I8) {
      const int waves = pl->MP / 32 + 1;  // a consumer per 32-row tile + the producer
      e = launch_i8(S, dim3(grid), dim3(64 * waves), i8_lds_bytes(S, prm.nslot), s, prm);
      if (e != hipSuccess) return fail(VL1_E_HIP, std::string("int8 split launch: ") + hipGetErrorString(e));
      if (error_est) {
        // two stages over the tiles (G partials per RHS, then one)
        const int G = std::max(1, std::min(256, sch->n_tiles / 256));
        if (G > 1)
          hipLaunchKernelGGL(i8_est_fold_kernel, dim3(G, Rc), dim3(256), 0, s, sch->n_tiles, sch->B,
                             Rc, st.est_part);
        hipLaunchKernelGGL(i8_rhs_error_kernel, dim3(Rc), dim3(256), 0, s, G > 1 ? G : sch->n_tiles,
                           G > 1 ? 1 : sch->B, Rc, r0, st.est_part, error_est);
        e = hipGetLastError();
        if (e != hipSuccess) return fail(VL1_E_HIP, std::string("error estimate launch: ") + hipGetErrorString(e));
      }
      continue;
    }
    if (form == VL1_FORM_SPLIT_BF16) {
      const size_t lds = split_lds_bytes(S, pl->MP);
      // one row tile per wave (blocks of at most 8 row tiles only)
      const bool one = (variant & kVarOneRow) && pl->MP <= 128;
      e = launch_split(S, one, dim3(grid), lds, s, prm);
      if (e != hipSuccess) return fail(VL1_E_HIP, std::string("split launch: ") + hipGetErrorString(e));
    } else {
      const size_t lds = (size_t)2 * kCols * kDPitch * 8 + 6 * kCols * 8;
      static std::atomic<unsigned> dense_attr_done[2];
      const bool one = (variant & kVarOneRow) && pl->MP <= 128;
      if (one) {
        e = allow_full_lds((const void*)dense_f64_kernel<1>, dense_attr_done[1]);
        if (e == hipSuccess) hipLaunchKernelGGL(dense_f64_kernel<1>, dim3(grid), dim3(512), lds, s, prm);
      } else {
        e = allow_full_lds((const void*)dense_f64_kernel<2>, dense_attr_done[0]);
        if (e == hipSuccess) hipLaunchKernelGGL(dense_f64_kernel<2>, dim3(grid), dim3(512), lds, s, prm);
      }
      if (e != hipSuccess) return fail(VL1_E_HIP, std::string("dense attr: ") + hipGetErrorString(e));
    }
    e = hipGetLastError();
    if (e != hipSuccess) return fail(VL1_E_HIP, std::string("tile kernel launch: ") + hipGetErrorString(e));
  }
  return VL1_OK;
}

// The plan's own staging buffers.
static Staging plan_staging(vl1_plan* pl) {
  Staging st;
  st.X = pl->X.as<__bf16>();
  st.colscale = pl->colscale.as<double>();
  st.nmdot = pl->nmdot.as<double>();
  st.XF = pl->XF.as<double>();
  st.X8 = pl->X8.as<int8_t>();
  st.sigma = pl->sigma.as<double>();
  st.est_part = pl->est_part.as<double>();
  return st;
}

int vl1_plan_stage(vl1_plan* pl, int form, const double* f, int64_t ldf, int n_rhs,
                   const vl1_split_opts* opts, void* stream) {
  clear_error();
  if (!pl || !f) return fail(VL1_E_ARG, "vl1_plan_stage: null argument");
  if (ldf < pl->n_points * (pl->cplx ? 2 : 1))
    return fail(VL1_E_ARG, "vl1_plan_stage: bad leading dimension");
  if (pl->n_tgt == 0 || pl->n_blocks == 0) return VL1_OK;
  hipStream_t s = (hipStream_t)stream;
  int S = 0;
  TileSchedule* sch = nullptr;
  int rc = plan_prepare(pl, form, n_rhs, opts, s, &S, &sch);
  if (rc != VL1_OK) return rc;
  const int R = n_rhs;
  const size_t ncolumns = (size_t)pl->nboxes * R;
  hipError_t e = hipSuccess;
  // the largest buffer first: a failed allocation then leaves the others
  // untouched
  if (form == VL1_FORM_SPLIT_I8) {
    e = pl->grow(pl->X8, (size_t)pl->nboxes * S * R * pl->KP);
    if (e == hipSuccess) e = pl->grow(pl->nmdot, ncolumns * 8);
    if (e == hipSuccess) e = pl->grow(pl->colscale, ncolumns * 8);
    if (e == hipSuccess) e = pl->grow(pl->sigma, (size_t)R * 8);
    if (e == hipSuccess) e = pl->grow(pl->est_part, (size_t)sch->n_tiles * 64 * 4 * 8);
  } else if (form == VL1_FORM_SPLIT_BF16) {
    e = pl->grow(pl->X, (size_t)pl->nboxes * S * R * pl->KP * 2);
    if (e == hipSuccess) e = pl->grow(pl->nmdot, ncolumns * 8);
    if (e == hipSuccess) e = pl->grow(pl->colscale, ncolumns * 8);
  } else {
    e = pl->grow(pl->XF, (size_t)pl->nboxes * R * pl->KP * 8);
    if (e == hipSuccess) e = pl->grow(pl->nmdot, ncolumns * 8);
  }
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("vl1_plan_stage alloc: ") + hipGetErrorString(e));
  const Staging st = plan_staging(pl);
  rc = stage_launch(pl, form, S, f, ldf, R, st, s);
  if (rc != VL1_OK) return rc;
  pl->staged_form = form;
  pl->staged_rhs = R;
  pl->staged_slices = S;
  return VL1_OK;
}

int vl1_plan_contract(vl1_plan* pl, int form, int n_rhs, double* y, int64_t ldy,
                      const vl1_split_opts* opts, void* stream) {
  clear_error();
  if (!pl || !y) return fail(VL1_E_ARG, "vl1_plan_contract: null argument");
  if (ldy < pl->n_points * (pl->cplx ? 2 : 1))
    return fail(VL1_E_ARG, "vl1_plan_contract: bad leading dimension");
  if (pl->n_tgt == 0 || pl->n_blocks == 0) return VL1_OK;
  hipStream_t s = (hipStream_t)stream;
  int S = 0;
  TileSchedule* sch = nullptr;
  int rc = plan_prepare(pl, form, n_rhs, opts, s, &S, &sch);
  if (rc != VL1_OK) return rc;
  if (pl->staged_form != form || pl->staged_rhs != n_rhs || pl->staged_slices != S)
    return fail(VL1_E_ARG, "vl1_plan_contract: densities were not staged for this form/RHS/slices");
  return contract_launch(pl, form, S, sch, n_rhs, y, ldy, plan_staging(pl), s,
                         form == VL1_FORM_SPLIT_I8 && opts ? opts->error_est : nullptr);
}

// ------------------------------------------------ allocation-free evaluation
// Staging region of vl1_plan_eval_noalloc: a fixed device array, one copy
// per GPU, resident until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host it only names the symbol): the
// device address comes from hipGetSymbolAddress, looked up once per device
// with that device current.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void* scratch_ptr[kMaxDevices] = {};

static int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(VL1_E_UNSUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(VL1_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return VL1_OK;
}

int vl1_plan_eval_noalloc(vl1_plan* pl, int form, const double* f, int64_t ldf, int n_rhs,
                          double* y, int64_t ldy, const vl1_split_opts* opts, void* stream) {
  clear_error();
  if (!pl || !f || !y) return fail(VL1_E_ARG, "vl1_plan_eval_noalloc: null argument");
  const int64_t need_ld = pl->n_points * (pl->cplx ? 2 : 1);
  if (ldf < need_ld || ldy < need_ld)
    return fail(VL1_E_ARG, "vl1_plan_eval_noalloc: bad leading dimension");
  if (pl->n_tgt == 0 || pl->n_blocks == 0) return VL1_OK;
  int S = 0;
  int rc = plan_check(pl, form, n_rhs, opts, &S);
  if (rc != VL1_OK) return rc;
  // the slice set and tile schedule must exist already: building them
  // allocates (one vl1_plan_eval with this form, slice count and RHS count
  // per tile builds both)
  const int Rc = std::min(n_rhs, kCols);
  const int key = Rc * 2 + ((plan_variant(pl, form, S, Rc) & kVarMorton) ? 1 : 0);
  auto it = pl->schedules.find(key);
  if ((form != VL1_FORM_F64 && !pl->slices(form, S)) || it == pl->schedules.end())
    return fail(VL1_E_UNSUPPORTED,
                "vl1_plan_eval_noalloc: the plan is not prepared for this form / slice count / "
                "RHS count (evaluate it once with vl1_plan_eval first)");
  if (form == VL1_FORM_SPLIT_I8) {
    rc = i8_check_range(pl, it->second, S);
    if (rc != VL1_OK) return rc;
  }
  const size_t need = staging_layout(pl, form, S, n_rhs, it->second->n_tiles, nullptr, nullptr);
  if (need > kScratchBytes)
    return fail(VL1_E_UNSUPPORTED, "vl1_plan_eval_noalloc: staging needs " + std::to_string(need) +
                                       " bytes, above the fixed " + std::to_string(kScratchBytes) +
                                       "-byte scratch (use vl1_plan_eval, or fewer RHS per call)");
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc != VL1_OK) return rc;
  Staging st;
  staging_layout(pl, form, S, n_rhs, it->second->n_tiles, static_cast<char*>(base), &st);
  hipStream_t s = (hipStream_t)stream;
  // every word the contraction reads is written by the staging kernel
  // (padding rows included), so nothing needs clearing between calls
  rc = stage_launch(pl, form, S, f, ldf, n_rhs, st, s);
  if (rc != VL1_OK) return rc;
  return contract_launch(pl, form, S, it->second, n_rhs, y, ldy, st, s,
                         form == VL1_FORM_SPLIT_I8 && opts ? opts->error_est : nullptr);
}

int vl1_plan_eval(vl1_plan* pl, int form, const double* f, int64_t ldf, int n_rhs, double* y,
                  int64_t ldy, const vl1_split_opts* opts, void* stream) {
  int rc = vl1_plan_stage(pl, form, f, ldf, n_rhs, opts, stream);
  if (rc != VL1_OK) return rc;
  return vl1_plan_contract(pl, form, n_rhs, y, ldy, opts, stream);
}

// Autotuning: time every valid tile-kernel variant on the caller's data and
// keep the fastest for this (form, slices, RHS-per-tile) -- see the variant
// bits above.  Synchronous: the host waits for the timing runs.
int vl1_plan_tune(vl1_plan* pl, int form, const double* f, int64_t ldf, int n_rhs, double* y,
                  int64_t ldy, const vl1_split_opts* opts, int reps, int32_t* chosen,
                  double* times_ms, void* stream) {
  clear_error();
  if (!pl || !f || !y) return fail(VL1_E_ARG, "vl1_plan_tune: null argument");
  if (reps < 1) return fail(VL1_E_ARG, "vl1_plan_tune: reps must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  if (times_ms)
    for (int v = 0; v < 8; ++v) times_ms[v] = -1.0;
  int rc = vl1_plan_stage(pl, form, f, ldf, n_rhs, opts, stream);
  if (rc != VL1_OK) return rc;
  const int S = pl->staged_slices, Rc = std::min(n_rhs, kCols);
  if (pl->n_tgt == 0 || pl->n_blocks == 0) {
    if (chosen) *chosen = default_variant(pl, form);
    return VL1_OK;
  }
  const bool multi_box = kCols / Rc > 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int best = -1;
  float best_ms = 0.f;
  for (int v = 0; v < 8 && e == hipSuccess; ++v) {
    if ((v & kVarOneRow) && pl->MP > 128) continue;
    // the int8 kernel has one row tile per consumer wave and no phase hint:
    // only the tile order is a choice
    if (form == VL1_FORM_SPLIT_I8 && (v & (kVarOneRow | kVarNoPhase))) continue;
    if ((v & kVarMorton) && !multi_box) continue;
    pl->forced_variant = v;
    rc = vl1_plan_contract(pl, form, n_rhs, y, ldy, opts, stream);  // warm (and schedule)
    if (rc != VL1_OK) break;
    e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == VL1_OK && e == hipSuccess; ++r)
      rc = vl1_plan_contract(pl, form, n_rhs, y, ldy, opts, stream);
    if (rc != VL1_OK) break;
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    if (times_ms) times_ms[v] = ms;
    if (best < 0 || ms < best_ms) {
      best = v;
      best_ms = ms;
    }
  }
  pl->forced_variant = -1;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != VL1_OK) return rc;
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("vl1_plan_tune: ") + hipGetErrorString(e));
  pl->tuned[tune_key(form, S, Rc)] = best;
  if (chosen) *chosen = best;
  return VL1_OK;
}

int vl1_plan_set_variant(vl1_plan* pl, int form, int n_slices, int n_rhs, int variant) {
  clear_error();
  if (!pl) return fail(VL1_E_ARG, "vl1_plan_set_variant: null plan");
  if (form != VL1_FORM_F64 && form != VL1_FORM_SPLIT_BF16 && form != VL1_FORM_SPLIT_I8)
    return fail(VL1_E_ARG, "vl1_plan_set_variant: unknown form");
  if (n_rhs < 1) return fail(VL1_E_ARG, "vl1_plan_set_variant: n_rhs must be >= 1");
  const int S = form != VL1_FORM_F64 ? n_slices : 0;
  if (form != VL1_FORM_F64 && (S < 1 || S > kMaxSlices))
    return fail(VL1_E_ARG, "vl1_plan_set_variant: n_slices must be 1..6");
  if (variant < -1 || variant > 7) return fail(VL1_E_ARG, "vl1_plan_set_variant: variant must be -1..7");
  const int key = tune_key(form, S, std::min(n_rhs, kCols));
  if (variant < 0) {
    pl->tuned.erase(key);
    return VL1_OK;
  }
  if ((variant & kVarOneRow) && pl->MP > 128)
    return fail(VL1_E_UNSUPPORTED, "vl1_plan_set_variant: one row tile per wave needs blocks of at most 128 rows");
  pl->tuned[key] = variant;
  return VL1_OK;
}

int vl1_plan_stats(const vl1_plan* pl, int64_t* out5) {
  clear_error();
  if (!pl || !out5) return fail(VL1_E_ARG, "vl1_plan_stats: null argument");
  out5[0] = pl->n_pairs;
  int64_t tiles = 0, steps = 0;
  if (!pl->schedules.empty()) {
    const TileSchedule* sch = pl->schedules.begin()->second;
    tiles = sch->n_tiles;
    steps = sch->n_steps;
  }
  out5[1] = tiles;
  out5[2] = steps;
  out5[3] = pl->n_keys;
  int64_t bytes = (int64_t)(pl->M.bytes + pl->MF.bytes + pl->rowscale.bytes + pl->X.bytes +
                            pl->XF.bytes + pl->colscale.bytes + pl->nmdot.bytes +
                            pl->mode_point.bytes);
  for (auto& kv : pl->A_by_slices) bytes += (int64_t)kv.second->bytes;
  for (const DevBuf* b : pl->retired) bytes += (int64_t)b->bytes;
  out5[4] = bytes;
  return VL1_OK;
}

}  // extern "C"
