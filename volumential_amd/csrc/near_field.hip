// Plain FP64 list-1 evaluation with the reference's general CSR semantics.
//
// Semantics follow the loopy `near_field` kernel of
// volumential/list1.py:1057-1148: for every target box of `target_boxes`
// (CSR order) and every target particle `tid` of the box,
//   result[tbeg + tid] = sum_{sbox in list1} sum_{sid}
//       coef[sbeg + sid] * (T[lev, row(case, s, t)] * scaling + displacement)
// with case = case_indices[int(sum_d ((c_t - c_s)/h_s*4 + shift + bias) base^k + 0.5)]
// (list1.py:439-459, 1083-1091), s/t the box-local mode ids of the particles,
// and the scaling/level policy of list1.py:361-437.
//
// GPU mapping: one workgroup per target box, taken in CSR order.  For every
// source box the densities of up to RC right-hand sides are staged in LDS;
// each lane owns one target node and walks the source modes.  The table
// values come from the payload's expanded per-(level, case) blocks
// D[s][t] (built once, on the first plain evaluation, while they stay
// under 8 GiB): one coalesced line per source mode.  Beyond that size the
// canonical orbit block is staged in LDS when it fits and gathered through
// the per-case row map.  For boxes with fewer than 256 nodes the lanes are
// split into P parts over the source modes and reduced through LDS at the
// end ("wavefront reduction over list-1 sources").
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>

#include <type_traits>

#include "internal.h"

namespace vl1 {

struct NFParams {
  // custom scaling and displacement factor per source level [VL1_MAX_LEVELS]
  // (device); first in the struct: further down, the ROCm 7.2 compiler
  // fails on the VALU kernel's load of them ("illegal instruction")
  const double* lsc;
  const double* ldisp;
  const double* centers;
  int nboxes;
  const uint8_t* levels;
  const int32_t* src_counts;
  const int32_t* src_starts;
  const int32_t* tgt_counts;
  const int32_t* tgt_starts;
  const int32_t* nbr_starts;
  const int32_t* nbr_lists;
  const int32_t* target_boxes;
  const int32_t* src_ids;
  const int32_t* tgt_ids;
  const int64_t* case_indices;
  int n_codes;
  const uint16_t* ranks;
  const int32_t* offs;
  const double* table;
  int64_t n_entries;
  int n_tables;
  const double* nmlz;
  const void* rowmap;
  const double* dense;  // expanded table (DENSE kernels), [n_tables][n_cases][n][n] (x VW)
  int n_cases;
  int dim, n;
  int scaling_mode, scaling_power, log_disp;
  double root_extent, table_root_extent;
  int table_starting_level;
  double enc_shift, enc_bias;
  int enc_base;
  int n_rhs;
  int64_t ldf, ldy;
  const double* f;
  double* y;
  int32_t* err;
  int tbuf;   // doubles of LDS for the staged orbit block
  int nt;     // lanes per target set (power of two)
};

constexpr int kNFBlock = 256;

// CPLX: complex table, densities and result (interleaved re, im doubles;
// leading dimensions in doubles); products are complex, the scaling real.
// DENSE: table values read from the expanded per-(level, case) blocks,
// [s][t] with t fastest, so the lanes of a wave (consecutive target nodes)
// read one coalesced line per source mode instead of a row-map load plus an
// LDS gather; no orbit block is staged.
template <typename IdxT, int RC, bool CPLX, bool DENSE>
__global__ void __launch_bounds__(kNFBlock) near_field_f64_kernel(const NFParams p) {
  extern __shared__ double smem[];
  constexpr int VW = CPLX ? 2 : 1;  // doubles per value
  const int n = p.n;
  const int P = kNFBlock / p.nt;
  double* sT = smem;
  double* sF = sT + p.tbuf;
  double* sD = sF + (size_t)n * RC * VW;
  double* red = sD + n;
  int* sMode = (int*)(red + (P > 1 ? kNFBlock * RC * VW : 0));

  const int tid = threadIdx.x;
  const int lane_t = tid % p.nt;
  const int part = tid / p.nt;

  const int tbox = blockIdx.x;
  const int tb = p.target_boxes[tbox];
  const int tbeg = p.tgt_starts[tb];
  const int ntb = min(p.tgt_counts[tb], n);
  const int l0 = p.nbr_starts[tbox];
  const int l1 = p.nbr_starts[tbox + 1];
  double ct[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < p.dim; ++d) ct[d] = p.centers[(int64_t)d * p.nboxes + tb];

  for (int r0 = 0; r0 < p.n_rhs; r0 += RC) {
    const int rc = min(RC, p.n_rhs - r0);
    for (int t0 = 0; t0 < ntb; t0 += p.nt) {
      const int ti = t0 + lane_t;
      const int tm_raw = ti < ntb ? p.tgt_ids[tbeg + ti] : -1;
      const bool tact = (unsigned)tm_raw < (unsigned)n;
      const int tmode = tact ? tm_raw : 0;
      double acc[RC], aci[RC];  // aci: imaginary parts (complex data only)
#pragma unroll
      for (int r = 0; r < RC; ++r) acc[r] = aci[r] = 0.0;

      for (int li = l0; li < l1; ++li) {
        const int sb = p.nbr_lists[li];
        const int lvl = p.levels[sb];
        const double sext = p.root_extent * (1.0 / (double)(1 << lvl));
        double code = 0.0;
        for (int d = 0; d < p.dim; ++d) {
          const double cs = p.centers[(int64_t)d * p.nboxes + sb];
          code = code * (double)p.enc_base + ((ct[d] - cs) / sext * 4.0 + p.enc_shift + p.enc_bias);
        }
        const int64_t vid = (int64_t)(code + 0.5);
        const int cid = (vid >= 0 && vid < p.n_codes) ? (int)p.case_indices[vid] : -1;
        __syncthreads();  // previous source box done with LDS
        if (cid < 0) {    // uniform over the workgroup
          if (tid == 0 && p.err) atomicOr(p.err, 1);
          continue;
        }
        int tlev = 0;
        double scal = 1.0, dscale = 0.0;
        if (p.scaling_mode == VL1_SCALING_CANONICAL) {
          double num = 1.0, den = 1.0;
          for (int k = 0; k < p.scaling_power; ++k) {
            num *= sext;
            den *= p.table_root_extent;
          }
          scal = num / den;
          if (p.log_disp) dscale = -0.5 / M_PI * scal * log(sext / p.table_root_extent);
        } else if (p.scaling_mode == VL1_SCALING_CUSTOM) {
          scal = p.lsc[min(lvl, VL1_MAX_LEVELS - 1)];
          dscale = p.ldisp[min(lvl, VL1_MAX_LEVELS - 1)];
        } else if (p.scaling_mode == VL1_SCALING_PER_LEVEL) {
          const int rel = lvl - p.table_starting_level;
          tlev = lvl < p.table_starting_level ? 0 : (rel < p.n_tables ? rel : p.n_tables - 1);
        }
        int off = 0, cnt = 0;  // the case's canonical orbit (row-map kernels only)
        if constexpr (!DENSE) {
          const int rank = p.ranks[cid];
          off = p.offs[rank];
          cnt = p.offs[rank + 1] - off;
        }
        const double* gT = p.table + ((int64_t)tlev * p.n_entries + off) * VW;
        const bool staged = !DENSE && cnt * VW <= p.tbuf;
        const double* gD =
            DENSE ? p.dense + (((int64_t)tlev * p.n_cases + cid) * n * n + tmode) * VW : nullptr;
        const int sbeg = p.src_starts[sb];
        const int ns = min(p.src_counts[sb], n);
        if (staged)
          for (int i = tid; i < cnt * VW; i += kNFBlock) sT[i] = gT[i];
        for (int i = tid; i < ns * RC * VW; i += kNFBlock) {
          const int c = i % VW, rs = i / VW;
          const int r = rs / ns, sid = rs % ns;
          sF[(sid * RC + r) * VW + c] =
              r < rc ? p.f[(int64_t)(r0 + r) * p.ldf + (int64_t)(sbeg + sid) * VW + c] : 0.0;
        }
        for (int i = tid; i < ns; i += kNFBlock) {
          sMode[i] = p.src_ids[sbeg + i];
          sD[i] = p.log_disp ? dscale * p.nmlz[(int64_t)tlev * n + i] : 0.0;
        }
        __syncthreads();
        if (tact) {
          const IdxT* rm = static_cast<const IdxT*>(p.rowmap) + (int64_t)cid * n * n + tmode;
          for (int sid = part; sid < ns; sid += P) {
            const int s = sMode[sid];
            if ((unsigned)s >= (unsigned)n) continue;  // malformed mode id: no contribution
            double tr, ti;  // table value (imaginary part for complex data)
            double esc = scal;  // scaling times the entry sign of signed kernels
            if constexpr (DENSE) {
              tr = gD[(int64_t)s * n * VW];
              ti = CPLX ? gD[(int64_t)s * n * VW + 1] : 0.0;
              // dense kind: no entry (NaN) -> no contribution, displacement
              // included (list1.py:1111-1116)
              if (tr != tr) continue;
            } else {
              int rel = (int)rm[(int64_t)s * n];
              if constexpr (std::is_signed<IdxT>::value) {
                double sg;
                rel = decode_signed_row(rel, sg);
                esc *= sg;
              }
              tr = staged ? sT[rel * VW] : gT[rel * VW];
              ti = CPLX ? (staged ? sT[rel * VW + 1] : gT[rel * VW + 1]) : 0.0;
            }
            const double* fs = sF + sid * RC * VW;
            if constexpr (!CPLX) {
              const double v = tr * esc + sD[sid];
#pragma unroll
              for (int r = 0; r < RC; ++r) acc[r] = fma(fs[r], v, acc[r]);
            } else {
              const double vr = tr * esc;
              const double vi = ti * esc;
#pragma unroll
              for (int r = 0; r < RC; ++r) {
                const double fr = fs[2 * r], fi = fs[2 * r + 1];
                acc[r] = fma(fr, vr, fma(-fi, vi, acc[r]));
                aci[r] = fma(fr, vi, fma(fi, vr, aci[r]));
              }
            }
          }
        }
      }

      if (P > 1) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RC; ++r) {
          red[((part * p.nt + lane_t) * RC + r) * VW] = acc[r];
          if (CPLX) red[((part * p.nt + lane_t) * RC + r) * VW + 1] = aci[r];
        }
        __syncthreads();
        if (part == 0) {
          for (int j = 1; j < P; ++j)
#pragma unroll
            for (int r = 0; r < RC; ++r) {
              acc[r] += red[((j * p.nt + lane_t) * RC + r) * VW];
              if (CPLX) aci[r] += red[((j * p.nt + lane_t) * RC + r) * VW + 1];
            }
        }
      }
      if (part == 0 && tact) {
        for (int r = 0; r < rc; ++r) {
          p.y[(int64_t)(r0 + r) * p.ldy + (int64_t)(tbeg + ti) * VW] = acc[r];
          if (CPLX) p.y[(int64_t)(r0 + r) * p.ldy + (int64_t)(tbeg + ti) * VW + 1] = aci[r];
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------
// FP64-MFMA form of the same evaluation (real tables, boxes of at most 256
// nodes): per target box and group of 64 densities, every list-1 pair is a
// product Y[tid][r] += sum_sid A[tid][sid] X[sid][r] with
//   A[tid][sid] = D[lev][case][mode(sid)][mode(tid)]  (expanded blocks),
//   X[sid][r]   = scaling * f[r][sbeg + sid],
// rows indexed by the target particles and K by the source particles of
// the pair, so ragged boxes, permuted, repeated or invalid mode ids keep the
// semantics of the scalar kernel above: an invalid id (or a missing entry,
// NaN in the expanded blocks) gives A = 0.  v_mfma_f64_16x16x4f64 tiles:
// wave w owns row tiles w, w + 8 (16 rows each) and all four 16-column
// tiles; K advances in chunks of 32 source particles.  A chunk's densities
// go through a two-slot LDS ring and its table fragments through
// registers, both loaded one chunk ahead; a pair's source modes (and its
// descriptor) are prepared one pair ahead.  The 2D log displacement is the
// same for every target of a pair: sum_sid nmlz[sid] f[sid], accumulated in
// LDS while the densities are staged and added when the pair ends (only
// used where no entry is missing: the scalar kernel drops the displacement
// of a missing entry, list1.py:1111-1116).
constexpr int kNFPitch = 34;  // doubles per column of a staged chunk (32 + 2)
constexpr int kNFMaxK = 256;  // source particles per box in the MFMA form

typedef __attribute__((ext_vector_type(4))) double nf_doublex4;

struct NFPair {
  int sbeg, ns;
  int64_t dbase;  // first value of the pair's expanded block
  double scal, dscale;
  int tlev;
  bool ok;
};

__device__ inline NFPair nf_pair(const NFParams& p, int li, const double (&ct)[3]) {
  NFPair d;
  const int sb = p.nbr_lists[li];
  const int lvl = p.levels[sb];
  const double sext = p.root_extent * (1.0 / (double)(1 << lvl));
  double code = 0.0;
  for (int k = 0; k < p.dim; ++k) {
    const double cs = p.centers[(int64_t)k * p.nboxes + sb];
    code = code * (double)p.enc_base + ((ct[k] - cs) / sext * 4.0 + p.enc_shift + p.enc_bias);
  }
  const int64_t vid = (int64_t)(code + 0.5);
  const int cid = (vid >= 0 && vid < p.n_codes) ? (int)p.case_indices[vid] : -1;
  d.ok = cid >= 0;
  d.tlev = 0;
  d.scal = 1.0;
  d.dscale = 0.0;
  if (p.scaling_mode == VL1_SCALING_CANONICAL) {
    double num = 1.0, den = 1.0;
    for (int k = 0; k < p.scaling_power; ++k) {
      num *= sext;
      den *= p.table_root_extent;
    }
    d.scal = num / den;
    if (p.log_disp) d.dscale = -0.5 / M_PI * d.scal * log(sext / p.table_root_extent);
  } else if (p.scaling_mode == VL1_SCALING_CUSTOM) {
    d.scal = p.lsc[min(lvl, VL1_MAX_LEVELS - 1)];
    d.dscale = p.ldisp[min(lvl, VL1_MAX_LEVELS - 1)];
  } else if (p.scaling_mode == VL1_SCALING_PER_LEVEL) {
    const int rel = lvl - p.table_starting_level;
    d.tlev = lvl < p.table_starting_level ? 0 : (rel < p.n_tables ? rel : p.n_tables - 1);
  }
  d.dbase = d.ok ? ((int64_t)d.tlev * p.n_cases + cid) * p.n * p.n : 0;
  d.sbeg = p.src_starts[sb];
  d.ns = min(p.src_counts[sb], p.n);
  if (!d.ok) d.ns = 0;
  return d;
}

template <int W, int NRT>
__global__ void __launch_bounds__(64 * W) near_field_mfma_kernel(const NFParams p, int KP) {
  constexpr int T = 64 * W;
  constexpr int VPT = 2048 / T;  // densities of a chunk per thread
  __shared__ __align__(16) double sX[2][64 * kNFPitch];
  __shared__ int sMode[2][kNFMaxK];
  __shared__ double sDisp[2][64];

  const int n = p.n;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;
  const int l16 = lane % 16, lg = lane / 16;
  const int tbox = blockIdx.x;
  const int r0 = blockIdx.y * 64;
  const int rc = min(64, p.n_rhs - r0);
  const int tb = p.target_boxes[tbox];
  const int tbeg = p.tgt_starts[tb];
  const int ntb = min(p.tgt_counts[tb], n);
  const int l0 = p.nbr_starts[tbox];
  const int npairs = p.nbr_starts[tbox + 1] - l0;
  const int nkc = KP / 32;
  double ct[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < p.dim; ++k) ct[k] = p.centers[(int64_t)k * p.nboxes + tb];

  // whether this wave's second row tile holds any of the box's targets
  const bool rows_rt1 = NRT > 1 && 16 * (wave + W) < ntb;
  // this lane's A rows (target particles) and their modes (-1: no row)
  int tmA[NRT];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt) {
    const int row = 16 * (wave + W * rt) + l16;
    const int m = row < ntb ? p.tgt_ids[tbeg + row] : -1;
    tmA[rt] = (unsigned)m < (unsigned)n ? m : -1;
  }

  // pair k's source modes into sMode[k & 1] (-1 past the box or invalid);
  // its displacement sums start at zero
  auto prep_pair = [&](const NFPair& d, int slot) {
    for (int i = tid; i < KP; i += T) {
      const int m = i < d.ns ? p.src_ids[d.sbeg + i] : -1;
      sMode[slot][i] = (unsigned)m < (unsigned)n ? m : -1;
    }
    if (tid < 64) sDisp[slot][tid] = 0.0;
  };
  // densities of chunk kc of a pair: thread value v is column (tid + T v) / 32,
  // source particle 32 kc + (tid + T v) % 32 (consecutive threads read
  // consecutive particles of one density)
  auto load_x = [&](const NFPair& d, int kc, double (&x)[VPT]) {
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
      const int u = tid + T * v;
      const int col = u / 32, sid = 32 * kc + u % 32;
      x[v] = (col < rc && sid < d.ns) ? p.f[(int64_t)(r0 + col) * p.ldf + d.sbeg + sid] : 0.0;
    }
  };
  auto store_x = [&](const NFPair& d, int kc, int slot, int ring, const double (&x)[VPT]) {
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
      const int u = tid + T * v;
      const int col = u / 32, k = u % 32, sid = 32 * kc + k;
      const bool live = sMode[slot][sid] >= 0;
      sX[ring][col * kNFPitch + k] = live ? d.scal * x[v] : 0.0;
      if (p.log_disp && live && x[v] != 0.0)
        atomicAdd(&sDisp[slot][col], p.nmlz[(int64_t)d.tlev * n + sid] * x[v]);
    }
  };
  // table fragments of chunk kc: lane (l16, lg) of k group u holds
  // A[row 16 rt16 + l16][sid 32 kc + 4 u + lg]
  auto load_a = [&](const NFPair& d, int kc, int slot, double (&a)[NRT][8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int m = sMode[slot][32 * kc + 4 * u + lg];
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) {
        double v = 0.0;
        if (m >= 0 && tmA[rt] >= 0) v = p.dense[d.dbase + (int64_t)m * n + tmA[rt]];
        a[rt][u] = v == v ? v : 0.0;  // NaN: no entry
      }
    }
  };

  nf_doublex4 yacc[NRT][4];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int c = 0; c < 4; ++c) yacc[rt][c] = nf_doublex4{0.0, 0.0, 0.0, 0.0};

  NFPair pdC{}, pdN{};
  double ac[NRT][8], an[NRT][8];
  const int total = npairs * nkc;
  if (total > 0) {
    pdC = nf_pair(p, l0, ct);
    if (tid == 0 && !pdC.ok && p.err) atomicOr(p.err, 1);
    prep_pair(pdC, 0);
    if (npairs > 1) {
      pdN = nf_pair(p, l0 + 1, ct);
      prep_pair(pdN, 1);
    }
    __syncthreads();
    double x[VPT];
    load_x(pdC, 0, x);
    store_x(pdC, 0, 0, 0, x);
    load_a(pdC, 0, 0, ac);
  }
  __syncthreads();

  for (int it = 0; it < total; ++it) {
    const int k = it / nkc, kc = it % nkc;
    if (kc == 0 && k > 0) {  // pair k begins: the next pair's descriptor and modes
      pdC = pdN;
      if (tid == 0 && !pdC.ok && p.err) atomicOr(p.err, 1);
      if (k + 1 < npairs) {
        pdN = nf_pair(p, l0 + k + 1, ct);
        prep_pair(pdN, (k + 1) & 1);
      }
      if (nkc == 1) __syncthreads();  // read within this same iteration
    }
    const bool has_next = it + 1 < total;
    const bool pair_end = kc + 1 == nkc;
    const int nk = pair_end ? k + 1 : k, nkc_ = pair_end ? 0 : kc + 1;
    const NFPair& pdX = pair_end ? pdN : pdC;
    double x[VPT];
    if (has_next) {
      load_x(pdX, nkc_, x);
      load_a(pdX, nkc_, nk & 1, an);
    }
    const double* rX = sX[it & 1];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      double bx[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) bx[c] = rX[(16 * c + l16) * kNFPitch + 4 * u + lg];
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
        if (rt == 0 || rows_rt1)  // a second row tile past the box's targets is skipped
#pragma unroll
          for (int c = 0; c < 4; ++c)
            yacc[rt][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[rt][u], bx[c], yacc[rt][c], 0, 0, 0);
    }
    if (pair_end && p.log_disp) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double ds = pdC.dscale * sDisp[k & 1][16 * c + l16];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
          for (int e = 0; e < 4; ++e) yacc[rt][c][e] += ds;
      }
    }
    if (has_next) {
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int u = 0; u < 8; ++u) ac[rt][u] = an[rt][u];
      store_x(pdX, nkc_, nk & 1, (it + 1) & 1, x);
    }
    __syncthreads();
  }

  // outputs: lane holds rows 16 rt16 + lg + 4 e, column 16 c + l16
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * (wave + W * rt) + lg + 4 * e;
      if (row >= ntb) continue;
      const int m = p.tgt_ids[tbeg + row];
      if ((unsigned)m >= (unsigned)n) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int col = 16 * c + l16;
        if (col < rc) p.y[(int64_t)(r0 + col) * p.ldy + tbeg + row] = yacc[rt][c][e];
      }
    }
  }
}

// Expanded table for the DENSE kernels: D[lev][case][s][t] = sign * T[lev,
// row(case, s, t)], read through the payload's row map.
template <typename IdxT>
__global__ void __launch_bounds__(256)
dense_table_kernel(int n, int n_cases, int n_tables, int vw, const IdxT* __restrict__ rowmap,
                   const uint16_t* __restrict__ ranks, const int32_t* __restrict__ offs,
                   const double* __restrict__ table, int64_t n_entries, double* __restrict__ D) {
  const int64_t nn = (int64_t)n * n;
  const int64_t total = (int64_t)n_tables * n_cases * nn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cst = i % ((int64_t)n_cases * nn);
    const int lev = (int)(i / ((int64_t)n_cases * nn));
    const int c = (int)(cst / nn);
    int rel = (int)rowmap[cst];
    double sg = 1.0;
    if constexpr (std::is_signed<IdxT>::value) rel = decode_signed_row(rel, sg);
    const double* T = table + ((int64_t)lev * n_entries + offs[ranks[c]] + rel) * vw;
    D[i * vw] = sg * T[0];
    if (vw == 2) D[i * vw + 1] = sg * T[1];
  }
}

// Build the payload's expanded table once (first plain evaluation).
static hipError_t ensure_dense(const vl1_payload* pl, hipStream_t s) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (pl->dense.bytes) return hipSuccess;
  const int vw = pl->is_complex ? 2 : 1;
  const int64_t total = (int64_t)pl->n_tables * pl->n_cases * pl->n * pl->n;
  hipError_t e = pl->dense.alloc((size_t)total * vw * sizeof(double));
  if (e != hipSuccess) return e;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
  if (pl->rowmap_u16)
    hipLaunchKernelGGL(dense_table_kernel<uint16_t>, dim3(grid), dim3(256), 0, s, pl->n,
                       pl->n_cases, pl->n_tables, vw, pl->rowmap.as<uint16_t>(),
                       pl->ranks.as<uint16_t>(), pl->offs.as<int32_t>(), pl->table.as<double>(),
                       pl->n_entries, pl->dense.as<double>());
  else
    hipLaunchKernelGGL(dense_table_kernel<int32_t>, dim3(grid), dim3(256), 0, s, pl->n,
                       pl->n_cases, pl->n_tables, vw, pl->rowmap.as<int32_t>(),
                       pl->ranks.as<uint16_t>(), pl->offs.as<int32_t>(), pl->table.as<double>(),
                       pl->n_entries, pl->dense.as<double>());
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) pl->dense.release();
  return e;
}

template <typename IdxT, int RC, bool CPLX, bool DENSE>
static hipError_t launch_nf(const NFParams& prm, int grid, size_t smem, hipStream_t s) {
  auto fn = near_field_f64_kernel<IdxT, RC, CPLX, DENSE>;
  hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kNFBlock), smem, s, prm);
  return hipGetLastError();
}

}  // namespace vl1

using namespace vl1;

// Potential kind 2 (list1.py:1097-1143): y[t] += f[t] x exterior_nmlz[lev_t][mode_t] x
// the target box's scaling, after the list-1 sums.  One thread per (RHS,
// target box, box-local target); real tables, no displacement.
__global__ void exterior_term_kernel(const NFParams p, const double* __restrict__ ext,
                                     int64_t ld_f, int64_t ld_y, int n_tgt) {
  const int64_t total = (int64_t)p.n_rhs * n_tgt * p.n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(i % p.n);
    const int tbox = (int)((i / p.n) % n_tgt);
    const int r = (int)(i / ((int64_t)p.n * n_tgt));
    const int tb = p.target_boxes[tbox];
    if (t >= min(p.tgt_counts[tb], p.n)) continue;
    const int lvl = p.levels[tb];
    const double text = p.root_extent * (1.0 / (double)(1 << lvl));
    double scal = 1.0;
    int tlev = 0;
    if (p.scaling_mode == VL1_SCALING_CANONICAL) {
      double num = 1.0, den = 1.0;
      for (int k = 0; k < p.scaling_power; ++k) {
        num *= text;
        den *= p.table_root_extent;
      }
      scal = num / den;
    } else if (p.scaling_mode == VL1_SCALING_CUSTOM) {
      scal = p.lsc[min(lvl, VL1_MAX_LEVELS - 1)];
    } else if (p.scaling_mode == VL1_SCALING_PER_LEVEL) {
      const int rel = lvl - p.table_starting_level;
      tlev = lvl < p.table_starting_level ? 0 : (rel < p.n_tables ? rel : p.n_tables - 1);
    }
    const int64_t pt = (int64_t)p.tgt_starts[tb] + t;
    const double e = ext[(int64_t)tlev * p.n + p.tgt_ids[pt]] * scal;
    p.y[(int64_t)r * ld_y + pt] += p.f[(int64_t)r * ld_f + pt] * e;
  }
}

static int exterior_pass(const NFParams& prm, const vl1_near_field_args* a, hipStream_t s) {
  if (!a->exterior_nmlz) return VL1_OK;
  const int64_t total = (int64_t)a->n_rhs * a->n_tgt_boxes * prm.n;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(exterior_term_kernel, dim3(grid), dim3(256), 0, s, prm, a->exterior_nmlz,
                     a->n_rhs > 1 ? a->ld_coefs : 0, a->n_rhs > 1 ? a->ld_result : 0,
                     a->n_tgt_boxes);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(VL1_E_HIP, std::string("exterior term launch: ") + hipGetErrorString(e));
  return VL1_OK;
}

extern "C" int vl1_near_field_f64(const vl1_payload* pl, const vl1_near_field_args* a,
                                  void* stream) {
  clear_error();
  if (!pl || !a) return fail(VL1_E_ARG, "vl1_near_field_f64: null argument");
  if (pl->is_complex && a->log_displacement && a->scaling_mode == VL1_SCALING_CANONICAL)
    return fail(VL1_E_UNSUPPORTED, "vl1_near_field_f64: the 2D log displacement applies to real tables");
  if (a->n_tgt_boxes < 0 || a->n_rhs < 1 || a->aligned_nboxes < 1)
    return fail(VL1_E_ARG, "vl1_near_field_f64: bad sizes");
  if (a->scaling_mode < 0 || a->scaling_mode > 3)
    return fail(VL1_E_ARG, "vl1_near_field_f64: bad scaling mode");
  if (a->scaling_mode == VL1_SCALING_CUSTOM && (!a->level_scaling || !a->level_disp))
    return fail(VL1_E_ARG, "vl1_near_field_f64: custom scaling needs level_scaling and level_disp");
  if (a->scaling_mode == VL1_SCALING_CANONICAL && (a->scaling_power < 0 || a->scaling_power > 3))
    return fail(VL1_E_ARG, "vl1_near_field_f64: bad scaling power");
  if (a->n_tgt_boxes == 0) return VL1_OK;
  if (!a->result || !a->mode_coefs || !a->box_centers || !a->box_levels ||
      !a->box_source_counts_nonchild || !a->box_source_starts ||
      !a->box_target_counts_nonchild || !a->box_target_starts ||
      !a->neighbor_source_boxes_starts || !a->neighbor_source_boxes_lists || !a->target_boxes ||
      !a->source_mode_ids || !a->target_point_ids)
    return fail(VL1_E_ARG, "vl1_near_field_f64: null device array");
  if (a->n_rhs > 1 && (a->ld_coefs < 1 || a->ld_result < 1))
    return fail(VL1_E_ARG, "vl1_near_field_f64: bad leading dimension");
  if (a->exterior_nmlz &&
      (pl->is_complex || (a->log_displacement && (a->scaling_mode == VL1_SCALING_CANONICAL ||
                                                  a->scaling_mode == VL1_SCALING_CUSTOM))))
    return fail(VL1_E_UNSUPPORTED,
                "vl1_near_field_f64: potential kind 2 takes real tables without a displacement");

  NFParams prm{};
  prm.centers = a->box_centers;
  prm.nboxes = a->aligned_nboxes;
  prm.levels = a->box_levels;
  prm.src_counts = a->box_source_counts_nonchild;
  prm.src_starts = a->box_source_starts;
  prm.tgt_counts = a->box_target_counts_nonchild;
  prm.tgt_starts = a->box_target_starts;
  prm.nbr_starts = a->neighbor_source_boxes_starts;
  prm.nbr_lists = a->neighbor_source_boxes_lists;
  prm.target_boxes = a->target_boxes;
  prm.src_ids = a->source_mode_ids;
  prm.tgt_ids = a->target_point_ids;
  prm.case_indices = pl->case_idx.as<int64_t>();
  prm.n_codes = (int)pl->case_indices.size();
  prm.ranks = pl->ranks.as<uint16_t>();
  prm.offs = pl->offs.as<int32_t>();
  prm.table = pl->table.as<double>();
  prm.n_entries = pl->n_entries;
  prm.n_tables = pl->n_tables;
  prm.nmlz = pl->nmlz.as<double>();
  prm.rowmap = pl->rowmap.ptr;
  prm.dim = pl->dim;
  prm.n = pl->n;
  prm.scaling_mode = a->scaling_mode;
  prm.scaling_power = a->scaling_power;
  prm.lsc = a->scaling_mode == VL1_SCALING_CUSTOM ? a->level_scaling : nullptr;
  prm.ldisp = a->scaling_mode == VL1_SCALING_CUSTOM ? a->level_disp : nullptr;
  prm.log_disp = a->log_displacement ? 1 : 0;
  if (a->scaling_mode == VL1_SCALING_CUSTOM) {
    // the displacement term is on unless the caller declares it zero
    prm.log_disp = a->log_displacement ? 1 : 0;
    if (prm.log_disp && pl->is_complex)
      return fail(VL1_E_UNSUPPORTED,
                  "vl1_near_field_f64: a custom displacement applies to real tables");
  }
  prm.root_extent = a->root_extent;
  prm.table_root_extent = a->table_root_extent;
  prm.table_starting_level = a->table_starting_level;
  prm.enc_shift = a->encoding_shift;
  prm.enc_bias = a->case_encoding_bias;
  prm.enc_base = a->encoding_base;
  prm.n_rhs = a->n_rhs;
  prm.ldf = a->ld_coefs;
  prm.ldy = a->ld_result;
  prm.f = a->mode_coefs;
  prm.y = a->result;
  prm.err = a->error_flags;
  if (prm.enc_base != 13) return fail(VL1_E_ARG, "encoding_base must be 13");

  int nt = 1;
  while (nt < pl->n && nt < kNFBlock) nt *= 2;
  prm.nt = nt;
  const int P = kNFBlock / nt;
  // densities per pass (32 measured 4.6x slower than 16 on 3D q=6: register
  // pressure)
  const int RC = a->n_rhs == 1 ? 1 : (a->n_rhs <= 4 ? 4 : 16);
  const size_t lds_cap = 150 * 1024;
  const size_t vw = pl->is_complex ? 2 : 1;
  hipStream_t s = (hipStream_t)stream;
  // expanded table: coalesced reads per source mode, no row-map gathers;
  // used while it stays under 8 GiB (a cache of the payload)
  // (VL1_PLAIN_TABLE=gather in the environment selects the row-map path)
  const char* mode_env = std::getenv("VL1_PLAIN_TABLE");
  const bool force_gather = mode_env && std::string(mode_env) == "gather";
  const bool dense = pl->dense_kind ||  // dense-kind payloads hold only the expanded blocks
                     (!force_gather && (double)pl->n_tables * pl->n_cases * pl->n * pl->n * vw *
                                               sizeof(double) <= 8.0 * (1ull << 30));
  if (dense) {
    hipError_t de = ensure_dense(pl, s);
    if (de != hipSuccess)
      return fail(VL1_E_HIP, std::string("vl1_near_field_f64: dense table: ") + hipGetErrorString(de));
    prm.dense = pl->dense.as<double>();
  }
  prm.n_cases = pl->n_cases;
  // FP64-MFMA form: real tables, boxes of at most 256 nodes, 16 or more
  // densities (a pass always computes 64 columns), and no missing entry
  // under the log displacement (VL1_PLAIN_KERNEL=scalar selects the scalar
  // kernel below)
  const char* kern_env = std::getenv("VL1_PLAIN_KERNEL");
  const bool force_scalar = kern_env && std::string(kern_env) == "scalar";
  if (dense && !force_scalar && !pl->is_complex && pl->n <= kNFMaxK && a->n_rhs >= 16 &&
      !(pl->has_missing && prm.log_disp)) {
    const int m16 = (pl->n + 15) / 16;
    const int KP = (pl->n + 31) / 32 * 32;
    const dim3 grid(a->n_tgt_boxes, (a->n_rhs + 63) / 64);
    hipStream_t st = (hipStream_t)stream;
    if (m16 <= 1)
      hipLaunchKernelGGL((near_field_mfma_kernel<1, 1>), grid, dim3(64), 0, st, prm, KP);
    else if (m16 <= 2)
      hipLaunchKernelGGL((near_field_mfma_kernel<2, 1>), grid, dim3(128), 0, st, prm, KP);
    else if (m16 <= 4)
      hipLaunchKernelGGL((near_field_mfma_kernel<4, 1>), grid, dim3(256), 0, st, prm, KP);
    else if (m16 <= 8)
      hipLaunchKernelGGL((near_field_mfma_kernel<8, 1>), grid, dim3(512), 0, st, prm, KP);
    else
      hipLaunchKernelGGL((near_field_mfma_kernel<8, 2>), grid, dim3(512), 0, st, prm, KP);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return fail(VL1_E_HIP, std::string("near_field_mfma launch: ") + hipGetErrorString(e));
    return exterior_pass(prm, a, st);
  }
  const size_t fixed =
      sizeof(double) * ((size_t)pl->n * RC * vw + pl->n + (P > 1 ? kNFBlock * RC * vw : 0)) +
      sizeof(int) * pl->n;
  if (fixed > lds_cap) return fail(VL1_E_UNSUPPORTED, "vl1_near_field_f64: box too large for LDS");
  size_t tbuf = dense ? 0
                      : std::min<size_t>((size_t)pl->max_orbit_count * vw,
                                         (lds_cap - fixed) / sizeof(double));
  prm.tbuf = (int)tbuf;
  const size_t smem = fixed + sizeof(double) * tbuf;
  hipError_t e;
  const bool u16 = pl->rowmap_u16 != 0;
  const int grid = a->n_tgt_boxes;
#define VL1_NF_LAUNCH(RCV, CPX)                                                                  \
  (dense ? launch_nf<uint16_t, RCV, CPX, true>(prm, grid, smem, s)                               \
         : (u16 ? launch_nf<uint16_t, RCV, CPX, false>(prm, grid, smem, s)                       \
                : launch_nf<int32_t, RCV, CPX, false>(prm, grid, smem, s)))
  if (pl->is_complex)
    e = RC == 1 ? VL1_NF_LAUNCH(1, true) : (RC == 4 ? VL1_NF_LAUNCH(4, true) : VL1_NF_LAUNCH(16, true));
  else
    e = RC == 1 ? VL1_NF_LAUNCH(1, false) : (RC == 4 ? VL1_NF_LAUNCH(4, false) : VL1_NF_LAUNCH(16, false));
#undef VL1_NF_LAUNCH
  if (e != hipSuccess)
    return fail(VL1_E_HIP, std::string("near_field_f64 launch: ") + hipGetErrorString(e));
  return exterior_pass(prm, a, s);
}
