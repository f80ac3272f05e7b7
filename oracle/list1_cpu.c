/*
 * Oracle / CPU baseline: the list-1 `near_field` kernel in C with OpenMP.
 * TEST INFRASTRUCTURE ONLY -- timed as the CPU baseline and used as a
 * checker; never linked into the product library.
 *
 * Restates volumential/list1.py:1057-1148: one task per target box
 * (dynamic schedule); for each source box the case id from the float
 * encoding (list1.py:439-459, 1083-1091), the scaling policy
 * (list1.py:361-437), then for every (target node, source node) the
 * arithmetic orbit reconstruction of the compact row
 * (list1.py:604-902 / orbit_arithmetic.py:372-634, with the entry sign of
 * axis-derivative kernels) and a table gather.
 * Like the reference, each density (RHS) is a separate evaluation.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Compact row of (s, t); *neg = 1 when the entry is read negated: an odd
 * number of flipped axes whose input axis has its bit set in sp_mask (the
 * axis-derivative sign of orbit_arithmetic.py:372-432). */
static inline int64_t compact_row(int dim, int q, const uint8_t* perm, const int8_t* sign,
                                  const uint8_t* group, int sp_mask, int s, int t, int* neg) {
  int sr[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
  for (int ax = dim - 1; ax >= 0; --ax) {
    sr[ax] = s % q;
    s /= q;
    tr[ax] = t % q;
    t /= q;
  }
  int pair[3];
  *neg = 0;
  for (int o = 0; o < dim; ++o) {
    int in = perm[o];
    int si = sr[in], ti = tr[in];
    int sf = q - 1 - si, tf = q - 1 - ti;
    int sg = sign[o];
    if (sg < 0 || (sg == 0 && (sf < si || (sf == si && tf < ti)))) {
      si = sf;
      ti = tf;
      if ((sp_mask >> in) & 1) *neg ^= 1;
    }
    pair[o] = si * q + ti;
  }
  const int64_t pc = (int64_t)q * q, fpc = (pc + 1) / 2;
  int64_t comp = 0;
  for (int g = 0; g < dim; ++g) {
    int64_t lo = pc, hi = -1, sum = 0;
    int k = 0, zero = 0;
    for (int o = 0; o < dim; ++o)
      if (group[o] == g) {
        int v = pair[o];
        if (v < lo) lo = v;
        if (v > hi) hi = v;
        sum += v;
        ++k;
        zero |= sign[o] == 0;
      }
    if (k == 0) continue;
    const int64_t A = zero ? fpc : pc;
    const int64_t mid = sum - lo - hi;
    int64_t r, c;
    if (k == 1) {
      r = lo;
      c = A;
    } else if (k == 2) {
      r = A * (A + 1) / 2 - (A - lo) * (A - lo + 1) / 2 + (hi - lo);
      c = A * (A + 1) / 2;
    } else {
      r = A * (A + 1) * (A + 2) / 6 - (A - lo) * (A - lo + 1) * (A - lo + 2) / 6 +
          (A - lo) * (A - lo + 1) / 2 - (A - mid) * (A - mid + 1) / 2 + (hi - mid);
      c = A * (A + 1) * (A + 2) / 6;
    }
    comp = comp * c + r;
  }
  return comp;
}

/* mode: 0 canonical, 1 fixed, 2 per-level.  Returns 0, or -1 on an
 * invalid case id. */
int oracle_near_field_f64(int dim, int q, int n_tgt_boxes, int nboxes, const double* centers,
                          const uint8_t* levels, const int32_t* box_starts,
                          const int32_t* box_counts, const int32_t* target_boxes,
                          const int32_t* nbr_starts, const int32_t* nbr_lists,
                          const int32_t* src_ids, const int32_t* tgt_ids, const int64_t* case_indices,
                          int n_codes, const uint16_t* ranks, const uint8_t* perm,
                          const int8_t* sign, const uint8_t* group, const int32_t* offsets,
                          const double* table, int64_t n_entries, int n_tables,
                          const double* nmlz, int mode, int power, int log_disp,
                          double root_extent, double table_root_extent, int tsl, double bias,
                          int n_rhs, const double* coefs, int64_t ldf, double* result,
                          int64_t ldy, int nthreads, int sp_mask) {
  int bad = 0;
  const int n = (int)lround(pow((double)q, (double)dim));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int r = 0; r < n_rhs; ++r) {
    const double* f = coefs + (int64_t)r * ldf;
    double* y = result + (int64_t)r * ldy;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int i = 0; i < n_tgt_boxes; ++i) {
      const int tb = target_boxes[i];
      const int tbeg = box_starts[tb], nt = box_counts[tb];
      double ct[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) ct[d] = centers[(int64_t)d * nboxes + tb];
      for (int tid = 0; tid < nt; ++tid) {
        const int t = tgt_ids[tbeg + tid];
        double acc = 0.0;
        for (int li = nbr_starts[i]; li < nbr_starts[i + 1]; ++li) {
          const int sb = nbr_lists[li];
          const int lvl = levels[sb];
          const double sext = root_extent * (1.0 / (double)(1 << lvl));
          double code = 0.0;
          for (int d = 0; d < dim; ++d)
            code = code * 13.0 +
                   ((ct[d] - centers[(int64_t)d * nboxes + sb]) / sext * 4.0 + 6.0 + bias);
          const int64_t vid = (int64_t)(code + 0.5);
          const int cid = (vid >= 0 && vid < n_codes) ? (int)case_indices[vid] : -1;
          if (cid < 0) {
            bad = 1;
            continue;
          }
          int lev = 0;
          double scal = 1.0, dsc = 0.0;
          if (mode == 0) {
            double num = 1.0, den = 1.0;
            for (int k = 0; k < power; ++k) {
              num *= sext;
              den *= table_root_extent;
            }
            scal = num / den;
            if (log_disp) dsc = -0.5 / M_PI * scal * log(sext / table_root_extent);
          } else if (mode == 2) {
            const int rel = lvl - tsl;
            lev = lvl < tsl ? 0 : (rel < n_tables ? rel : n_tables - 1);
          }
          const int rk = ranks[cid];
          const double* T = table + (int64_t)lev * n_entries + offsets[rk];
          const uint8_t* pp = perm + cid * dim;
          const int8_t* ss = sign + cid * dim;
          const uint8_t* gg = group + cid * dim;
          const int sbeg = box_starts[sb], ns = box_counts[sb];
          for (int sid = 0; sid < ns; ++sid) {
            const int s = src_ids[sbeg + sid];
            int neg;
            const double tv = T[compact_row(dim, q, pp, ss, gg, sp_mask, s, t, &neg)];
            const double integ = (neg ? -tv : tv) * scal +
                                 (log_disp ? dsc * nmlz[(int64_t)lev * n + sid] : 0.0);
            acc += f[sbeg + sid] * integ;
          }
        }
        y[tbeg + tid] = acc;
      }
    }
  }
  return bad ? -1 : 0;
}
