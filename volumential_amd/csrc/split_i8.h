// Ozaki split on int8 MFMA with exact cross-box integer sums.
//
// Rows of the (case, level) blocks and the densities are cut into S int8
// digits: a value v = scale * sum_i d_i / (127 * 254^i), d_i in [-127, 127]
// (d_0 = rint(127 x), then d_i = rint(254 r_{i-1}) of the remainders, so
// every digit fits int8 and the expansion carries 7 + 8 (S - 1) bits).
// The row scale of block row t is the largest |entry| of row t over all the
// plan's blocks (rho[t]); the column scale of RHS r is the largest |f alpha|
// over all boxes (sigma[r]), with the canonical pair scaling alpha of the
// source box folded into its digits.  With scales that do not change from
// one (case, level) step to the next, the digit products of a whole tile --
// every K chunk of every list-1 source box -- accumulate exactly in int32,
// one accumulator set per diagonal d = i + j < S, and are turned into FP64
// once per tile:
//     y[t][c] = rho[t] sigma[r] sum_d acc_d[t][c] / (127^2 254^d) + sum disp.
// v_mfma_i32_32x32x32_i8 runs at twice the BF16 MFMA rate.
//
// Error estimate: the last diagonal kept, D_{S-1} = acc_{S-1} W_{S-1},
// bounds the size of the first dropped one (about D_{S-1} / 254, S + 1
// terms against S); est[t][c] = 4 |D_{S-1}| / 254 rho sigma per output.
// Included by plan.hip inside namespace vl1, after TileParams, xcd_tile,
// out_index and static_for.
#pragma once

typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;

constexpr int kI8MaxSlices = 6;
constexpr int kI8Ring = 8;  // density ring slots (chunks in flight)
#ifndef VL1_I8_PRODUCER_DEPTH
#define VL1_I8_PRODUCER_DEPTH 2
#endif
// chunks whose global loads the producer wave keeps in flight (registers)
constexpr int kI8ProducerDepth = VL1_I8_PRODUCER_DEPTH;

// 1 / (127^2 254^d)
__host__ __device__ constexpr double i8_weight(int d) {
  double w = 1.0 / (127.0 * 127.0);
  for (int i = 0; i < d; ++i) w /= 254.0;
  return w;
}

// Digit products (i, j), i, j < S, on diagonal i + j = d.
__host__ __device__ constexpr int i8_diag_terms(int S, int d) { return d < S ? d + 1 : 2 * S - 1 - d; }

// rho[g][t] = max over the plan's (case, level) blocks of max_s |M[k G + g][t][s]|
// (row group g, row t); one wavefront per (g, t).
__global__ void i8_rowmax_kernel(int MP, int KP, int n_keys, int G, const double* __restrict__ M,
                                 double* __restrict__ rho) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (row >= G * MP) return;
  const int g = row / MP, t = row % MP;
  double mx = 0.0;
  for (int k = 0; k < n_keys; ++k) {
    const double* m = M + ((int64_t)(k * G + g) * MP + t) * KP;
    for (int s = lane; s < KP; s += 64) mx = fmax(mx, fabs(m[s]));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if (lane == 0) rho[row] = mx;
}

// Digits of one value in [-1, 1]: d_0 = rint(127 x), d_i = rint(254 r).
template <int S>
__device__ inline void i8_digits(double x, int8_t (&d)[S]) {
  double t = 127.0 * x;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const double r = rint(t);
    d[i] = (int8_t)(int)r;
    t = 254.0 * (t - r);
  }
}

// Block rows -> int8 digit slices in the fragment order of
// v_mfma_i32_32x32x32_i8: for row tile rt (32 rows), K chunk kc (32
// columns) and slice i, lane l holds row 32 rt + l % 32, columns
// 32 kc + 16 (l / 32) + 0..15 as 16 bytes; one 1 KiB block per
// (block, rt, kc, i).  One thread per (block, row, 16-column group).
template <int S>
__global__ void i8_split_rows_kernel(int MP, int KP, int n_blocks, int G,
                                     const double* __restrict__ M, const double* __restrict__ rho,
                                     int8_t* __restrict__ A8) {
  const int64_t total = (int64_t)n_blocks * MP * (KP / 16);
  const int nrt = MP / 32, nkc = KP / 32;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int sg = (int)(i % (KP / 16));
    const int t = (int)((i / (KP / 16)) % MP);
    const int b = (int)(i / ((int64_t)(KP / 16) * MP));
    const double r = rho[(b % G) * MP + t];
    const double inv = r > 0.0 ? 1.0 / r : 0.0;
    const double* m = M + ((int64_t)b * MP + t) * KP + sg * 16;
    int8_t dg[16][S];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      // |M| <= rho: x in [-1, 1] (the division may round above 1 by an ulp;
      // clamp so d_0 stays in [-127, 127])
      const double x = fmin(1.0, fmax(-1.0, m[u] * inv));
      i8_digits<S>(x, dg[u]);
    }
    const int rt = t / 32, kc = sg / 2, h = sg % 2;
    const int lane = t % 32 + 32 * h;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      uint32_t w[4];
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4)
        w[q4] = (uint32_t)(uint8_t)dg[4 * q4][j] | ((uint32_t)(uint8_t)dg[4 * q4 + 1][j] << 8) |
                ((uint32_t)(uint8_t)dg[4 * q4 + 2][j] << 16) | ((uint32_t)(uint8_t)dg[4 * q4 + 3][j] << 24);
      *reinterpret_cast<uint4*>(A8 + (((((int64_t)b * nrt + rt) * nkc + kc) * S + j) * 64 + lane) * 16) =
          make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

// Density staging, pass 1: colmax[box][r] = max |f| alpha(box) over the
// column and, for the 2D log term, nmdot[box][r] = sum_sid nmlz[sid]
// f[sbeg + sid].  A source box's particles are contiguous from
// box_starts[box], so both read the column in place.  One wavefront per RHS
// row and 8 consecutive boxes (4 when KW = 4), a column taking KW loads of
// 64 values (columns of up to 64 KW values; longer ones loop): all loads
// issued before the first reduction.  A box that is no list-1 source has
// alpha 0 (the plan rejects a zero pair scaling) and gets 0 without reading.
// Boxes of at most 128 values per density take i8_colmax_lanes_kernel.
template <int KW>
__global__ void i8_colmax_kernel(int n, int cplx, int nboxes, int R, const double* __restrict__ f,
                                 int64_t ldf, const int32_t* __restrict__ box_starts,
                                 const double* __restrict__ box_alpha,
                                 const double* __restrict__ nmlz, int log_disp,
                                 double* __restrict__ colmax, double* __restrict__ nmdot) {
  constexpr int NB = KW == 4 ? 4 : 8;  // boxes per wavefront (registers)
  const int ngrp = (nboxes + NB - 1) / NB;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (w >= (int64_t)ngrp * R) return;
  const int r = (int)(w / ngrp), b0 = (int)(w % ngrp) * NB;
  const double* fr = f + (int64_t)r * ldf;
  const int rows = cplx ? 2 * n : n;
  double alpha[NB], mx[NB];
  const double* fb[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int box = b0 + b;
    alpha[b] = box < nboxes ? box_alpha[box] : 0.0;
    const int64_t bs = alpha[b] != 0.0 ? box_starts[box] : 0;
    fb[b] = fr + (cplx ? 2 * bs : bs);
    mx[b] = 0.0;
  }
  if (rows <= 64 * KW) {
    double v[NB][KW];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int k = 0; k < KW; ++k)
        v[b][k] = (alpha[b] != 0.0 && lane + 64 * k < rows) ? fb[b][lane + 64 * k] : 0.0;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int k = 0; k < KW; ++k) mx[b] = fmax(mx[b], fabs(v[b][k]));
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (alpha[b] != 0.0)
        for (int q = lane; q < rows; q += 64) mx[b] = fmax(mx[b], fabs(fb[b][q]));
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    double m = mx[b], dot = 0.0;
    if (log_disp && alpha[b] != 0.0)
      for (int q = lane; q < n; q += 64) dot += nmlz[q] * fb[b][q];
    for (int o = 32; o > 0; o >>= 1) {
      m = fmax(m, __shfl_xor(m, o));
      dot += __shfl_xor(dot, o);
    }
    const int box = b0 + b;
    if (lane == 0 && box < nboxes) {
      colmax[(int64_t)box * R + r] = m * fabs(alpha[b]);
      nmdot[(int64_t)box * R + r] = dot;
    }
  }
}

// Column maxima, one lane per (box, RHS): a lane streams its box's `rows`
// values (16-byte loads where the box starts 16-byte aligned) and keeps the
// maximum (and the 2D log term's dot with the normalisers) in registers --
// no cross-lane reduction, every lane busy whatever the box size.  The lanes
// of a wavefront take consecutive boxes of one RHS, so a load instruction's
// addresses advance box by box and the lines each lane touches are reused
// by its next loads from L1.
__global__ void __launch_bounds__(256)
    i8_colmax_lanes_kernel(int n, int cplx, int nboxes, int R, const double* __restrict__ f,
                           int64_t ldf, const int32_t* __restrict__ box_starts,
                           const double* __restrict__ box_alpha, const double* __restrict__ nmlz,
                           int log_disp, double* __restrict__ colmax, double* __restrict__ nmdot) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nbp = ((int64_t)nboxes + 63) / 64 * 64;  // boxes padded to whole waves
  if (gid >= nbp * R) return;
  const int r = (int)(gid / nbp), box = (int)(gid % nbp);
  if (box >= nboxes) return;
  const double alpha = box_alpha[box];
  double mx = 0.0, dot = 0.0;
  if (alpha != 0.0) {
    const int rows = cplx ? 2 * n : n;
    const int64_t bs = box_starts[box];
    const double* fb = f + (int64_t)r * ldf + (cplx ? 2 * bs : bs);
    int q = 0;
    if ((reinterpret_cast<uintptr_t>(fb) & 15) == 0) {
#pragma unroll 8
      for (; q + 1 < rows; q += 2) {
        const double2 v = *reinterpret_cast<const double2*>(fb + q);
        mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
        if (log_disp) dot = fma(nmlz[q], v.x, fma(nmlz[q + 1], v.y, dot));
      }
    }
    for (; q < rows; ++q) {
      const double v = fb[q];
      mx = fmax(mx, fabs(v));
      if (log_disp) dot = fma(nmlz[q], v, dot);
    }
  }
  colmax[(int64_t)box * R + r] = mx * fabs(alpha);
  nmdot[(int64_t)box * R + r] = dot;
}

// Density staging, pass 1, streaming form (no 2D log term): partial[g][r] =
// the largest |f[r][v]| alpha(v) over the kI8StreamChunk values of chunk g,
// alpha(v) = alpha_lut[pt_class[v >> cshift]] (class 255: a point of no
// list-1 source box, alpha 0; cshift 1 for interleaved complex values).  The
// per-RHS scale needs only these maxima, so the rows are read as flat
// streams -- each load instruction covers 512 consecutive bytes, with no
// per-box structure -- kI8StreamRows RHS rows per block.  Same products
// |f| alpha as the per-box form, so the same maxima.
constexpr int kI8StreamChunk = 2048;  // values per block: 256 threads x 8
constexpr int kI8StreamRows = 8;      // RHS rows per block

__global__ void __launch_bounds__(256)
    i8_colmax_stream_kernel(int64_t n_values, int cshift, int R, const double* __restrict__ f,
                            int64_t ldf, const uint8_t* __restrict__ pt_class,
                            const double* __restrict__ alpha_lut, double* __restrict__ partial) {
  __shared__ double red[kI8StreamRows][4];
  constexpr int PER = kI8StreamChunk / 256;
  const int64_t v0 = (int64_t)blockIdx.x * kI8StreamChunk + threadIdx.x;
  const int r0 = blockIdx.y * kI8StreamRows;
  const int nr = R - r0 < kI8StreamRows ? R - r0 : kI8StreamRows;
  double a[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t v = v0 + 256 * k;
    a[k] = v < n_values ? alpha_lut[pt_class[v >> cshift]] : 0.0;
  }
  double mx[kI8StreamRows];
#pragma unroll
  for (int rr = 0; rr < kI8StreamRows; ++rr) {
    mx[rr] = 0.0;
    if (rr < nr) {
      const double* row = f + (int64_t)(r0 + rr) * ldf;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int64_t v = v0 + 256 * k;
        if (v < n_values) mx[rr] = fmax(mx[rr], fabs(row[v]) * a[k]);
      }
    }
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int rr = 0; rr < kI8StreamRows; ++rr) {
    double m = mx[rr];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (lane == 0) red[rr][wave] = m;
  }
  __syncthreads();
  if ((int)threadIdx.x < nr) {
    const int rr = threadIdx.x;
    partial[(int64_t)blockIdx.x * R + r0 + rr] =
        fmax(fmax(red[rr][0], red[rr][1]), fmax(red[rr][2], red[rr][3]));
  }
}

// Two-stage maxima over many boxes: block (g, r) folds column r of boxes g,
// g + G, g + 2G, ... into colmax[g][r] (in place: no other block touches
// those entries); then i8_sigma_kernel over the first G boxes.
__global__ void __launch_bounds__(256) i8_colmax_fold_kernel(int nboxes, int R,
                                                             double* __restrict__ colmax) {
  __shared__ double red[256];
  const int G = gridDim.x, g = blockIdx.x, r = blockIdx.y;
  double mx = 0.0;
  for (int64_t b = g + (int64_t)G * threadIdx.x; b < nboxes; b += (int64_t)G * blockDim.x)
    mx = fmax(mx, colmax[b * R + r]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) colmax[(int64_t)g * R + r] = red[0];
}

// sigma[r] = max over boxes of colmax[box][r]; one block per RHS.
__global__ void __launch_bounds__(256) i8_sigma_kernel(int nboxes, int R,
                                                       const double* __restrict__ colmax,
                                                       double* __restrict__ sigma) {
  __shared__ double red[256];
  const int r = blockIdx.x;
  double mx = 0.0;
  for (int b = threadIdx.x; b < nboxes; b += blockDim.x) mx = fmax(mx, colmax[(int64_t)b * R + r]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) sigma[r] = red[0];
}

// Density staging, pass 2: digits of f alpha / sigma into
// X8[box][kc][j][r][32 bytes] (K chunk kc, slice j), zero past the rows.
// One block per box.  The box's K -> particle offsets (mode order) go to
// LDS once; rounds of RHS columns are then read into LDS in place (a box's
// particles are contiguous; row stride rows + 1 doubles spreads the rows
// over the banks), and a thread turns 16 consecutive K of one column into S
// 16-byte digit words, consecutive threads writing consecutive 16-byte
// pieces.  Boxes that are no source (alpha 0) are skipped: no step reads
// their digits.
constexpr int kI8ColsLds = 4096;  // doubles at most: offsets + one round of columns

// Columns per round of i8_split_cols_kernel: as many as kI8ColsLds holds, a
// multiple of 4 where possible, so a round's digit words fill whole 128-byte
// lines (4 columns x 32 bytes per K chunk); the block's LDS is
// (rows + 1) / 2 + RR (rows + 1) doubles, its size 64 to 256 threads by the
// round's work (small boxes take small blocks).
inline int i8_cols_round(int rows, int R) {
  const int cap = (kI8ColsLds - (rows + 1) / 2) / (rows + 1);  // >= 1: checked by the plan
  const int c4 = cap >= 4 ? cap / 4 * 4 : cap;
  return c4 < R ? c4 : R;
}

template <int S>
__global__ void __launch_bounds__(256)
    i8_split_cols_kernel(int n, int cplx, int KP, int nboxes, int R, int RR,
                         const double* __restrict__ f, int64_t ldf,
                         const int32_t* __restrict__ mode_point,
                         const int32_t* __restrict__ box_starts,
                         const double* __restrict__ box_alpha, const double* __restrict__ sigma,
                         int8_t* __restrict__ X8) {
  extern __shared__ double sf[];
  const int box = blockIdx.x;
  const double alpha = box_alpha[box];
  if (alpha == 0.0) return;
  const int rows = cplx ? 2 * n : n;
  const int nkc = KP / 32;
  const int LS = rows + 1;
  int* soff = reinterpret_cast<int*>(sf);  // rows ints
  double* sv = sf + (rows + 1) / 2;
  const int64_t bs = box_starts[box];
  const int64_t fo = cplx ? 2 * bs : bs;
  const int32_t* mp = mode_point + (int64_t)box * n;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int k = threadIdx.x; k < rows; k += blockDim.x) {
    const int pl = (int)(mp[k % n] - bs);
    soff[k] = cplx ? 2 * pl + k / n : pl;
  }
  for (int r0 = 0; r0 < R; r0 += RR) {
    const int nr = R - r0 < RR ? R - r0 : RR;
    for (int rr = wave; rr < nr; rr += blockDim.x / 64) {
      const double* src = f + (int64_t)(r0 + rr) * ldf + fo;
      double* dst = sv + rr * LS;
#pragma unroll 4
      for (int q = lane; q < rows; q += 64) dst[q] = src[q];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nr * nkc * 2; t += blockDim.x) {
      const int h = t % 2, rr = (t / 2) % nr, kc = t / (2 * nr);
      const int r = r0 + rr;
      const double sg = sigma[r];
      const double scale = sg > 0.0 ? alpha / sg : 0.0;
      const double* sr = sv + rr * LS;
      int8_t dg[16][S];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int k = 32 * kc + 16 * h + u;
        const double v = k < rows ? sr[soff[k]] : 0.0;
        i8_digits<S>(fmin(1.0, fmax(-1.0, v * scale)), dg[u]);
      }
#pragma unroll
      for (int j = 0; j < S; ++j) {
        uint32_t w[4];
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4)
          w[q4] = (uint32_t)(uint8_t)dg[4 * q4][j] | ((uint32_t)(uint8_t)dg[4 * q4 + 1][j] << 8) |
                  ((uint32_t)(uint8_t)dg[4 * q4 + 2][j] << 16) |
                  ((uint32_t)(uint8_t)dg[4 * q4 + 3][j] << 24);
        *reinterpret_cast<uint4*>(X8 + ((((int64_t)box * nkc + kc) * S + j) * R + r) * 32 + 16 * h) =
            make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
    __syncthreads();
  }
}

// Stage 1 of the error estimates over many tiles: block (g, r) folds the
// partials of RHS r (columns b Rc + r, b < B) of tiles g, g + G, ... into
// the entry of tile g, column r (in place: no other block touches those
// entries); i8_rhs_error_kernel then reads tiles < G with B = 1.
__global__ void __launch_bounds__(256) i8_est_fold_kernel(int n_tiles, int B, int Rc,
                                                          double* __restrict__ part) {
  __shared__ double red[4][256];
  const int G = gridDim.x, g = blockIdx.x, r = blockIdx.y;
  double em = 0.0, e2 = 0.0, ym = 0.0, y2 = 0.0;
  const int64_t per = ((int64_t)n_tiles - g + G - 1) / G;  // tiles g, g + G, ...
  for (int64_t i = threadIdx.x; i < per * B; i += blockDim.x) {
    const int64_t tile = g + (i / B) * G;
    const int c = (int)(i % B) * Rc + r;
    const double* q = part + (tile * 64 + c) * 4;
    em = fmax(em, q[0]);
    e2 += q[1];
    ym = fmax(ym, q[2]);
    y2 += q[3];
  }
  red[0][threadIdx.x] = em;
  red[1][threadIdx.x] = e2;
  red[2][threadIdx.x] = ym;
  red[3][threadIdx.x] = y2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + o]);
      red[3][threadIdx.x] += red[3][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* q = part + ((int64_t)g * 64 + r) * 4;
    q[0] = red[0][0];
    q[1] = red[1][0];
    q[2] = red[2][0];
    q[3] = red[3][0];
  }
}

// Per-RHS relative error estimates from the per-(tile, column) partials:
// out[r][0] = max est / max |y| (Linf), out[r][1] = sqrt(sum est^2 / sum y^2)
// (L2), over the plan's targets.  One block per RHS.
__global__ void i8_rhs_error_kernel(int n_tiles, int B, int Rc, int r0,
                                    const double* __restrict__ part, double* __restrict__ out) {
  const int r = blockIdx.x;  // RHS within this launch group
  __shared__ double red[4][256];
  double em = 0.0, e2 = 0.0, ym = 0.0, y2 = 0.0;
  for (int64_t i = threadIdx.x; i < (int64_t)n_tiles * B; i += blockDim.x) {
    const int64_t tile = i / B;
    const int c = (int)(i % B) * Rc + r;
    const double* q = part + (tile * 64 + c) * 4;
    em = fmax(em, q[0]);
    e2 += q[1];
    ym = fmax(ym, q[2]);
    y2 += q[3];
  }
  red[0][threadIdx.x] = em;
  red[1][threadIdx.x] = e2;
  red[2][threadIdx.x] = ym;
  red[3][threadIdx.x] = y2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + o]);
      red[3][threadIdx.x] += red[3][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * (r0 + r)] = red[2][0] > 0.0 ? red[0][0] / red[2][0] : 0.0;
    out[2 * (r0 + r) + 1] = red[3][0] > 0.0 ? sqrt(red[1][0] / red[3][0]) : 0.0;
  }
}

// Tile kernel.  Workgroup = MP/32 consumer waves (wave w owns block rows
// 32w .. 32w + 31 and all 64 columns of the tile) + 1 producer wave.
//
// The producer streams each (step, K chunk) of the tile -- the 32-column
// chunk of the S digit slices of the step's source box densities -- into a
// ring of NS LDS slots, NS - 1 chunks ahead of the consumers, together with
// the step's block id; one workgroup barrier per chunk separates writes of
// a slot from its reads.  Consumers take the table's digit fragments for
// their rows straight from global memory (fragment order, 1 KiB per wave
// load), refilled for the next chunk right after their last use, and the
// density fragments from the ring; products (i, j), i + j < S, accumulate
// into set i + j for the whole tile.  Ring slot layout:
// [S][2 K halves][64 columns][16 bytes], so a wave's fragment read of 32
// columns touches 512 consecutive bytes.
inline size_t i8_lds_bytes(int S, int NS) {
  return (size_t)NS * S * 2048 + 64 * 8 + (size_t)NS * 4;
}

template <int S>
__global__ void __launch_bounds__(512) split_i8_kernel(const TileParams p) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int SLOT = S * 2048;
  const int NS = p.nslot;
  unsigned char* sR = smem_raw;                                            // [NS][SLOT]
  double* sDsp = reinterpret_cast<double*>(smem_raw + (size_t)NS * SLOT);  // [64]
  int* sBlk = reinterpret_cast<int*>(sDsp + 64);                           // [NS]

  const int tile = xcd_tile(blockIdx.x, p.n_tiles);
  if (tile >= p.n_tiles) return;  // uniform over the workgroup
  const int nrt = p.MP / 32;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  const int ncols = p.B * p.Rc;
  const int nkc = p.KP / 32;
  const int st0 = p.tile_step_start[tile], st1 = p.tile_step_start[tile + 1];
  const int total = (st1 - st0) * nkc;
  const int D = NS - 1;  // chunks the producer runs ahead
  const int64_t jstride = (int64_t)p.R * 32;  // bytes between the slices of a chunk

  if (wave == nrt) {
    // ------------------------------------------------------------ producer
    // Chunk c's global loads are issued one chunk before its LDS stores, so
    // they have a whole consumer iteration to arrive; the step descriptors
    // (source boxes of this lane's two 16-byte pieces, block id) are read
    // one step ahead.
    double dsum = 0.0;  // lane c: sum over the tile's steps of the column's 2D log term
    const int nst = st1 - st0;
    int col[2], half[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      col[m] = (lane + 64 * m) / 2;  // 16-byte piece lane + 64 m: column, K half
      half[m] = (lane + 64 * m) % 2;
    }
    auto load_desc = [&](int k, int (&sb)[2], int& blk) {
      blk = 0;
#pragma unroll
      for (int m = 0; m < 2; ++m) sb[m] = -1;
      if (k >= nst) return;
      const int st = st0 + k;
      blk = p.step_block[st];
#pragma unroll
      for (int m = 0; m < 2; ++m)
        if (col[m] < ncols) sb[m] = p.step_src[st * p.B + col[m] / p.Rc];
    };
    // every slot holds a valid block id from the start (consumers prefetch
    // table fragments for the chunk after the last one from a stale slot)
    if (lane < NS) sBlk[lane] = 0;
    int sbC[2], sbN[2], blkC = 0, blkN = 0;
    load_desc(0, sbN, blkN);
    // NPB chunks' loads in flight: chunk c lives in register buffer
    // (c - D) % NPB and is loaded NPB iterations before its LDS stores; the
    // loads are unconditional (a column without a source reads offset 0 and
    // stores zeros), so no branch separates a load from its wait
    constexpr int NPB = kI8ProducerDepth;
    i32x4 buf[NPB][2][S];
    int buf_blk[NPB];
    bool live[NPB][2];
    auto issue = [&](int c, auto pc) {
      constexpr int P = decltype(pc)::value;
      const int k = c / nkc, kc = c % nkc;
      if (kc == 0) {  // a new step: its descriptors, then prefetch the next step's
#pragma unroll
        for (int m = 0; m < 2; ++m) sbC[m] = sbN[m];
        blkC = blkN;
        load_desc(k + 1, sbN, blkN);
        if (p.log_disp && lane < ncols) {
          const int r = lane % p.Rc;
          const int sb = p.step_src[(st0 + k) * p.B + lane / p.Rc];
          if (sb >= 0) dsum += p.box_disp[sb] * p.nmdot[(int64_t)sb * p.R + p.r0 + r];
        }
      }
      buf_blk[P] = blkC;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        live[P][m] = sbC[m] >= 0;
        const int64_t base =
            live[P][m]
                ? ((((int64_t)sbC[m] * nkc + kc) * S) * p.R + p.r0 + col[m] % p.Rc) * 32 + 16 * half[m]
                : 0;
#pragma unroll
        for (int j = 0; j < S; ++j)
          buf[P][m][j] = *reinterpret_cast<const i32x4*>(p.X8 + base + (live[P][m] ? j * jstride : 0));
      }
    };
    auto commit = [&](int c, auto pc) {
      constexpr int P = decltype(pc)::value;
      const int slot = c % NS;
      if (lane == 0) sBlk[slot] = buf_blk[P];
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int j = 0; j < S; ++j)
          *reinterpret_cast<i32x4*>(sR + (size_t)slot * SLOT + j * 2048 + half[m] * 1024 +
                                    col[m] * 16) = live[P][m] ? buf[P][m][j] : i32x4{0, 0, 0, 0};
    };
    using Z = std::integral_constant<int, 0>;
    const int pre = total < D ? total : D;
    for (int c = 0; c < pre; ++c) {
      issue(c, Z{});
      commit(c, Z{});
    }
    static_for<0, NPB>([&](auto qc) {  // in flight across the first barrier
      if (D + decltype(qc)::value < total) issue(D + decltype(qc)::value, qc);
    });
    __syncthreads();
    // iteration it stores chunk it + D (slot (it + D) % NS was last read in
    // iteration it - 1) and issues the loads of chunk it + D + NPB
    auto step = [&](int it, auto pc) {
      const int c = it + D;
      if (c < total) {
        commit(c, pc);
        if (c + NPB < total) issue(c + NPB, pc);
      }
      __syncthreads();
    };
    int it = 0;
    for (; it + NPB - 1 < total; it += NPB)
      static_for<0, NPB>([&](auto qc) { step(it + decltype(qc)::value, qc); });
    static_for<0, NPB - 1>([&](auto qc) {
      if (it + decltype(qc)::value < total) step(it + decltype(qc)::value, qc);
    });
    sDsp[lane] = dsum;
    __syncthreads();  // end of the step loop
  } else {
    // ------------------------------------------------------------ consumers
    i32x16 acc[S][2];
#pragma unroll
    for (int d = 0; d < S; ++d)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[d][ct][v] = 0;
    const int8_t* arow = p.A8 + (int64_t)wave * nkc * S * 1024 + lane * 16;
    const int64_t blk_stride = (int64_t)nrt * nkc * S * 1024;
    auto afrag = [&](int blk, int kc, int i) {
      return *reinterpret_cast<const i32x4*>(arow + blk * blk_stride + ((int64_t)kc * S + i) * 1024);
    };
    __syncthreads();  // the first D chunks and their block ids are in the ring
    // Table fragments: slices i < NB of chunk it + NA - 1 are loaded at the
    // top of chunk it into one of NA register sets used in rotation (the
    // loop runs NA chunks per trip, so no register copies sit between a load
    // and its use; S <= 5 loads two chunks ahead, which covers a miss in L2);
    // the others (S = 6 only, for registers) are refilled in place right
    // after their last use (slice i meets density slices 0..S-1-i).  Density
    // slice 0 of a chunk takes the table slices in the order they were
    // loaded: the rotating ones, then S-1 down to NB.  The loads are
    // unconditional (past the last chunk they reload an older, valid block),
    // so no branch separates a load from its wait.
    constexpr int NB = S <= 5 ? S : 3;
    constexpr int NSH = S - NB > 0 ? S - NB : 1;
    // NA register sets of the NB ping-pong slices: the loads for chunk it + NA - 1
    // are issued at the top of chunk it (two chunks ahead where registers allow)
    constexpr int NA = S <= 5 ? 3 : 2;
    i32x4 app[NA][NB], ash[NSH];
    if (total > 0) {
      const int blk = sBlk[0];
#pragma unroll
      for (int i = 0; i < S; ++i) {
        if (i < NB) app[0][i] = afrag(blk, 0, i);
        else ash[i - NB] = afrag(blk, 0, i);
      }
    }
    if (NA == 3 && total > 1) {
      const int blk = sBlk[1 % NS];
#pragma unroll
      for (int i = 0; i < NB; ++i) app[NA - 2][i] = afrag(blk, 1 % nkc, i);
    }
    const unsigned char* rsl = sR + (lane / 32) * 1024 + (lane % 32) * 16;
    // Density fragments (S <= 5; S = 6 has no registers for it): slice j + 1
    // (after the last slice, slice 0 of the next chunk) is read from the
    // ring before slice j's MFMAs are issued, so its LDS latency runs under
    // them.  The next chunk was stored before the previous barrier (the
    // producer runs NS - 1 >= 2 chunks ahead); after the last chunk the read
    // hits a stale slot and is not used.
    constexpr bool AHEAD_B = S <= 5;
    i32x4 bc0, bc1;
    if (AHEAD_B && total > 0) {
      bc0 = *reinterpret_cast<const i32x4*>(rsl);
      bc1 = *reinterpret_cast<const i32x4*>(rsl + 512);
    }
    auto chunk = [&](int it, auto pc) {
      constexpr int P = decltype(pc)::value;  // register set of this chunk (it % NA)
      // the block ids of the next NA - 1 chunks were written at least one
      // barrier ago (after the last chunk a slot holds an older, still valid,
      // block id)
      const int nblk = sBlk[(it + 1) % NS];
      const int nkc_ = (it + 1) % nkc;
      const int ablk = NA == 2 ? nblk : sBlk[(it + NA - 1) % NS];
      const int akc = NA == 2 ? nkc_ : (it + NA - 1) % nkc;
#pragma unroll
      for (int i = 0; i < NB; ++i) app[(P + NA - 1) % NA][i] = afrag(ablk, akc, i);
      // keep these loads at the top: the scheduler would otherwise sink them
      // below the MFMAs, leaving their latency to the next chunk's start
      __builtin_amdgcn_sched_barrier(0);
      const unsigned char* rs = rsl + (size_t)(it % NS) * SLOT;
      const unsigned char* rn = rsl + (size_t)((it + 1) % NS) * SLOT;
      static_for<0, S>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        i32x4 b0, b1;
        if constexpr (AHEAD_B) {
          b0 = bc0;
          b1 = bc1;
          const unsigned char* src = j + 1 < S ? rs + (j + 1) * 2048 : rn;
          bc0 = *reinterpret_cast<const i32x4*>(src);
          bc1 = *reinterpret_cast<const i32x4*>(src + 512);
          __builtin_amdgcn_sched_barrier(0);  // the reads go out before the MFMAs
        } else {
          b0 = *reinterpret_cast<const i32x4*>(rs + j * 2048);
          b1 = *reinterpret_cast<const i32x4*>(rs + j * 2048 + 512);
        }
        static_for<0, S - j>([&](auto ic) {
          constexpr int u = decltype(ic)::value;
          constexpr int i = (j == 0 && u >= NB) ? S - 1 - (u - NB) : u;
          const i32x4& ai = i < NB ? app[P][i < NB ? i : 0] : ash[i < NB ? 0 : i - NB];
          acc[i + j][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai, b0, acc[i + j][0], 0, 0, 0);
          acc[i + j][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai, b1, acc[i + j][1], 0, 0, 0);
        });
        // table slice S-1-j had its last partner: refill it for the next chunk
        if constexpr (S - 1 - j >= NB) ash[S - 1 - j - NB] = afrag(nblk, nkc_, S - 1 - j);
      });
      __syncthreads();  // ring slot it % NS may be refilled
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, NA == 3 ? 2 : 0>;
    int it = 0;
    if constexpr (NA == 3) {
      for (; it + 2 < total; it += 3) {
        chunk(it, I0{});
        chunk(it + 1, I1{});
        chunk(it + 2, I2{});
      }
      if (it < total) chunk(it, I0{});
      if (it + 1 < total) chunk(it + 1, I1{});
    } else {
      for (; it + 1 < total; it += 2) {
        chunk(it, I0{});
        chunk(it + 1, I1{});
      }
      if (it < total) chunk(it, I0{});
    }
    __syncthreads();  // end of the step loop (the producer's log-term sums)

    // epilogue: y = rho sigma sum_d acc_d W_d + sum disp, per output; the
    // error estimate 4 |acc_{S-1}| W_{S-1} / 254 rho sigma; per-column
    // partials (max est, sum est^2, max |y|, sum y^2) for the RHS estimates
    const int g = tile % p.n_rg;
    const double* rho = p.rho + (int64_t)g * p.MP;
    double sg[2], dsp[2];
    int tbx[2], rr[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int c = 32 * ct + lane % 32;
      const bool vc = c < ncols;
      const int b = vc ? c / p.Rc : 0, r = vc ? c % p.Rc : 0;
      tbx[ct] = vc ? p.tile_boxes[tile * p.B + b] : -1;
      rr[ct] = r;
      sg[ct] = vc ? p.sigma[p.r0 + r] : 0.0;
      dsp[ct] = sDsp[c];
    }
    double em[2] = {0.0, 0.0}, e2[2] = {0.0, 0.0}, ym[2] = {0.0, 0.0}, y2[2] = {0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int row = 32 * wave + 8 * (v / 4) + 4 * (lane / 32) + v % 4;
      const int grow = g * p.rg_size + row;
      const bool vr = row < p.rg_size && grow < p.n;
      const double rh = vr ? rho[row] : 0.0;
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        // sum_d acc_d 254^-d in FP64 (Horner from the smallest diagonal),
        // then the common factor 1/127^2
        double val = (double)acc[S - 1][ct][v];
#pragma unroll
        for (int d = S - 2; d >= 0; --d) val = fma(val, 1.0 / 254.0, (double)acc[d][ct][v]);
        val *= i8_weight(0);
        const double sc = rh * sg[ct];
        const double y = fma(val, sc, dsp[ct]);
        const double est =
            4.0 / 254.0 * fabs((double)acc[S - 1][ct][v]) * i8_weight(S - 1) * fabs(sc);
        if (vr && tbx[ct] >= 0) {
          p.y[(int64_t)(p.r0 + rr[ct]) * p.ldy + out_index(p, tbx[ct], grow)] = y;
          em[ct] = fmax(em[ct], est);
          e2[ct] = fma(est, est, e2[ct]);
          ym[ct] = fmax(ym[ct], fabs(y));
          y2[ct] = fma(y, y, y2[ct]);
        }
      }
    }
    if (p.est_part) {
      double* scr = reinterpret_cast<double*>(smem_raw);  // the ring is free now: [nrt][64][4]
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        em[ct] = fmax(em[ct], __shfl_xor(em[ct], 32));
        e2[ct] += __shfl_xor(e2[ct], 32);
        ym[ct] = fmax(ym[ct], __shfl_xor(ym[ct], 32));
        y2[ct] += __shfl_xor(y2[ct], 32);
        if (lane < 32) {
          double* q = scr + ((int64_t)wave * 64 + 32 * ct + lane) * 4;
          q[0] = em[ct];
          q[1] = e2[ct];
          q[2] = ym[ct];
          q[3] = y2[ct];
        }
      }
    }
  }
  if (p.est_part) {
    __syncthreads();
    if (threadIdx.x < 64) {
      const double* scr = reinterpret_cast<const double*>(smem_raw);
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      for (int w = 0; w < nrt; ++w) {
        const double* q = scr + ((int64_t)w * 64 + threadIdx.x) * 4;
        a0 = fmax(a0, q[0]);
        a1 += q[1];
        a2 = fmax(a2, q[2]);
        a3 += q[3];
      }
      double* o = p.est_part + ((int64_t)tile * 64 + threadIdx.x) * 4;
      o[0] = a0;
      o[1] = a1;
      o[2] = a2;
      o[3] = a3;
    }
  }
}

