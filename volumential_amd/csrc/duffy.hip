// FP64 Duffy quadrature of canonical near-field table entries.
//
// Entry e = (case, s, t): target point x (case geometry), decomposition
// vertex v = clamp(x, 0, H), source mode s with per-axis indices m_k.
//   T_e = sum_nodes  G(x - y) * jac * prod_k len_k * prod_k l_{m_k}(y_k)
// with len_k = v_k (sign < 0) or H - v_k (sign > 0), y = v + sign*len*u,
// a node active only if every len_k > 0 and |x - y|^2 > 0, and the
// barycentric Lagrange basis l evaluated with the node-hit rule at
// tolerance 8 eps H.  This is the computation of the fused loopy kernel
// `duffy_invariant_fused_table` (volumential/nearfield_potential_table.py:
// 2098-2201), launched over every canonical entry of every case at once.
//
// GPU mapping (per-entry form): one lane per entry; the node loop is
// sequential and all lanes of a wavefront read the same node (uniform,
// scalar-cached loads).  A sign block whose sub-box is empty for the entry
// (some len_k == 0) is skipped as a whole: every node in it is inactive.
// The segmented form (duffy_seg_kernel, below) shares each node's geometry
// and kernel value among the entries of one target point.
#include <cmath>
#include <string>

#include "internal.h"

namespace vl1 {

struct DuffyParams {
  int dim, q, kind, deriv_axis;
  double k;
  int64_t n_entries, n_nodes, nodes_per_sign;
  double H;
  const double* xt;
  const double* vt;
  const int32_t* mode;
  const double* u;
  const double* sg;
  const double* jac;
  const double* xi;
  const double* bw;
  double* out;
  const int64_t* chunk_start;  // segmented form: [n_chunks + 1], entries sorted by target point
  int64_t n_chunks;
};

// Lagrange basis l_m(y) on the q interpolation nodes, product form:
//   l_m(y) = c_m * prod_{j != m} (y - xi_j),  c_m = 1 / prod_{j != m} (xi_m - xi_j)
// (no divisions per node), with the reference's node-hit rule: if y lies
// within tol of some node, l_m(y) is 1 at that node and 0 elsewhere
// (lagrange.py:45-92, nearfield_potential_table.py:2172-2178).
template <int Q>
__device__ inline double lagrange_basis(double y, int m, int q, double cm, const double* xi,
                                        double tol) {
  const int qq = Q > 0 ? Q : q;
  bool hit_any = false, hit_m = false;
  double pr = cm;
#pragma unroll
  for (int j = 0; j < (Q > 0 ? Q : 16); ++j) {
    if (Q == 0 && j >= qq) break;
    const double d = y - xi[j];
    const bool h = fabs(d) < tol;
    hit_any = hit_any || h;
    hit_m = hit_m || (h && j == m);
    pr *= (j == m) ? 1.0 : d;
  }
  return hit_any ? (hit_m ? 1.0 : 0.0) : pr;
}

// Kernel values, sumpy conventions (expression times global scaling):
// Laplace 2D -log(r)/(2 pi), 3D 1/(4 pi r); Helmholtz 2D (i/4) H0^(1)(k r)
// = (-Y0(kr) + i J0(kr)) / 4, 3D e^{ikr}/(4 pi r).  Axis derivatives of
// Laplace (da = component of d = target - source on the derivative axis):
// target dG/dd_a = 2D -da/(2 pi r^2), 3D -da/(4 pi r^3); source = minus.
// Axis derivatives of Helmholtz: G'(r) da / r with G' = 2D -(ik/4) H1^(1)(kr)
// = (k/4)(Y1(kr) - i J1(kr)), 3D e^{ikr}(ikr - 1)/(4 pi r^2).  Radial terms
// of the Helmholtz split: r^p and r^p log r, no scaling constant (the
// exponent p travels in the k slot).
__host__ __device__ constexpr bool kind_is_complex(int kind) {
  return kind == VL1_KERNEL_HELMHOLTZ || kind == VL1_KERNEL_HELMHOLTZ_TARGET_DERIV ||
         kind == VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV;
}
__host__ __device__ constexpr bool kind_is_power_deriv(int kind) {
  return kind == VL1_KERNEL_POWER_TARGET_DERIV || kind == VL1_KERNEL_POWER_SOURCE_DERIV ||
         kind == VL1_KERNEL_POWER_LOG_TARGET_DERIV || kind == VL1_KERNEL_POWER_LOG_SOURCE_DERIV;
}
__host__ __device__ constexpr bool kind_is_deriv(int kind) {
  return kind == VL1_KERNEL_LAPLACE_TARGET_DERIV || kind == VL1_KERNEL_LAPLACE_SOURCE_DERIV ||
         kind == VL1_KERNEL_HELMHOLTZ_TARGET_DERIV || kind == VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV ||
         kind_is_power_deriv(kind);
}
template <int DIM, int KIND>
__device__ inline void kernel_value(double r2, double da, double k, double& re, double& im) {
  if (KIND == VL1_KERNEL_LAPLACE_TARGET_DERIV || KIND == VL1_KERNEL_LAPLACE_SOURCE_DERIV) {
    constexpr double sgn = KIND == VL1_KERNEL_LAPLACE_TARGET_DERIV ? -1.0 : 1.0;
    if (DIM == 2) {
      re = sgn * 0.5 / M_PI * da / r2;
    } else {
      const double ir = rsqrt(r2);
      re = sgn * 0.25 / M_PI * da * ir * ir * ir;
    }
    im = 0.0;
  } else if (KIND == VL1_KERNEL_HELMHOLTZ_TARGET_DERIV || KIND == VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV) {
    constexpr double sgn = KIND == VL1_KERNEL_HELMHOLTZ_TARGET_DERIV ? 1.0 : -1.0;
    const double ir = rsqrt(r2);
    const double r = r2 * ir;
    const double kr = k * r;
    if (DIM == 2) {
      const double f = sgn * 0.25 * k * da * ir;
      re = f * y1(kr);
      im = -f * j1(kr);
    } else {
      double sn, cs;
      sincos(kr, &sn, &cs);
      const double f = sgn * 0.25 / M_PI * da * ir * ir * ir;
      re = f * (-cs - kr * sn);
      im = f * (kr * cs - sn);
    }
  } else if (KIND == VL1_KERNEL_CONSTANT) {
    re = 1.0;
    im = 0.0;
  } else if (kind_is_power_deriv(KIND)) {
    // d_a |d|^(p-2) times p, or times (p log|d| + 1) with the log (k carries p)
    constexpr bool tgt =
        KIND == VL1_KERNEL_POWER_TARGET_DERIV || KIND == VL1_KERNEL_POWER_LOG_TARGET_DERIV;
    constexpr bool lg =
        KIND == VL1_KERNEL_POWER_LOG_TARGET_DERIV || KIND == VL1_KERNEL_POWER_LOG_SOURCE_DERIV;
    const double base = da * pow(r2, 0.5 * k - 1.0);
    re = (tgt ? 1.0 : -1.0) * base * (lg ? fma(0.5 * k, log(r2), 1.0) : k);
    im = 0.0;
  } else if (KIND == VL1_KERNEL_POWER || KIND == VL1_KERNEL_POWER_LOG) {
    // |d|^p (k carries p for these kinds), times log|d| for POWER_LOG
    re = pow(r2, 0.5 * k);
    if (KIND == VL1_KERNEL_POWER_LOG) re *= 0.5 * log(r2);
    im = 0.0;
  } else if (KIND == VL1_KERNEL_LAPLACE) {
    if (DIM == 2) {
      re = -0.25 / M_PI * log(r2);
    } else {
      re = 0.25 / M_PI * rsqrt(r2);
    }
    im = 0.0;
  } else if (DIM == 2) {
    const double kr = k * sqrt(r2);
    re = -0.25 * y0(kr);
    im = 0.25 * j0(kr);
  } else {
    const double ir = rsqrt(r2);
    double sn, cs;
    sincos(k * (r2 * ir), &sn, &cs);
    re = 0.25 / M_PI * cs * ir;
    im = 0.25 / M_PI * sn * ir;
  }
}

// One lane per entry; every lane of a wave walks the same node sequence, so
// node data are wave-uniform loads.
template <int DIM, int KIND, int Q>
__global__ void __launch_bounds__(256) duffy_kernel(const DuffyParams p) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p.n_entries) return;
  constexpr int kMaxQ = 16;
  const int q = Q > 0 ? Q : p.q;
  const double* __restrict__ nu = p.u;
  const double* __restrict__ nsg = p.sg;
  const double* __restrict__ njac = p.jac;
  double xi[kMaxQ];
#pragma unroll
  for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j)
    if (Q > 0 || j < q) xi[j] = p.xi[j];
  double x[DIM], v[DIM], cm[DIM];
  int m[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    x[a] = p.xt[a * p.n_entries + e];
    v[a] = p.vt[a * p.n_entries + e];
    m[a] = p.mode[a * p.n_entries + e];
    double den = 1.0;
#pragma unroll
    for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j)
      if (Q > 0 || j < q) den *= (j == m[a]) ? 1.0 : (xi[m[a]] - xi[j]);
    cm[a] = 1.0 / den;
  }
  const double tol = 8.0 * 2.220446049250313e-16 * p.H;
  double acc_re = 0.0, acc_im = 0.0;
  const int64_t nsign = p.n_nodes / p.nodes_per_sign;
  for (int64_t b = 0; b < nsign; ++b) {
    const int64_t base = b * p.nodes_per_sign;
    double len[DIM];
    double sgn[DIM];
    bool pos = true;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      sgn[a] = nsg[a * p.n_nodes + base];
      len[a] = sgn[a] < 0.0 ? v[a] : p.H - v[a];
      pos = pos && (len[a] > 0.0);
    }
    if (!pos) continue;  // empty sub-box: every node of the block inactive
    double sl[DIM], lenprod = 1.0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      sl[a] = sgn[a] * len[a];
      lenprod *= len[a];
    }
    for (int64_t i = 0; i < p.nodes_per_sign; ++i) {
      const int64_t node = base + i;
      double r2 = 0.0, da = 0.0;
      double yk[DIM];
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        yk[a] = v[a] + sl[a] * nu[a * p.n_nodes + node];
        const double d = x[a] - yk[a];
        r2 += d * d;
        if (a == p.deriv_axis) da = d;
      }
      if (!(r2 > 0.0)) continue;
      double basis = njac[node] * lenprod;
#pragma unroll
      for (int a = 0; a < DIM; ++a) basis *= lagrange_basis<Q>(yk[a], m[a], q, cm[a], xi, tol);
      double kr, ki;
      kernel_value<DIM, KIND>(r2, da, p.k, kr, ki);
      acc_re = fma(kr, basis, acc_re);
      if (kind_is_complex(KIND)) acc_im = fma(ki, basis, acc_im);
    }
  }
  if (kind_is_complex(KIND)) {
    p.out[2 * e] = acc_re;
    p.out[2 * e + 1] = acc_im;
  } else {
    p.out[e] = acc_re;
  }
}

// Segmented form.  The entries of one (case, target node) pair share the
// target point x, the decomposition vertex v and therefore every node y and
// kernel value G(x - y); they differ only in the source mode m, i.e. in the
// Lagrange factor prod_k l_{m_k}(y_k).  One wavefront takes a chunk of at
// most 64 such entries.  Per round each lane produces one node record in
// LDS -- K l_0..l_{q-1} of axis 0 (K = G(x - y) jac prod_k len_k folded in,
// complex for Helmholtz kinds) and the q basis values of each other axis --
// and then every lane accumulates its entry over the wave's 64 records:
// (K l_{m_0}(y_0)) * prod_{k>0} l_{m_k}(y_k), a few LDS reads and
// multiplies per (entry, node) instead of the whole geometry and kernel
// evaluation.  With fewer than 64 entries, R = 64 / ne' lanes share an
// entry (ne' = ne rounded up to a power of two) and split the records,
// then combine by shuffles.  Nodes, signs and the skip rule are those of
// duffy_kernel; only the summation order differs.
//
// LDS holds the records field-major, rec[field][record] with a row pitch of
// kSegPitch doubles (rows start 16 bytes apart modulo the 128-byte bank
// window), so a lane reads one field of two consecutive records with one
// 16-byte load.  Records past the end of a sign block are zero.
constexpr int kSegNodes = 64;
constexpr int kSegPitch = kSegNodes + 2;

template <int DIM, int KIND, int Q>
__global__ void __launch_bounds__(64) duffy_seg_kernel(const DuffyParams p) {
  constexpr int kMaxQ = 16;
  constexpr int VW = kind_is_complex(KIND) ? 2 : 1;
  const int q = Q > 0 ? Q : p.q;
  extern __shared__ double rec[];  // [VW * q + (DIM - 1) * q][kSegPitch]
  const int lane = threadIdx.x;
  const int64_t e0 = p.chunk_start[blockIdx.x];
  const int64_t ne64 = p.chunk_start[blockIdx.x + 1] - e0;
  // A chunk holds 1..64 entries (one lane per entry at least); any other
  // count breaks the lane layout below.  Such a chunk writes NaN to the
  // entries it names inside the result (the host's finiteness check then
  // rejects the table) and returns.
  if (ne64 < 1 || ne64 > kSegNodes || e0 < 0 || e0 + ne64 > p.n_entries) {
    if (e0 >= 0 && ne64 > 0)
      for (int64_t i = lane; i < ne64 && e0 + i < p.n_entries; i += kSegNodes) {
        p.out[VW * (e0 + i)] = __builtin_nan("");
        if (VW == 2) p.out[VW * (e0 + i) + 1] = __builtin_nan("");
      }
    return;
  }
  const int ne = (int)ne64;
  int nep = 1;
  while (nep < ne) nep *= 2;                 // ne' (<= 64)
  const int R = kSegNodes / nep;             // lanes per entry
  const int el = lane % nep, part = lane / nep;
  const bool active = el < ne;
  double x[DIM], v[DIM];
  int m[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    x[a] = p.xt[a * p.n_entries + e0];
    v[a] = p.vt[a * p.n_entries + e0];
    m[a] = active ? p.mode[a * p.n_entries + e0 + el] : 0;
  }
  double xi[kMaxQ], cm[kMaxQ];
#pragma unroll
  for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j)
    if (Q > 0 || j < q) xi[j] = p.xi[j];
#pragma unroll
  for (int i = 0; i < (Q > 0 ? Q : kMaxQ); ++i) {
    if (Q == 0 && i >= q) break;
    double den = 1.0;
#pragma unroll
    for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j)
      if (Q > 0 || j < q) den *= (j == i) ? 1.0 : (xi[i] - xi[j]);
    cm[i] = 1.0 / den;
  }
  const double tol = 8.0 * 2.220446049250313e-16 * p.H;
  double acc_re = 0.0, acc_im = 0.0;
  const int64_t nsign = p.n_nodes / p.nodes_per_sign;
  double* my = rec + lane;  // this lane's record: field f at my[f * kSegPitch]
  for (int64_t b = 0; b < nsign; ++b) {
    const int64_t base = b * p.nodes_per_sign;
    double len[DIM], sl[DIM], lenprod = 1.0;
    bool pos = true;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const double sg = p.sg[a * p.n_nodes + base];
      len[a] = sg < 0.0 ? v[a] : p.H - v[a];
      pos = pos && (len[a] > 0.0);
      sl[a] = sg * len[a];
      lenprod *= len[a];
    }
    if (!pos) continue;  // empty sub-box (uniform over the wave)
    for (int64_t r0 = 0; r0 < p.nodes_per_sign; r0 += kSegNodes) {
      // produce this lane's node record
      const int64_t i = r0 + lane;
      double kr = 0.0, ki = 0.0;
      double bas[DIM][kMaxQ];
      bool ok = false;
      if (i < p.nodes_per_sign) {
        const int64_t node = base + i;
        double r2 = 0.0, da = 0.0, yk[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          yk[a] = v[a] + sl[a] * p.u[a * p.n_nodes + node];
          const double d = x[a] - yk[a];
          r2 += d * d;
          if (a == p.deriv_axis) da = d;
        }
        if (r2 > 0.0) {
          ok = true;
          kernel_value<DIM, KIND>(r2, da, p.k, kr, ki);
          const double w = p.jac[node] * lenprod;
          kr *= w;
          ki *= w;
#pragma unroll
          for (int a = 0; a < DIM; ++a) {
            // l_i(y) = c_i prod_{j != i} (y - xi_j) by prefix / suffix
            // products, with the node-hit rule of lagrange_basis
            double d[kMaxQ], pre[kMaxQ];
            int hit = -1;
            double acc = 1.0;
#pragma unroll
            for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j) {
              if (Q == 0 && j >= q) break;
              d[j] = yk[a] - xi[j];
              if (hit < 0 && fabs(d[j]) < tol) hit = j;
              pre[j] = acc;
              acc *= d[j];
            }
            double suf = 1.0;
#pragma unroll
            for (int j = (Q > 0 ? Q : kMaxQ) - 1; j >= 0; --j) {
              if (Q == 0 && j >= q) continue;
              bas[a][j] = hit >= 0 ? (j == hit ? 1.0 : 0.0) : cm[j] * pre[j] * suf;
              suf *= d[j];
            }
          }
        }
      }
      __syncthreads();  // the previous round's records have been read
      // [K l(y_0) re][K l(y_0) im (complex)][l(y_1)]...[l(y_{DIM-1})]
#pragma unroll
      for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j) {
        if (Q == 0 && j >= q) break;
        my[j * kSegPitch] = ok ? kr * bas[0][j] : 0.0;
        if (VW == 2) my[(q + j) * kSegPitch] = ok ? ki * bas[0][j] : 0.0;
      }
#pragma unroll
      for (int a = 1; a < DIM; ++a)
#pragma unroll
        for (int j = 0; j < (Q > 0 ? Q : kMaxQ); ++j) {
          if (Q == 0 && j >= q) break;
          my[((VW + a - 1) * q + j) * kSegPitch] = ok ? bas[a][j] : 0.0;
        }
      __syncthreads();
      // consume: this lane's share of the wave's records for its entry
      const int nrec = (int)min<int64_t>(kSegNodes, p.nodes_per_sign - r0);
      if (active) {
        // record pairs (t, t + 1); the zero records past nrec add nothing
        const double* f0 = rec + (size_t)m[0] * kSegPitch;
        const double* f0i = rec + (size_t)(q + m[0]) * kSegPitch;
        const double* f1 = rec + (size_t)(VW * q + m[1]) * kSegPitch;
        const double* f2 = rec + (size_t)((VW + 1) * q + m[DIM - 1]) * kSegPitch;
        for (int t = 2 * part; t < nrec; t += 2 * R) {
          const double2 b1 = *reinterpret_cast<const double2*>(f1 + t);
          double2 bp = b1;
          if (DIM == 3) {
            const double2 b2 = *reinterpret_cast<const double2*>(f2 + t);
            bp.x *= b2.x;
            bp.y *= b2.y;
          }
          const double2 k0 = *reinterpret_cast<const double2*>(f0 + t);
          acc_re = fma(k0.x, bp.x, fma(k0.y, bp.y, acc_re));
          if (VW == 2) {
            const double2 k0i = *reinterpret_cast<const double2*>(f0i + t);
            acc_im = fma(k0i.x, bp.x, fma(k0i.y, bp.y, acc_im));
          }
        }
      }
    }
  }
  // combine the R lanes of each entry (lanes el + nep * part)
  for (int off = nep; off < kSegNodes; off *= 2) {
    acc_re += __shfl_xor(acc_re, off);
    if (VW == 2) acc_im += __shfl_xor(acc_im, off);
  }
  if (active && part == 0) {
    if (VW == 2) {
      p.out[2 * (e0 + el)] = acc_re;
      p.out[2 * (e0 + el) + 1] = acc_im;
    } else {
      p.out[e0 + el] = acc_re;
    }
  }
}

template <int DIM, int KIND, int Q>
static hipError_t launch_duffy_q(const DuffyParams& prm, hipStream_t s) {
  if (prm.chunk_start) {
    const int vw = kind_is_complex(KIND) ? 2 : 1;
    const size_t smem = sizeof(double) * kSegPitch * ((size_t)(vw + DIM - 1) * prm.q);
    hipLaunchKernelGGL((duffy_seg_kernel<DIM, KIND, Q>), dim3((unsigned)prm.n_chunks),
                       dim3(kSegNodes), smem, s, prm);
    return hipGetLastError();
  }
  const int block = 256;
  const int64_t grid = (prm.n_entries + block - 1) / block;
  hipLaunchKernelGGL((duffy_kernel<DIM, KIND, Q>), dim3((unsigned)grid), dim3(block), 0, s, prm);
  return hipGetLastError();
}

template <int DIM, int KIND>
static hipError_t launch_duffy(const DuffyParams& prm, hipStream_t s) {
  switch (prm.q) {
    case 1: return launch_duffy_q<DIM, KIND, 1>(prm, s);
    case 2: return launch_duffy_q<DIM, KIND, 2>(prm, s);
    case 3: return launch_duffy_q<DIM, KIND, 3>(prm, s);
    case 4: return launch_duffy_q<DIM, KIND, 4>(prm, s);
    case 5: return launch_duffy_q<DIM, KIND, 5>(prm, s);
    case 6: return launch_duffy_q<DIM, KIND, 6>(prm, s);
    case 8: return launch_duffy_q<DIM, KIND, 8>(prm, s);
    default: return launch_duffy_q<DIM, KIND, 0>(prm, s);
  }
}

}  // namespace vl1

using namespace vl1;

extern "C" int vl1_duffy_build(const vl1_duffy_args* a, void* stream) {
  clear_error();
  if (!a) return fail(VL1_E_ARG, "vl1_duffy_build: null argument");
  if (a->dim != 2 && a->dim != 3) return fail(VL1_E_UNSUPPORTED, "vl1_duffy_build: dim must be 2 or 3");
  if (a->quad_order < 1 || a->quad_order > 16) return fail(VL1_E_ARG, "vl1_duffy_build: quad_order");
  if (a->n_entries < 0 || a->n_nodes < 1 || a->nodes_per_sign < 1 ||
      a->n_nodes % a->nodes_per_sign != 0)
    return fail(VL1_E_ARG, "vl1_duffy_build: bad node counts");
  if (!(a->source_box_extent > 0)) return fail(VL1_E_ARG, "vl1_duffy_build: extent must be > 0");
  if (a->n_entries == 0) return VL1_OK;
  if (!a->target_points || !a->decomposition || !a->mode_axes || !a->node_u || !a->node_sign ||
      !a->node_jac || !a->interp_nodes || !a->bary_w || !a->result)
    return fail(VL1_E_ARG, "vl1_duffy_build: null device array");
  const bool deriv = kind_is_deriv(a->kernel_kind);
  if (deriv && (a->deriv_axis < 0 || a->deriv_axis >= a->dim))
    return fail(VL1_E_ARG, "vl1_duffy_build: derivative axis out of range");
  if (a->chunk_start && (a->n_chunks < 1 || a->n_chunks > INT32_MAX))
    return fail(VL1_E_ARG, "vl1_duffy_build: bad chunk count");
  const bool radial = a->kernel_kind == VL1_KERNEL_POWER || a->kernel_kind == VL1_KERNEL_POWER_LOG ||
                      kind_is_power_deriv(a->kernel_kind);
  const bool needs_pos = a->kernel_kind != VL1_KERNEL_POWER;  // log terms, derivatives
  if (radial && (a->power < 0 || a->power > 32 || (needs_pos && a->power == 0)))
    return fail(VL1_E_ARG,
                "vl1_duffy_build: power must be 0..32 (1..32 with the log or a derivative)");
  DuffyParams prm{a->dim, a->quad_order, a->kernel_kind, deriv ? a->deriv_axis : -1,
                  radial ? (double)a->power : a->helmholtz_k, a->n_entries,
                  a->n_nodes, a->nodes_per_sign, a->source_box_extent, a->target_points,
                  a->decomposition, a->mode_axes, a->node_u, a->node_sign, a->node_jac,
                  a->interp_nodes, a->bary_w, a->result, a->chunk_start,
                  a->chunk_start ? a->n_chunks : 0};
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  const int kind = a->kernel_kind;
  if (a->dim == 2) {
    if (kind == VL1_KERNEL_CONSTANT) e = launch_duffy<2, VL1_KERNEL_CONSTANT>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE) e = launch_duffy<2, VL1_KERNEL_LAPLACE>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ) e = launch_duffy<2, VL1_KERNEL_HELMHOLTZ>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE_TARGET_DERIV)
      e = launch_duffy<2, VL1_KERNEL_LAPLACE_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE_SOURCE_DERIV)
      e = launch_duffy<2, VL1_KERNEL_LAPLACE_SOURCE_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ_TARGET_DERIV)
      e = launch_duffy<2, VL1_KERNEL_HELMHOLTZ_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV)
      e = launch_duffy<2, VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER) e = launch_duffy<2, VL1_KERNEL_POWER>(prm, s);
    else if (kind == VL1_KERNEL_POWER_LOG) e = launch_duffy<2, VL1_KERNEL_POWER_LOG>(prm, s);
    else if (kind == VL1_KERNEL_POWER_TARGET_DERIV)
      e = launch_duffy<2, VL1_KERNEL_POWER_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER_SOURCE_DERIV)
      e = launch_duffy<2, VL1_KERNEL_POWER_SOURCE_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER_LOG_TARGET_DERIV)
      e = launch_duffy<2, VL1_KERNEL_POWER_LOG_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER_LOG_SOURCE_DERIV)
      e = launch_duffy<2, VL1_KERNEL_POWER_LOG_SOURCE_DERIV>(prm, s);
    else return fail(VL1_E_UNSUPPORTED, "vl1_duffy_build: kernel kind");
  } else {
    if (kind == VL1_KERNEL_CONSTANT) e = launch_duffy<3, VL1_KERNEL_CONSTANT>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE) e = launch_duffy<3, VL1_KERNEL_LAPLACE>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ) e = launch_duffy<3, VL1_KERNEL_HELMHOLTZ>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE_TARGET_DERIV)
      e = launch_duffy<3, VL1_KERNEL_LAPLACE_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE_SOURCE_DERIV)
      e = launch_duffy<3, VL1_KERNEL_LAPLACE_SOURCE_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ_TARGET_DERIV)
      e = launch_duffy<3, VL1_KERNEL_HELMHOLTZ_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV)
      e = launch_duffy<3, VL1_KERNEL_HELMHOLTZ_SOURCE_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER) e = launch_duffy<3, VL1_KERNEL_POWER>(prm, s);
    else if (kind == VL1_KERNEL_POWER_LOG) e = launch_duffy<3, VL1_KERNEL_POWER_LOG>(prm, s);
    else if (kind == VL1_KERNEL_POWER_TARGET_DERIV)
      e = launch_duffy<3, VL1_KERNEL_POWER_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER_SOURCE_DERIV)
      e = launch_duffy<3, VL1_KERNEL_POWER_SOURCE_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER_LOG_TARGET_DERIV)
      e = launch_duffy<3, VL1_KERNEL_POWER_LOG_TARGET_DERIV>(prm, s);
    else if (kind == VL1_KERNEL_POWER_LOG_SOURCE_DERIV)
      e = launch_duffy<3, VL1_KERNEL_POWER_LOG_SOURCE_DERIV>(prm, s);
    else return fail(VL1_E_UNSUPPORTED, "vl1_duffy_build: kernel kind");
  }
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("duffy launch: ") + hipGetErrorString(e));
  return VL1_OK;
}
