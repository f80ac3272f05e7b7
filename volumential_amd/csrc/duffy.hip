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
// GPU mapping: one lane per entry; the node loop is sequential and all lanes
// of a wavefront read the same node (uniform, scalar-cached loads).  A sign
// block whose sub-box is empty for the entry (some len_k == 0) is skipped as
// a whole: every node in it is inactive.
#include <cmath>
#include <string>

#include "internal.h"

namespace vl1 {

struct DuffyParams {
  int dim, q, kind;
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
};

__device__ inline double lagrange_basis(double y, int m, int q, const double* xi, const double* bw,
                                        double tol) {
  if (q == 1) return 1.0;
  int hit_any = 0, hit_mode = 0;
  double den = 0.0, num = 0.0;
  for (int j = 0; j < q; ++j) {
    const double diff = y - xi[j];
    const bool hit = fabs(diff) < tol;
    hit_any |= hit;
    if (j == m) hit_mode = hit;
    if (!hit) {
      const double term = bw[j] / diff;
      den += term;
      if (j == m) num = term;
    }
  }
  return hit_any ? (double)hit_mode : num / den;
}

// Kernel values, sumpy conventions (expression times global scaling):
// Laplace 2D -log(r)/(2 pi), 3D 1/(4 pi r); Helmholtz 3D e^{ikr}/(4 pi r).
template <int DIM, int KIND>
__device__ inline void kernel_value(double r2, double k, double& re, double& im) {
  if (KIND == VL1_KERNEL_CONSTANT) {
    re = 1.0;
    im = 0.0;
  } else if (KIND == VL1_KERNEL_LAPLACE) {
    if (DIM == 2) {
      re = -0.25 / M_PI * log(r2);
    } else {
      re = 0.25 / M_PI / sqrt(r2);
    }
    im = 0.0;
  } else {
    const double r = sqrt(r2);
    double sn, cs;
    sincos(k * r, &sn, &cs);
    re = 0.25 / M_PI * cs / r;
    im = 0.25 / M_PI * sn / r;
  }
}

template <int DIM, int KIND>
__global__ void __launch_bounds__(256) duffy_kernel(const DuffyParams p) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p.n_entries) return;
  constexpr int kMaxQ = 16;
  double xi[kMaxQ], bw[kMaxQ];
  for (int j = 0; j < p.q; ++j) {
    xi[j] = p.xi[j];
    bw[j] = p.bw[j];
  }
  double x[DIM], v[DIM];
  int m[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    x[a] = p.xt[a * p.n_entries + e];
    v[a] = p.vt[a * p.n_entries + e];
    m[a] = p.mode[a * p.n_entries + e];
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
      sgn[a] = p.sg[a * p.n_nodes + base];
      len[a] = sgn[a] < 0.0 ? v[a] : p.H - v[a];
      pos = pos && (len[a] > 0.0);
    }
    if (!pos) continue;
    for (int64_t i = 0; i < p.nodes_per_sign; ++i) {
      const int64_t node = base + i;
      double r2 = 0.0;
      double yk[DIM];
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        yk[a] = v[a] + sgn[a] * len[a] * p.u[a * p.n_nodes + node];
        const double d = x[a] - yk[a];
        r2 += d * d;
      }
      if (!(r2 > 0.0)) continue;
      double w = p.jac[node];
#pragma unroll
      for (int a = 0; a < DIM; ++a) w *= len[a];
      double basis = 1.0;
#pragma unroll
      for (int a = 0; a < DIM; ++a) basis *= lagrange_basis(yk[a], m[a], p.q, xi, bw, tol);
      double kr, ki;
      kernel_value<DIM, KIND>(r2, p.k, kr, ki);
      const double wb = w * basis;
      acc_re = fma(kr, wb, acc_re);
      if (KIND == VL1_KERNEL_HELMHOLTZ) acc_im = fma(ki, wb, acc_im);
    }
  }
  if (KIND == VL1_KERNEL_HELMHOLTZ) {
    p.out[2 * e] = acc_re;
    p.out[2 * e + 1] = acc_im;
  } else {
    p.out[e] = acc_re;
  }
}

template <int DIM, int KIND>
static hipError_t launch_duffy(const DuffyParams& prm, hipStream_t s) {
  const int block = 256;
  const int64_t grid = (prm.n_entries + block - 1) / block;
  hipLaunchKernelGGL((duffy_kernel<DIM, KIND>), dim3((unsigned)grid), dim3(block), 0, s, prm);
  return hipGetLastError();
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
  if (a->kernel_kind == VL1_KERNEL_HELMHOLTZ && a->dim == 2)
    return fail(VL1_E_UNSUPPORTED, "vl1_duffy_build: 2D Helmholtz tables are not implemented");
  if (a->n_entries == 0) return VL1_OK;
  if (!a->target_points || !a->decomposition || !a->mode_axes || !a->node_u || !a->node_sign ||
      !a->node_jac || !a->interp_nodes || !a->bary_w || !a->result)
    return fail(VL1_E_ARG, "vl1_duffy_build: null device array");
  DuffyParams prm{a->dim, a->quad_order, a->kernel_kind, a->helmholtz_k, a->n_entries,
                  a->n_nodes, a->nodes_per_sign, a->source_box_extent, a->target_points,
                  a->decomposition, a->mode_axes, a->node_u, a->node_sign, a->node_jac,
                  a->interp_nodes, a->bary_w, a->result};
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  const int kind = a->kernel_kind;
  if (a->dim == 2) {
    if (kind == VL1_KERNEL_CONSTANT) e = launch_duffy<2, VL1_KERNEL_CONSTANT>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE) e = launch_duffy<2, VL1_KERNEL_LAPLACE>(prm, s);
    else return fail(VL1_E_UNSUPPORTED, "vl1_duffy_build: kernel kind");
  } else {
    if (kind == VL1_KERNEL_CONSTANT) e = launch_duffy<3, VL1_KERNEL_CONSTANT>(prm, s);
    else if (kind == VL1_KERNEL_LAPLACE) e = launch_duffy<3, VL1_KERNEL_LAPLACE>(prm, s);
    else if (kind == VL1_KERNEL_HELMHOLTZ) e = launch_duffy<3, VL1_KERNEL_HELMHOLTZ>(prm, s);
    else return fail(VL1_E_UNSUPPORTED, "vl1_duffy_build: kernel kind");
  }
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("duffy launch: ") + hipGetErrorString(e));
  return VL1_OK;
}
