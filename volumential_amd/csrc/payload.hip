// Payload: device-resident near-field table data and orbit descriptors.
//
// Mirrors _prepare_table_data_and_entry_map / _get_cached_nearfield_payload
// (volumential/expansion_wrangler_fpnd.py:1805-2066, 2939-3158): the table
// levels arrive already packed in compact-row order; this uploads them with
// the case descriptors and expands, once, the per-case map
//     rowmap[case][s][t] = compact row of (case, s, t) - first row of its orbit
// so evaluation kernels gather table values without redoing the orbit
// arithmetic for every (box pair, s, t).
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "internal.h"

namespace vl1 {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

template <typename IdxT>
__global__ void __launch_bounds__(256)
rowmap_kernel(int dim, int q, int n, int n_cases, int sp_mask, const uint8_t* __restrict__ perm,
              const int8_t* __restrict__ sign, const uint8_t* __restrict__ group,
              IdxT* __restrict__ out) {
  const int64_t total = (int64_t)n_cases * n * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(i % n);
    const int64_t cs = i / n;
    const int s = (int)(cs % n);
    const int c = (int)(cs / n);
    const int64_t rel = compact_rel_row(dim, q, perm + c * dim, sign + c * dim, group + c * dim, s, t);
    // sp_mask != 0 only with int32 maps: a negated entry is stored as ~rel
    const bool neg = entry_sign(dim, q, perm + c * dim, sign + c * dim, sp_mask, s, t) < 0;
    out[i] = (IdxT)(neg ? ~rel : rel);
  }
}

hipError_t launch_rowmap(const vl1_payload& p, const uint8_t* d_perm, const int8_t* d_sign,
                         const uint8_t* d_group, hipStream_t s) {
  const int64_t total = (int64_t)p.n_cases * p.n * p.n;
  const int block = 256;
  const int grid = (int)std::min<int64_t>((total + block - 1) / block, 65536);
  if (p.rowmap_u16) {
    hipLaunchKernelGGL(rowmap_kernel<uint16_t>, dim3(grid), dim3(block), 0, s, p.dim, p.q, p.n,
                       p.n_cases, 0, d_perm, d_sign, d_group, p.rowmap.as<uint16_t>());
  } else {
    hipLaunchKernelGGL(rowmap_kernel<int32_t>, dim3(grid), dim3(block), 0, s, p.dim, p.q, p.n,
                       p.n_cases, p.sign_mask, d_perm, d_sign, d_group, p.rowmap.as<int32_t>());
  }
  return hipGetLastError();
}

// Dense kind: D[lev][c][s][t] = T[lev][id] * entry_scales[f] * mode_case_scale[s][c]
// with f = (mcm[s][c] n + mqm[s][s]) n + mqm[s][t], id = ids[f]; a quiet NaN
// (both parts) where id < 0.  Complex values multiply as complex numbers.
__global__ void __launch_bounds__(256)
dense_kind_kernel(int n, int n_cases, int n_tables, int cplx, int64_t n_entries,
                  const double* __restrict__ table, const int32_t* __restrict__ ids,
                  const double* __restrict__ scales, const int32_t* __restrict__ mqm,
                  const int32_t* __restrict__ mcm, const double* __restrict__ mcs,
                  double* __restrict__ D) {
  const int64_t nn = (int64_t)n * n;
  const int64_t total = (int64_t)n_tables * n_cases * nn;
  const int vw = cplx ? 2 : 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(i % n);
    const int s = (int)((i / n) % n);
    const int c = (int)((i / nn) % n_cases);
    const int lev = (int)(i / (nn * n_cases));
    const int64_t f = ((int64_t)mcm[(int64_t)s * n_cases + c] * n + mqm[(int64_t)s * n + s]) * n +
                      mqm[(int64_t)s * n + t];
    const int id = ids[f];
    if (id < 0) {
      D[i * vw] = __builtin_nan("");
      if (cplx) D[i * vw + 1] = __builtin_nan("");
      continue;
    }
    const double* v = table + ((int64_t)lev * n_entries + id) * vw;
    const double* a = scales + f * vw;
    const double* b = mcs + ((int64_t)s * n_cases + c) * vw;
    if (!cplx) {
      D[i] = v[0] * a[0] * b[0];
    } else {
      const double re = v[0] * a[0] - v[1] * a[1], im = v[0] * a[1] + v[1] * a[0];
      D[i * 2] = re * b[0] - im * b[1];
      D[i * 2 + 1] = re * b[1] + im * b[0];
    }
  }
}

}  // namespace vl1

using namespace vl1;

extern "C" {

const char* vl1_last_error(void) { return g_last_error.c_str(); }

int vl1_abi_version(void) { return VL1_ABI_VERSION; }

int vl1_payload_create(vl1_payload** out, const vl1_payload_desc* d, void* stream) {
  clear_error();
  if (!out || !d) return fail(VL1_E_ARG, "vl1_payload_create: null argument");
  *out = nullptr;
  if (d->dim != 2 && d->dim != 3)
    return fail(VL1_E_UNSUPPORTED, "arithmetic orbit payloads support dim 2 or 3");
  if (d->quad_order < 1 || d->quad_order > 16) return fail(VL1_E_ARG, "quad_order out of range");
  if (d->n_tables < 1 || d->n_table_entries < 1 || d->n_cases < 1 || d->n_case_orbits < 1)
    return fail(VL1_E_ARG, "empty table payload");
  if (!d->table_data || !d->mode_nmlz || !d->case_indices || !d->case_orbit_ranks ||
      !d->case_axis_perm || !d->case_axis_sign || !d->case_axis_group || !d->case_value_offsets)
    return fail(VL1_E_ARG, "vl1_payload_create: null array");
  hipStream_t s = (hipStream_t)stream;

  auto p = new vl1_payload();
  p->dim = d->dim;
  p->q = d->quad_order;
  p->n = 1;
  for (int i = 0; i < d->dim; ++i) p->n *= d->quad_order;
  p->n_cases = d->n_cases;
  p->n_orbits = d->n_case_orbits;
  p->n_tables = d->n_tables;
  p->is_complex = d->is_complex ? 1 : 0;
  p->n_entries = d->n_table_entries;

  // validate descriptors on the host
  p->offsets.assign(d->case_value_offsets, d->case_value_offsets + d->n_case_orbits + 1);
  p->orbit_ranks.assign(d->case_orbit_ranks, d->case_orbit_ranks + d->n_cases);
  p->perm.assign(d->case_axis_perm, d->case_axis_perm + d->n_cases * d->dim);
  p->sign.assign(d->case_axis_sign, d->case_axis_sign + d->n_cases * d->dim);
  p->group.assign(d->case_axis_group, d->case_axis_group + d->n_cases * d->dim);
  int64_t n_codes = 1;
  for (int i = 0; i < d->dim; ++i) n_codes *= 13;
  p->case_indices.assign(d->case_indices, d->case_indices + n_codes);
  if (p->offsets[0] != 0 || p->offsets.back() != d->n_table_entries) {
    delete p;
    return fail(VL1_E_TABLE, "case_value_offsets do not span the compact table");
  }
  for (int r = 0; r < d->n_case_orbits; ++r) {
    const int cnt = p->offsets[r + 1] - p->offsets[r];
    if (cnt <= 0) {
      delete p;
      return fail(VL1_E_TABLE, "empty canonical orbit in case_value_offsets");
    }
    p->max_orbit_count = std::max(p->max_orbit_count, cnt);
  }
  for (int c = 0; c < d->n_cases; ++c) {
    if (p->orbit_ranks[c] >= d->n_case_orbits) {
      delete p;
      return fail(VL1_E_TABLE, "case_orbit_ranks out of range");
    }
    for (int a = 0; a < d->dim; ++a)
      if (p->perm[c * d->dim + a] >= d->dim || p->group[c * d->dim + a] >= d->dim) {
        delete p;
        return fail(VL1_E_TABLE, "case axis descriptors out of range");
      }
  }
  const int64_t nvals = (int64_t)d->n_tables * d->n_table_entries * (p->is_complex ? 2 : 1);
  for (int64_t i = 0; i < nvals; ++i) {
    if (!std::isfinite(d->table_data[i])) {
      delete p;
      return fail(VL1_E_TABLE, "near-field table contains non-finite entries");
    }
  }
  if (d->axis_sign_power)
    for (int a = 0; a < d->dim; ++a) {
      if (d->axis_sign_power[a] > 1) {
        delete p;
        return fail(VL1_E_TABLE, "axis_sign_power entries must be 0 or 1");
      }
      p->sign_mask |= (int)d->axis_sign_power[a] << a;
    }
  // signed kernels keep the entry sign in the row map (int32, ~row negated)
  p->rowmap_u16 = (p->max_orbit_count <= 65535 && p->sign_mask == 0) ? 1 : 0;

  hipError_t e = hipSuccess;
  const size_t vbytes = sizeof(double) * (p->is_complex ? 2 : 1);
  e = upload(p->table, d->table_data, vbytes * d->n_tables * d->n_table_entries, s);
  if (e == hipSuccess) e = upload(p->nmlz, d->mode_nmlz, vbytes * d->n_tables * p->n, s);
  if (e == hipSuccess)
    e = upload(p->case_idx, p->case_indices.data(), sizeof(int64_t) * n_codes, s);
  if (e == hipSuccess)
    e = upload(p->ranks, p->orbit_ranks.data(), sizeof(uint16_t) * d->n_cases, s);
  if (e == hipSuccess)
    e = upload(p->offs, p->offsets.data(), sizeof(int32_t) * p->offsets.size(), s);
  DevBuf dperm, dsign, dgroup;
  if (e == hipSuccess) e = upload(dperm, p->perm.data(), p->perm.size(), s);
  if (e == hipSuccess) e = upload(dsign, p->sign.data(), p->sign.size(), s);
  if (e == hipSuccess) e = upload(dgroup, p->group.data(), p->group.size(), s);
  if (e == hipSuccess)
    e = p->rowmap.alloc((size_t)p->n_cases * p->n * p->n * (p->rowmap_u16 ? 2 : 4));
  if (e == hipSuccess)
    e = launch_rowmap(*p, dperm.as<uint8_t>(), dsign.as<int8_t>(), dgroup.as<uint8_t>(), s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete p;
    return fail(VL1_E_HIP, std::string("vl1_payload_create: ") + hipGetErrorString(e));
  }
  *out = p;
  return VL1_OK;
}

int vl1_payload_create_dense(vl1_payload** out, const vl1_dense_payload_desc* d, void* stream) {
  clear_error();
  if (!out || !d) return fail(VL1_E_ARG, "vl1_payload_create_dense: null argument");
  *out = nullptr;
  if (d->dim != 2 && d->dim != 3)
    return fail(VL1_E_UNSUPPORTED, "dense payloads support dim 2 or 3");
  if (d->quad_order < 1 || d->quad_order > 16) return fail(VL1_E_ARG, "quad_order out of range");
  if (d->n_tables < 1 || d->n_table_entries < 1 || d->n_cases < 1)
    return fail(VL1_E_ARG, "empty table payload");
  if (!d->table_data || !d->mode_nmlz || !d->case_indices || !d->table_entry_ids ||
      !d->table_entry_scales || !d->mode_qpoint_map || !d->mode_case_map || !d->mode_case_scale)
    return fail(VL1_E_ARG, "vl1_payload_create_dense: null array");
  if (d->n_table_entries > INT32_MAX)
    return fail(VL1_E_UNSUPPORTED, "dense payload: entry ids are int32");
  hipStream_t s = (hipStream_t)stream;
  int n = 1;
  for (int i = 0; i < d->dim; ++i) n *= d->quad_order;
  const int vw = d->is_complex ? 2 : 1;
  const int64_t nn = (int64_t)n * n, nfull = (int64_t)d->n_cases * nn;
  const double dense_bytes = (double)d->n_tables * nfull * vw * sizeof(double);
  if (dense_bytes > 8.0 * (1ull << 30))
    return fail(VL1_E_UNSUPPORTED, "dense payload: expanded blocks exceed 8 GiB");

  // host checks: every index the reconstruction forms is in range
  for (int64_t i = 0; i < nn; ++i)
    if (d->mode_qpoint_map[i] < 0 || d->mode_qpoint_map[i] >= n)
      return fail(VL1_E_TABLE, "mode_qpoint_map entries must lie in [0, n_q_points)");
  for (int64_t i = 0; i < (int64_t)n * d->n_cases; ++i)
    if (d->mode_case_map[i] < 0 || d->mode_case_map[i] >= d->n_cases)
      return fail(VL1_E_TABLE, "mode_case_map entries must lie in [0, n_cases)");
  for (int64_t i = 0; i < nfull; ++i)
    if (d->table_entry_ids[i] >= d->n_table_entries)
      return fail(VL1_E_TABLE, "table_entry_ids beyond the table data");
  for (int64_t i = 0; i < (int64_t)d->n_tables * d->n_table_entries * vw; ++i)
    if (!std::isfinite(d->table_data[i]))
      return fail(VL1_E_TABLE, "near-field table contains non-finite entries");
  for (int64_t i = 0; i < nfull * vw; ++i)
    if (!std::isfinite(d->table_entry_scales[i]))
      return fail(VL1_E_TABLE, "table_entry_scales contains non-finite entries");
  for (int64_t i = 0; i < (int64_t)n * d->n_cases * vw; ++i)
    if (!std::isfinite(d->mode_case_scale[i]))
      return fail(VL1_E_TABLE, "mode_case_scale contains non-finite entries");
  // entries without data that an evaluation can reach
  int missing = 0;
  for (int sm = 0; sm < n && !missing; ++sm) {
    const int32_t* mq = d->mode_qpoint_map + (int64_t)sm * n;
    for (int c = 0; c < d->n_cases && !missing; ++c) {
      const int64_t base = ((int64_t)d->mode_case_map[(int64_t)sm * d->n_cases + c] * n + mq[sm]) * n;
      for (int t = 0; t < n; ++t)
        if (d->table_entry_ids[base + mq[t]] < 0) {
          missing = 1;
          break;
        }
    }
  }

  auto p = new vl1_payload();
  p->dim = d->dim;
  p->q = d->quad_order;
  p->n = n;
  p->n_cases = d->n_cases;
  p->n_orbits = 0;
  p->n_tables = d->n_tables;
  p->is_complex = d->is_complex ? 1 : 0;
  p->n_entries = d->n_table_entries;
  p->dense_kind = 1;
  p->has_missing = missing;
  int64_t n_codes = 1;
  for (int i = 0; i < d->dim; ++i) n_codes *= 13;
  p->case_indices.assign(d->case_indices, d->case_indices + n_codes);
  for (int64_t v : p->case_indices)
    if (v >= d->n_cases) {
      delete p;
      return fail(VL1_E_TABLE, "case_indices beyond n_cases");
    }

  const size_t vbytes = sizeof(double) * vw;
  hipError_t e = upload(p->table, d->table_data, vbytes * d->n_tables * d->n_table_entries, s);
  if (e == hipSuccess) e = upload(p->nmlz, d->mode_nmlz, vbytes * d->n_tables * n, s);
  if (e == hipSuccess)
    e = upload(p->case_idx, p->case_indices.data(), sizeof(int64_t) * n_codes, s);
  DevBuf dids, dsc, dmq, dmc, dms;
  if (e == hipSuccess) e = upload(dids, d->table_entry_ids, sizeof(int32_t) * nfull, s);
  if (e == hipSuccess) e = upload(dsc, d->table_entry_scales, vbytes * nfull, s);
  if (e == hipSuccess) e = upload(dmq, d->mode_qpoint_map, sizeof(int32_t) * nn, s);
  if (e == hipSuccess) e = upload(dmc, d->mode_case_map, sizeof(int32_t) * n * d->n_cases, s);
  if (e == hipSuccess) e = upload(dms, d->mode_case_scale, vbytes * n * d->n_cases, s);
  if (e == hipSuccess) e = p->dense.alloc((size_t)d->n_tables * nfull * vbytes);
  if (e == hipSuccess) {
    const int64_t total = (int64_t)d->n_tables * nfull;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(dense_kind_kernel, dim3(grid), dim3(256), 0, s, n, d->n_cases, d->n_tables,
                       p->is_complex, d->n_table_entries, p->table.as<double>(),
                       dids.as<int32_t>(), dsc.as<double>(), dmq.as<int32_t>(), dmc.as<int32_t>(),
                       dms.as<double>(), p->dense.as<double>());
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // before the uploads are freed
  if (e != hipSuccess) {
    delete p;
    return fail(VL1_E_HIP, std::string("vl1_payload_create_dense: ") + hipGetErrorString(e));
  }
  *out = p;
  return VL1_OK;
}

int vl1_payload_has_missing(const vl1_payload* p) { return p ? p->has_missing : 0; }

int vl1_payload_destroy(vl1_payload* p) {
  delete p;
  return VL1_OK;
}

int64_t vl1_payload_device_bytes(const vl1_payload* p) { return p ? p->device_bytes() : 0; }

// Host evaluation of the compact-row function the device kernels use; lets
// CPU-only tests check it against the orbit enumeration.
int vl1_orbit_rows_host(int dim, int q, const uint8_t* perm, const int8_t* sign,
                        const uint8_t* group, const int32_t* s, const int32_t* t, int64_t count,
                        int64_t* out) {
  clear_error();
  if ((dim != 2 && dim != 3) || q < 1 || !perm || !sign || !group || !s || !t || !out)
    return fail(VL1_E_ARG, "vl1_orbit_rows_host: bad argument");
  for (int64_t i = 0; i < count; ++i) out[i] = compact_rel_row(dim, q, perm, sign, group, s[i], t[i]);
  return VL1_OK;
}

int vl1_orbit_signs_host(int dim, int q, const uint8_t* perm, const int8_t* sign,
                         const uint8_t* axis_sign_power, const int32_t* s, const int32_t* t,
                         int64_t count, int8_t* out) {
  clear_error();
  if ((dim != 2 && dim != 3) || q < 1 || !perm || !sign || !axis_sign_power || !s || !t || !out)
    return fail(VL1_E_ARG, "vl1_orbit_signs_host: bad argument");
  int mask = 0;
  for (int a = 0; a < dim; ++a) mask |= (axis_sign_power[a] & 1) << a;
  for (int64_t i = 0; i < count; ++i) out[i] = (int8_t)entry_sign(dim, q, perm, sign, mask, s[i], t[i]);
  return VL1_OK;
}

}  // extern "C"
