// Internal object layouts shared by the translation units of libvl1.
#pragma once

#include <vector>

#include "common.h"

struct vl1_payload {
  int dim = 0, q = 0, n = 0, n_cases = 0, n_orbits = 0, n_tables = 0, is_complex = 0;
  int64_t n_entries = 0;   // compact rows per table level
  int rowmap_u16 = 1;      // rowmap element type: 1 = uint16, 0 = int32
  int sign_mask = 0;       // bit a set: axis a has sign power 1 (signed kernels)
  int max_orbit_count = 0; // largest canonical orbit (rows)
  int dense_kind = 0;      // 1: the reference's "dense" kind (expanded blocks only)
  int has_missing = 0;     // dense kind: some (case, s, t) has no entry (NaN in `dense`)
  // host copies of the small descriptors
  std::vector<int64_t> case_indices;
  std::vector<uint16_t> orbit_ranks;
  std::vector<uint8_t> perm, group;
  std::vector<int8_t> sign;
  std::vector<int32_t> offsets;
  // device buffers
  vl1::DevBuf table;       // [n_tables][n_entries] (x2 complex)
  vl1::DevBuf nmlz;        // [n_tables][n] (x2 complex)
  vl1::DevBuf case_idx;    // int64 [13^dim]
  vl1::DevBuf ranks;       // uint16 [n_cases]
  vl1::DevBuf offs;        // int32 [n_orbits+1]
  vl1::DevBuf rowmap;      // [n_cases][n (s)][n (t)] relative compact rows; for signed
                           // kernels (int32) ~row marks an entry read negated
  // [n_tables][n_cases][n (s)][n (t)] table values (sign applied; x2 complex),
  // expanded on the first plain evaluation (mutable: a cache of the payload);
  // built at creation for the dense kind, where a quiet NaN marks an entry
  // without data
  mutable vl1::DevBuf dense;
  int64_t device_bytes() const {
    return (int64_t)(table.bytes + nmlz.bytes + case_idx.bytes + ranks.bytes + offs.bytes +
                     rowmap.bytes + dense.bytes);
  }
};

namespace vl1 {
// Launchers implemented in the .hip translation units.
hipError_t launch_rowmap(const vl1_payload& p, const uint8_t* d_perm, const int8_t* d_sign,
                         const uint8_t* d_group, hipStream_t s);
// Decode a signed row-map element: row >= 0 as is, ~row read negated.
__host__ __device__ inline int decode_signed_row(int v, double& sgn) {
  sgn = v < 0 ? -1.0 : 1.0;
  return v < 0 ? ~v : v;
}
}  // namespace vl1
