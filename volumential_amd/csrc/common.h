// Shared host/device helpers of the vl1 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "vl1.h"

namespace vl1 {

// Thread-local error message returned by vl1_last_error().
void set_error(const std::string& msg);
void clear_error();

// Record `msg` and return `code`.
int fail(int code, const std::string& msg);

#define VL1_HIP_TRY(expr)                                                        \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      return ::vl1::fail(VL1_E_HIP, std::string(#expr) + ": " +                  \
                                        hipGetErrorString(_e));                  \
    }                                                                            \
  } while (0)

// Device buffer owned by payloads and plans.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  // On failure the buffer stays empty (null, 0 bytes) and the runtime's
  // recorded error is cleared, so a later call neither sees a stale
  // capacity nor inherits the failed allocation's error from
  // hipGetLastError().
  hipError_t alloc(size_t n) {
    release();
    if (n == 0) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
    ptr = p;
    bytes = n;
    return hipSuccess;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
};

inline hipError_t upload(DevBuf& b, const void* host, size_t bytes, hipStream_t s) {
  hipError_t e = b.alloc(bytes);
  if (e != hipSuccess || bytes == 0) return e;
  e = hipMemcpyAsync(b.ptr, host, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

// Compact row of table entry (case, s, t) relative to the first row of the
// case's canonical orbit (orbit_arithmetic.py:372-432 and 566-634).  The
// pair ranks only depend on the multiset of each group's pair codes, so the
// compare-exchange network of the reference does not change the result and
// is replaced by sorting the (at most three) codes of each group.
__host__ __device__ inline int64_t compact_rel_row(int dim, int q, const uint8_t* perm,
                                                  const int8_t* sign, const uint8_t* group,
                                                  int s, int t) {
  int sr[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
  for (int ax = dim - 1; ax >= 0; --ax) {
    sr[ax] = s % q;
    s /= q;
    tr[ax] = t % q;
    t /= q;
  }
  int pair[3];
  for (int o = 0; o < dim; ++o) {
    const int in = perm[o];
    int si = sr[in], ti = tr[in];
    const int sf = q - 1 - si, tf = q - 1 - ti;
    const int sg = sign[o];
    const bool flip = (sg < 0) || (sg == 0 && (sf < si || (sf == si && tf < ti)));
    if (flip) {
      si = sf;
      ti = tf;
    }
    pair[o] = si * q + ti;
  }
  const int64_t pc = (int64_t)q * q;
  const int64_t fpc = (pc + 1) / 2;
  int64_t compact = 0;
  for (int g = 0; g < dim; ++g) {
    int v[3];
    int k = 0;
    bool zero = false;
    for (int o = 0; o < dim; ++o) {
      if (group[o] == g) {
        v[k++] = pair[o];
        zero = zero || (sign[o] == 0);
      }
    }
    if (k == 0) continue;
    const int64_t A = zero ? fpc : pc;
    // sort up to three codes
    if (k >= 2 && v[1] < v[0]) { int x = v[0]; v[0] = v[1]; v[1] = x; }
    if (k == 3) {
      if (v[2] < v[1]) { int x = v[1]; v[1] = v[2]; v[2] = x; }
      if (v[1] < v[0]) { int x = v[0]; v[0] = v[1]; v[1] = x; }
    }
    int64_t rank, count;
    if (k == 1) {
      rank = v[0];
      count = A;
    } else if (k == 2) {
      const int64_t lo = v[0], hi = v[1];
      rank = A * (A + 1) / 2 - (A - lo) * (A - lo + 1) / 2 + (hi - lo);
      count = A * (A + 1) / 2;
    } else {
      const int64_t lo = v[0], mid = v[1], hi = v[2];
      rank = A * (A + 1) * (A + 2) / 6 - (A - lo) * (A - lo + 1) * (A - lo + 2) / 6 +
             (A - lo) * (A - lo + 1) / 2 - (A - mid) * (A - mid + 1) / 2 + (hi - mid);
      count = A * (A + 1) * (A + 2) / 6;
    }
    compact = compact * count + rank;
  }
  return compact;
}

// Sign of entry (s, t) of one case for signed (axis-derivative) kernels:
// -1 when the canonicalisation above flips an odd number of axes whose input
// axis has its bit set in sp_mask (orbit_arithmetic.py:372-432, entry_sign;
// list1.py:799-832, entry_sign_factor).  +1 for scalar kernels (sp_mask 0).
__host__ __device__ inline int entry_sign(int dim, int q, const uint8_t* perm, const int8_t* sign,
                                          int sp_mask, int s, int t) {
  if (!sp_mask) return 1;
  int sr[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
  for (int ax = dim - 1; ax >= 0; --ax) {
    sr[ax] = s % q;
    s /= q;
    tr[ax] = t % q;
    t /= q;
  }
  int neg = 0;
  for (int o = 0; o < dim; ++o) {
    const int in = perm[o];
    const int si = sr[in], ti = tr[in];
    const int sf = q - 1 - si, tf = q - 1 - ti;
    const int sg = sign[o];
    const bool flip = (sg < 0) || (sg == 0 && (sf < si || (sf == si && tf < ti)));
    if (flip && ((sp_mask >> in) & 1)) neg ^= 1;
  }
  return neg ? -1 : 1;
}

}  // namespace vl1
