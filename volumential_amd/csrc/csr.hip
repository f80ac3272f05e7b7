// List-1 CSR producer on the GPU: for every leaf of a 2:1-balanced 2^d-tree
// on a cube, the leaves touching it (itself included), sorted by leaf id --
// the neighbour lists the reference gets from boxtree's traversal
// (meshgen.py:391-444, tree_interactive_build.py:202-470) and that the
// list-1 evaluation consumes (neighbor_source_boxes_starts / _lists).
//
// Leaves are identified by (level, integer index at that level).  The host
// sorts the keys (level, Morton code of the index) once; one thread per leaf
// then walks the 3^d neighbour cells of its own size and finds, by binary
// search in the sorted keys, the leaf at the same level, its parent's level
// (a coarser neighbour) or the 2^d children touching the leaf (finer
// neighbours).  A count pass sizes every list, the host scans the counts, and
// a fill pass writes the deduplicated, sorted lists.
#include <algorithm>
#include <string>
#include <vector>

#include "internal.h"

namespace vl1 {

constexpr int kMaxNbrs = 64;  // 2:1 balance: at most 57 in 3D (self + 56 finer)

__host__ __device__ inline uint64_t morton_key(int dim, int level, const int32_t* ix) {
  uint64_t m = 0;
  for (int b = 0; b < 19; ++b)
    for (int a = 0; a < dim; ++a) m |= (uint64_t)((ix[a] >> b) & 1) << (b * dim + (dim - 1 - a));
  return ((uint64_t)level << 58) | m;
}

__device__ inline int find_leaf(const uint64_t* keys, const int32_t* ids, int n, uint64_t k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1; else hi = mid;
  }
  return (lo < n && keys[lo] == k) ? ids[lo] : -1;
}

// Neighbour list of leaf b into out[]; returns the count, or -1 if the tree
// is not 2:1 balanced / the leaves do not tile the domain.
__device__ inline int leaf_neighbours(int dim, int b, const uint8_t* levels, const int32_t* index,
                                      const uint64_t* keys, const int32_t* ids, int n, int lmax,
                                      int32_t* out) {
  const int lv = levels[b];
  int ib[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) ib[a] = index[(int64_t)b * dim + a];
  const int span = 1 << lv;
  int cnt = 0;
  auto add = [&](int id) {
    for (int i = 0; i < cnt; ++i)
      if (out[i] == id) return true;
    if (cnt >= kMaxNbrs) return false;
    out[cnt++] = id;
    return true;
  };
  const int noff = dim == 2 ? 9 : 27;
  for (int o = 0; o < noff; ++o) {
    int nb[3] = {0, 0, 0};
    int oo = o;
    bool inside = true;
    for (int a = dim - 1; a >= 0; --a) {
      nb[a] = ib[a] + (oo % 3) - 1;
      oo /= 3;
      inside = inside && nb[a] >= 0 && nb[a] < span;
    }
    if (!inside) continue;
    int id = find_leaf(keys, ids, n, morton_key(dim, lv, nb));
    if (id >= 0) {
      if (!add(id)) return -1;
      continue;
    }
    if (lv > 0) {
      int pb[3] = {nb[0] >> 1, nb[1] >> 1, nb[2] >> 1};
      id = find_leaf(keys, ids, n, morton_key(dim, lv - 1, pb));
      if (id >= 0) {
        if (!add(id)) return -1;
        continue;
      }
    }
    if (lv >= lmax) return -1;  // cell covered by nothing
    // finer neighbours: children of cell nb touching leaf b (closed boxes
    // [c, c+1] at level lv+1 and [2 ib, 2 ib + 2])
    for (int ch = 0; ch < (1 << dim); ++ch) {
      int c[3] = {0, 0, 0};
      bool touch = true;
      for (int a = 0; a < dim; ++a) {
        c[a] = 2 * nb[a] + ((ch >> (dim - 1 - a)) & 1);
        touch = touch && (c[a] + 1 >= 2 * ib[a]) && (c[a] <= 2 * ib[a] + 2);
      }
      if (!touch) continue;
      id = find_leaf(keys, ids, n, morton_key(dim, lv + 1, c));
      if (id < 0 || !add(id)) return -1;  // not 2:1 balanced
    }
  }
  return cnt;
}

__global__ void csr_count_kernel(int dim, int n, int lmax, const uint8_t* levels,
                                 const int32_t* index, const uint64_t* keys, const int32_t* ids,
                                 int32_t* counts) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  int32_t buf[kMaxNbrs];
  counts[b] = leaf_neighbours(dim, b, levels, index, keys, ids, n, lmax, buf);
}

__global__ void csr_fill_kernel(int dim, int n, int lmax, const uint8_t* levels,
                                const int32_t* index, const uint64_t* keys, const int32_t* ids,
                                const int64_t* starts, int32_t* lists) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  int32_t buf[kMaxNbrs];
  const int cnt = leaf_neighbours(dim, b, levels, index, keys, ids, n, lmax, buf);
  for (int i = 1; i < cnt; ++i) {  // insertion sort by leaf id
    const int32_t v = buf[i];
    int j = i - 1;
    while (j >= 0 && buf[j] > v) {
      buf[j + 1] = buf[j];
      --j;
    }
    buf[j + 1] = v;
  }
  for (int i = 0; i < cnt; ++i) lists[starts[b] + i] = buf[i];
}

}  // namespace vl1

using namespace vl1;

struct vl1_csr_builder {
  int dim = 0, n = 0, lmax = 0;
  int64_t n_pairs = 0;
  DevBuf keys, ids, levels, index, starts64;
  std::vector<int32_t> starts32;
};

extern "C" int vl1_csr_builder_create(vl1_csr_builder** out, int dim, int n_leaves,
                                      const uint8_t* levels, const int32_t* index,
                                      int64_t* n_pairs, void* stream) {
  clear_error();
  if (!out || (dim != 2 && dim != 3) || n_leaves < 1 || !levels || !index || !n_pairs)
    return fail(VL1_E_ARG, "vl1_csr_builder_create: bad arguments");
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  std::vector<std::pair<uint64_t, int32_t>> kv(n_leaves);
  int lmax = 0;
  for (int b = 0; b < n_leaves; ++b) {
    const int lv = levels[b];
    if (lv > 18) return fail(VL1_E_ARG, "vl1_csr_builder_create: level above 18");
    for (int a = 0; a < dim; ++a) {
      const int v = index[(int64_t)b * dim + a];
      if (v < 0 || v >= (1 << lv)) return fail(VL1_E_ARG, "vl1_csr_builder_create: index out of range");
    }
    lmax = std::max(lmax, lv);
    kv[b] = {morton_key(dim, lv, index + (int64_t)b * dim), b};
  }
  std::sort(kv.begin(), kv.end());
  for (int b = 1; b < n_leaves; ++b)
    if (kv[b].first == kv[b - 1].first)
      return fail(VL1_E_LAYOUT, "vl1_csr_builder_create: duplicate leaf");
  std::vector<uint64_t> keys(n_leaves);
  std::vector<int32_t> ids(n_leaves);
  for (int b = 0; b < n_leaves; ++b) {
    keys[b] = kv[b].first;
    ids[b] = kv[b].second;
  }
  auto bl = new vl1_csr_builder();
  bl->dim = dim;
  bl->n = n_leaves;
  bl->lmax = lmax;
  DevBuf dcnt;
  hipError_t e = upload(bl->keys, keys.data(), keys.size() * 8, s);
  if (e == hipSuccess) e = upload(bl->ids, ids.data(), ids.size() * 4, s);
  if (e == hipSuccess) e = upload(bl->levels, levels, (size_t)n_leaves, s);
  if (e == hipSuccess) e = upload(bl->index, index, (size_t)n_leaves * dim * 4, s);
  if (e == hipSuccess) e = dcnt.alloc((size_t)n_leaves * 4);
  const int block = 128, grid = (n_leaves + block - 1) / block;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(csr_count_kernel, dim3(grid), dim3(block), 0, s, dim, n_leaves, lmax,
                       bl->levels.as<uint8_t>(), bl->index.as<int32_t>(), bl->keys.as<uint64_t>(),
                       bl->ids.as<int32_t>(), dcnt.as<int32_t>());
    e = hipGetLastError();
  }
  std::vector<int32_t> cnt(n_leaves);
  if (e == hipSuccess)
    e = hipMemcpyAsync(cnt.data(), dcnt.ptr, cnt.size() * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete bl;
    return fail(VL1_E_HIP, std::string("vl1_csr_builder_create: ") + hipGetErrorString(e));
  }
  std::vector<int64_t> st64(n_leaves + 1, 0);
  for (int b = 0; b < n_leaves; ++b) {
    if (cnt[b] < 0) {
      delete bl;
      return fail(VL1_E_LAYOUT, "vl1_csr_builder_create: the leaves are not a 2:1-balanced tiling");
    }
    st64[b + 1] = st64[b] + cnt[b];
  }
  if (st64[n_leaves] > INT32_MAX) {
    delete bl;
    return fail(VL1_E_UNSUPPORTED, "vl1_csr_builder_create: too many pairs");
  }
  bl->starts32.resize(n_leaves + 1);
  for (int b = 0; b <= n_leaves; ++b) bl->starts32[b] = (int32_t)st64[b];
  e = upload(bl->starts64, st64.data(), st64.size() * 8, s);
  if (e != hipSuccess) {
    delete bl;
    return fail(VL1_E_HIP, std::string("vl1_csr_builder_create: ") + hipGetErrorString(e));
  }
  bl->n_pairs = st64[n_leaves];
  *n_pairs = bl->n_pairs;
  *out = bl;
  return VL1_OK;
}

extern "C" int vl1_csr_builder_fill(vl1_csr_builder* bl, int32_t* starts, int32_t* lists,
                                    void* stream) {
  clear_error();
  if (!bl || !starts || (!lists && bl->n_pairs > 0))
    return fail(VL1_E_ARG, "vl1_csr_builder_fill: null argument");
  hipStream_t s = (hipStream_t)stream;
  const int block = 128, grid = (bl->n + block - 1) / block;
  hipLaunchKernelGGL(csr_fill_kernel, dim3(grid), dim3(block), 0, s, bl->dim, bl->n, bl->lmax,
                     bl->levels.as<uint8_t>(), bl->index.as<int32_t>(), bl->keys.as<uint64_t>(),
                     bl->ids.as<int32_t>(), bl->starts64.as<int64_t>(), lists);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(starts, bl->starts32.data(), bl->starts32.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(VL1_E_HIP, std::string("vl1_csr_builder_fill: ") + hipGetErrorString(e));
  return VL1_OK;
}

extern "C" int vl1_csr_builder_destroy(vl1_csr_builder* bl) {
  delete bl;
  return VL1_OK;
}
