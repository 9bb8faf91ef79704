// Connected components of a graph held as CSR, edges taken in both directions.
//
// Reference: sparse/csgraph/_components.pyx:86-177 (find_components_sparse,
// an iterative DFS over CSR and CSC) and :15-81 (find_components, the dense
// crawl).  Both number components by their smallest vertex, in increasing
// order.  Here every vertex carries a label that only decreases: a hooking
// pass sets, for every edge whose endpoints sit in different trees, the
// larger root's label to the smaller root (atomicMin), and a compression
// pass points every vertex at its root.  Rounds repeat until no edge hooks.
// A component's smallest vertex is never hooked (every label in it is at
// least that vertex), so it ends as the root: labels are the components'
// smallest vertices, which rank into the reference's component ids.
#include "rbg_common.hpp"

namespace rbg {

static __global__ void k_cc_init(int32_t* __restrict__ lab, int32_t n, int32_t* flag) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    lab[v] = (int32_t)v;
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

static __device__ inline int32_t cc_root(const int32_t* lab, int32_t x) {
  // labels strictly decrease along the chain until a root (lab[r] == r)
  int32_t y = __atomic_load_n(&lab[x], __ATOMIC_RELAXED);
  while (y != x) {
    x = y;
    y = __atomic_load_n(&lab[x], __ATOMIC_RELAXED);
  }
  return x;
}

// One wave per vertex v, a lane per edge (v, w): the roots of both ends, and
// the larger one hooked under the smaller.  v's root is looked up again for
// every 64 edges; a root that another hook has just lowered only hooks a
// vertex of the same component, and the rounds repeat until no edge hooks.
static __global__ __launch_bounds__(256) void k_cc_hook(const int64_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices, int32_t n, int32_t* lab,
                                                        int32_t* flag) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * 4;
  bool changed = false;
  for (int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); v < n; v += waves) {
    const int64_t e1 = indptr[v + 1];
    for (int64_t e = indptr[v] + lane; e - lane < e1; e += 64) {
      const int32_t a = cc_root(lab, (int32_t)v);
      if (e < e1) {
        const int32_t b = cc_root(lab, indices[e]);
        if (a != b) {
          const int32_t hi = a > b ? a : b, lo = a > b ? b : a;
          atomicMin(&lab[hi], lo);
          changed = true;
        }
      }
    }
  }
  if (__ballot(changed) && lane == 0) atomicOr(flag, 1);
}

static __global__ void k_cc_compress(int32_t* lab, int32_t n, int32_t* flag_next) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    lab[v] = cc_root(lab, (int32_t)v);
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag_next = 0;
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int rbg_components(const int64_t* indptr, const int32_t* indices, int32_t n, int32_t* labels, int32_t* flags,
                   int32_t* rounds, rbg_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || (n > 0 && (!indptr || !labels || !flags))) return RBG_E_ARG;
  if (rounds) *rounds = 0;
  if (n == 0) return 0;
  const int grid = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  const int hook_grid = (int)((n + 3) / 4 < 8192 ? (n + 3) / 4 : 8192);  // a wave per vertex
  hipLaunchKernelGGL(k_cc_init, dim3(grid), dim3(256), 0, s, labels, n, flags);
  RBG_CHECK_LAUNCH();
  // flags[0] / flags[1] alternate as "some edge hooked in this round"
  // every round that hooks lowers some label, so rounds are bounded by n;
  // graphs of family trees need a handful
  for (int32_t r = 0; r <= n; r++) {
    int32_t* cur = flags + (r & 1);
    int32_t* nxt = flags + ((r & 1) ^ 1);
    hipLaunchKernelGGL(k_cc_hook, dim3(hook_grid), dim3(256), 0, s, indptr, indices, n, labels, cur);
    RBG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_cc_compress, dim3(grid), dim3(256), 0, s, labels, n, nxt);
    RBG_CHECK_LAUNCH();
    int32_t h = 0;
    hipError_t e = hipMemcpyAsync(&h, cur, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return (int)e;
    if (rounds) *rounds = r + 1;
    if (!h) return 0;
  }
  return RBG_E_ARG;  // not reached
}

}  // extern "C"
