// Sparse AVOS closure of a DAG, level-synchronous on the GPU.
//
// Reference: redblackgraph/sparse/csgraph/_transitive_closure_dag.pyx:168-351
// (closure row of v = direct edges of v (+) A[v][w] (x) C[w] over successors
// w != v, rows built sinks first) and redblackgraph/gpu/transitive_closure.py:
// 474-652 (level-by-level expansion).  Here:
//   * levels: a vertex joins level r once every non-self successor sits in a
//     level < r ("pull" rounds over the shrinking set of unplaced vertices);
//     a round that places nothing while vertices remain means a cycle;
//   * one wavefront closes one vertex: the direct edges and each successor's
//     (already sorted) closure row, mapped through the AVOS product, are
//     folded together with a merge-path merge in LDS, duplicates combined
//     with the AVOS sum and zero products dropped; rows land sorted;
//   * row storage comes from a bump-allocated arena (atomicAdd), so a level
//     needs no host round trip; the final CSR is gathered at the end.
#include "rbg_common.hpp"

namespace rbg {

constexpr int DAG_WAVES = 4;          // waves per workgroup in the merge kernel
constexpr int DAG_CAP = 512;          // LDS entries per buffer per wave
constexpr int SCAN_NT = 256;

// ---------------------------------------------------------------------------
// exclusive scan of int64 (3 kernels: block sums, scan of sums, add)
// ---------------------------------------------------------------------------
__global__ void k_scan_blocks(const int64_t* in, int64_t* out, int64_t n, int64_t* block_sums, int per_thread) {
  __shared__ int64_t s[SCAN_NT];
  const int tid = threadIdx.x;
  const int64_t base = ((int64_t)blockIdx.x * SCAN_NT + tid) * per_thread;
  int64_t local = 0;
  for (int q = 0; q < per_thread; q++) {
    int64_t e = base + q;
    if (e < n) local += in[e];
  }
  s[tid] = local;
  __syncthreads();
  for (int off = 1; off < SCAN_NT; off <<= 1) {
    int64_t v = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += v;
    __syncthreads();
  }
  int64_t run = s[tid] - local;  // exclusive prefix of this thread within the block
  for (int q = 0; q < per_thread; q++) {
    int64_t e = base + q;
    if (e < n) {
      int64_t v = in[e];
      out[e] = run;
      run += v;
    }
  }
  if (tid == SCAN_NT - 1) block_sums[blockIdx.x] = s[tid];
}

__global__ void k_scan_sums(int64_t* block_sums, int nblocks, int64_t* total) {
  // single workgroup, sequential chunks
  __shared__ int64_t s[SCAN_NT];
  const int tid = threadIdx.x;
  int64_t carry = 0;
  for (int c0 = 0; c0 < nblocks; c0 += SCAN_NT) {
    int idx = c0 + tid;
    int64_t v = idx < nblocks ? block_sums[idx] : 0;
    s[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_NT; off <<= 1) {
      int64_t t = tid >= off ? s[tid - off] : 0;
      __syncthreads();
      s[tid] += t;
      __syncthreads();
    }
    if (idx < nblocks) block_sums[idx] = carry + s[tid] - v;
    carry += s[SCAN_NT - 1];
    __syncthreads();
  }
  if (tid == 0 && total) *total = carry;
}

__global__ void k_scan_add(int64_t* out, int64_t n, const int64_t* block_sums, int per_thread) {
  const int64_t base = ((int64_t)blockIdx.x * SCAN_NT + threadIdx.x) * per_thread;
  const int64_t add = block_sums[blockIdx.x];
  for (int q = 0; q < per_thread; q++) {
    int64_t e = base + q;
    if (e < n) out[e] += add;
  }
}

static int scan_blocks_for(int64_t n, int per_thread) {
  return (int)((n + (int64_t)SCAN_NT * per_thread - 1) / ((int64_t)SCAN_NT * per_thread));
}

// out[0..n) = exclusive scan of in[0..n); *total = sum.  tmp holds >= nblocks int64.
static int scan_exclusive(const int64_t* in, int64_t* out, int64_t n, int64_t* tmp, int64_t* total, hipStream_t s) {
  const int per = 8;
  int nb = scan_blocks_for(n, per);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_NT), 0, s, in, out, n, tmp, per);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_NT), 0, s, tmp, nb, total);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_NT), 0, s, out, n, (const int64_t*)tmp, per);
  RBG_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// exclusive scan with a device-side element count: block-local scans plus a
// scan of the block sums; a consumer reads element i as
// local[i] + sums[i / SCAN_CHUNK].
// ---------------------------------------------------------------------------
constexpr int SCAN_PER = 8;
constexpr int64_t SCAN_CHUNK = (int64_t)SCAN_NT * SCAN_PER;

// The scanned length is *n_ptr - *n_sub (n_sub may be null) plus `extra`
// trailing zeros, at most max_n.
__device__ inline int64_t scan_len(const int64_t* n_ptr, const int64_t* n_sub) {
  return *n_ptr - (n_sub ? *n_sub : 0);
}

__global__ void k_scan_local(const int64_t* __restrict__ in, int64_t* __restrict__ local, const int64_t* n_ptr,
                             const int64_t* n_sub, int64_t extra, int64_t max_n, int64_t* __restrict__ sums) {
  __shared__ int64_t sh[SCAN_NT];
  const int64_t nr = scan_len(n_ptr, n_sub);
  int64_t n = nr + extra;
  n = n < max_n ? n : max_n;
  const int tid = threadIdx.x;
  for (int64_t blk = blockIdx.x; blk * SCAN_CHUNK < n; blk += gridDim.x) {
    const int64_t base = blk * SCAN_CHUNK + (int64_t)tid * SCAN_PER;
    int64_t v[SCAN_PER];
    int64_t acc = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
      v[q] = (base + q < nr) ? in[base + q] : 0;
      acc += v[q];
    }
    sh[tid] = acc;
    __syncthreads();
    for (int off = 1; off < SCAN_NT; off <<= 1) {
      int64_t t = tid >= off ? sh[tid - off] : 0;
      __syncthreads();
      sh[tid] += t;
      __syncthreads();
    }
    int64_t run = sh[tid] - acc;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
      if (base + q < n) local[base + q] = run;
      run += v[q];
    }
    if (tid == SCAN_NT - 1) sums[blk] = sh[tid];
    __syncthreads();
  }
}

// scan of the block sums in place (single workgroup); *total = sum of all
__global__ void k_scan_blocksums(int64_t* sums, const int64_t* n_ptr, const int64_t* n_sub, int64_t extra,
                                 int64_t max_n, int64_t* total) {
  __shared__ int64_t sh[SCAN_NT];
  const int tid = threadIdx.x;
  int64_t n = scan_len(n_ptr, n_sub) + extra;
  n = n < max_n ? n : max_n;
  const int64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += SCAN_NT) {
    int64_t idx = c0 + tid;
    int64_t v = idx < nb ? sums[idx] : 0;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_NT; off <<= 1) {
      int64_t t = tid >= off ? sh[tid - off] : 0;
      __syncthreads();
      sh[tid] += t;
      __syncthreads();
    }
    if (idx < nb) sums[idx] = carry + sh[tid] - v;
    carry += sh[SCAN_NT - 1];
    __syncthreads();
  }
  if (tid == 0 && total) *total = carry;
}

__device__ inline int64_t scan_at(const int64_t* local, const int64_t* sums, int64_t i) {
  return local[i] + sums[i / SCAN_CHUNK];
}

static int scan_dev(const int64_t* in, int64_t* local, int64_t* sums, const int64_t* n_ptr, const int64_t* n_sub,
                    int64_t extra, int64_t max_n, int64_t* total, hipStream_t s) {
  int64_t nb = (max_n + extra + SCAN_CHUNK - 1) / SCAN_CHUNK;
  int grid = (int)(nb < 1 ? 1 : (nb > 4096 ? 4096 : nb));
  hipLaunchKernelGGL(k_scan_local, dim3(grid), dim3(SCAN_NT), 0, s, in, local, n_ptr, n_sub, extra, max_n + extra,
                     sums);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_blocksums, dim3(1), dim3(SCAN_NT), 0, s, sums, n_ptr, n_sub, extra, max_n + extra,
                     total);
  RBG_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------
enum {
  C_PLACED = 0,   // vertices placed so far
  C_UNUSED1 = 1,
  C_UNUSED2 = 2,
  C_LSTART = 3,   // first index in `order` of the current level (its size: C_PLACED - C_LSTART)
  C_LPREV = 4,    // first index in `order` of the previous level
  C_M = 5,        // candidate entries of the current level
  C_LONG = 6,     // rows of the current level left to the long-row step
  C_ARENA = 7,    // arena entries used
  C_CMAX = 8,     // max |candidate|
  C_FMAX = 9,     // max |final value| > 1
  C_STUCK = 10,   // a round placed nothing while vertices remain (cycle)
  C_CAP = 11,     // a capacity was exceeded
  C_NLEV = 12,    // levels so far
  C_NNZ = 13,     // total closure entries (gather)
  C_DONE = 14,    // every vertex placed
  C_ROUND = 15,   // round counter
  C_CYCV = 16,    // a vertex on a cycle
  C_TEMP = 17,    // temp arena entries used (SpGEMM long rows)
  C_HINT = 18,    // a level did not match its hint (the rounds stop; the caller reruns without hints)
  C_NCTR = 32,
};

struct DagWork {
  int32_t* level;      // [n] level of each vertex (INT32_MAX = unplaced)
  int32_t* order;      // [n] vertices grouped by level, sinks first
  int32_t* rem[2];     // [n] unplaced successors of each vertex; predecessor fill cursors
  int64_t* row_off;    // [n] arena offset of each closure row
  int64_t* row_len;    // [n+1] closure row lengths
  int64_t* cnt;        // [n+1] candidate counts of the current level's vertices
  int64_t* cand_loc;   // [n+1] block-local scan of cnt
  int64_t* sums_a;     // scan block sums (vertex scans)
  int64_t* counters;   // [C_NCTR]
  // the current level's vertices by slot (written by the Kahn round): vertex,
  // first entry, direct entries, lists; and for rows with at most SLOT_D
  // direct entries, each non-self successor's arena row, length and edge value
  int32_t* slot_v;     // [n]
  int64_t* slot_e0;    // [n]
  int32_t* slot_nd;    // [n]
  int32_t* slot_nl;    // [n]
  int64_t* slot_ro;    // [n * SLOT_D]
  int32_t* slot_len;   // [n * SLOT_D]
  int64_t* slot_av;    // [n * SLOT_D] (edge values, stored as int64)
  int32_t* long_list;  // [n] slots of the current level left to the long-row step
  uint8_t* kinds;      // [n] per level: LK_UNITS (merged by k_lvl_units) | LK_LONG (it had long rows)
  int64_t* pptr;       // [n+1] predecessor lists (the reversed non-self edges) ...
  int32_t* preds;      // [ecap] ... carved from the workspace's tail by the closure
  int64_t ecap;
  static constexpr int SLOT_D = 7;
  static int64_t bytes(int64_t n) {
    int64_t b = 4 * n * 4 + 8 * n + 8 * (n + 1) * 4 + 8 * (n / SCAN_CHUNK + 8) + 8 * C_NCTR;
    b += n * (4 + 8 + 4 + 4 + 4 + 1) + n * SLOT_D * (8 + 4 + 8);
    return b + 26 * 256;
  }
  // with room for the predecessor lists of a matrix with nnz entries
  static int64_t bytes(int64_t n, int64_t nnz) { return bytes(n) + 4 * (nnz > 0 ? nnz : 0) + 256; }
  static DagWork carve(void* base, int64_t n) {
    DagWork w;
    char* p = (char*)(((uintptr_t)base + 255) & ~(uintptr_t)255);
    auto take = [&](int64_t bytes) {
      char* r = p;
      p += (bytes + 255) & ~(int64_t)255;
      return r;
    };
    w.level = (int32_t*)take(4 * n);
    w.order = (int32_t*)take(4 * n);
    w.rem[0] = (int32_t*)take(4 * n);
    w.rem[1] = (int32_t*)take(4 * n);
    w.row_off = (int64_t*)take(8 * n);
    w.row_len = (int64_t*)take(8 * (n + 1));
    w.cnt = (int64_t*)take(8 * (n + 1));
    w.cand_loc = (int64_t*)take(8 * (n + 1));
    w.sums_a = (int64_t*)take(8 * (n / SCAN_CHUNK + 8));
    w.counters = (int64_t*)take(8 * C_NCTR);
    w.slot_v = (int32_t*)take(4 * n);
    w.slot_e0 = (int64_t*)take(8 * n);
    w.slot_nd = (int32_t*)take(4 * n);
    w.slot_nl = (int32_t*)take(4 * n);
    w.slot_ro = (int64_t*)take(8 * n * SLOT_D);
    w.slot_len = (int32_t*)take(4 * n * SLOT_D);
    w.slot_av = (int64_t*)take(8 * n * SLOT_D);
    w.long_list = (int32_t*)take(4 * n);
    w.kinds = (uint8_t*)take(n);
    w.pptr = (int64_t*)take(8 * (n + 1));
    w.preds = (int32_t*)p;
    w.ecap = 0;
    return w;
  }
};

__global__ void k_dag_init(DagWork ws, int32_t n) {
  int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v < n) {
    ws.level[v] = INT32_MAX;
    ws.rem[1][v] = 0;
    ws.row_len[v] = 0;
    ws.cnt[v] = 0;
  }
  if (v == 0) {
    for (int q = 0; q < C_NCTR; q++) ws.counters[q] = 0;
    ws.counters[C_CYCV] = -1;
  }
}

__device__ inline bool dag_halted(const int64_t* c) { return c[C_STUCK] | c[C_CAP] | c[C_DONE] | c[C_HINT]; }

// Levels by Kahn's method on the reversed edges: level 0 is every vertex
// without non-self successors, and a vertex joins level r + 1 when its last
// unplaced non-self successor is placed in level r, so its level is the
// longest path to a sink.  Placement offsets in `order` come from one atomic
// per workgroup (a workgroup-local count first).
__device__ inline void place_ready(DagWork& ws, int32_t v, bool ready, int32_t r) {
  __shared__ int s_n;
  __shared__ int64_t s_base;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const int slot = ready ? atomicAdd(&s_n, 1) : 0;
  __syncthreads();
  if (threadIdx.x == 0 && s_n) s_base = (int64_t)atomicAdd((unsigned long long*)&ws.counters[C_PLACED], (unsigned long long)s_n);
  __syncthreads();
  if (ready) {
    ws.level[v] = r;
    ws.order[s_base + slot] = v;
  }
  __syncthreads();
}

// Setup, step 1: each vertex's unplaced-successor count, each vertex's
// predecessor count, and level 0.
__global__ __launch_bounds__(256) void k_kahn_count(const int64_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices, DagWork ws, int32_t n) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + threadIdx.x;
    int32_t c = 0;
    if (v < n) {
      for (int64_t e = indptr[v]; e < indptr[v + 1]; e++) {
        const int32_t w = indices[e];
        if (w != v) {
          c++;
          atomicAdd((unsigned long long*)&ws.cnt[w], 1ull);
        }
      }
      ws.rem[0][v] = c;
    }
    place_ready(ws, (int32_t)v, v < n && c == 0, 0);
  }
}

// Setup, step 2 (after the predecessor counts are scanned into pptr): the
// predecessor lists.
__global__ __launch_bounds__(256) void k_kahn_fill(const int64_t* __restrict__ indptr,
                                                   const int32_t* __restrict__ indices, DagWork ws, int32_t n) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t e = indptr[v]; e < indptr[v + 1]; e++) {
      const int32_t w = indices[e];
      if (w != v) {
        const int64_t q = ws.pptr[w] + atomicAdd(&ws.rem[1][w], 1);
        if (q < ws.ecap) ws.preds[q] = (int32_t)v;
        else ws.counters[C_CAP] = 1;  // a workspace sized without the matrix's nnz
      }
    }
  }
}

// Vertex v's slot t of the current level: its candidate count (its direct
// entries plus the closure rows of its non-self successors, all placed in
// earlier levels) and the slot data the merge step reads.
template <typename T>
__device__ inline void level_slot(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                  const T* __restrict__ data, const DagWork& ws, int64_t t, int32_t v) {
  const int64_t e0 = indptr[v], e1 = indptr[v + 1];
  const bool keep = e1 - e0 <= DagWork::SLOT_D;
  int64_t c = 0;
  int k = 0;
  for (int64_t e = e0; e < e1; e++) {
    const int32_t w = indices[e];
    if (w == v) {
      c += 1;
      continue;
    }
    const int64_t len = ws.row_len[w];
    c += 1 + len;
    if (keep) {
      ws.slot_ro[t * DagWork::SLOT_D + k] = ws.row_off[w];
      ws.slot_len[t * DagWork::SLOT_D + k] = (int32_t)len;
      ws.slot_av[t * DagWork::SLOT_D + k] = (int64_t)data[e];
    }
    k++;
  }
  ws.cnt[t] = c;
  ws.slot_v[t] = v;
  ws.slot_e0[t] = e0;
  ws.slot_nd[t] = (int32_t)(e1 - e0);
  ws.slot_nl[t] = 1 + k;
}

// Round r, step 1: level r and its slot data.  Round 0 only records the slot
// data of level 0 (placed with the predecessor counts).  Round r >= 1 builds
// level r from level r - 1 (the slots [C_LPREV, C_LSTART) of `order`): each
// vertex of level r - 1 releases its predecessors, and a predecessor whose
// count reaches zero joins level r.  A workgroup takes 256 vertices of level
// r - 1 and walks their predecessor lists as one flattened range (a prefix of
// the list lengths in LDS), so every thread issues about (edges / 256)
// decrements however the lists are spread; the vertices it frees are queued
// in LDS, placed with one atomic, and given their slot data.
// SLOTS = false (the topological sort) places the vertices only.
constexpr int KR_NT = 256;
constexpr int KR_QCAP = 4096;
template <typename T, bool SLOTS = true>
__global__ __launch_bounds__(KR_NT) void k_kahn_round(const int64_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ indices, const T* __restrict__ data,
                                                      DagWork ws, int32_t r) {
  int64_t* ctr = ws.counters;
  if (dag_halted(ctr)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t L0 = ctr[C_LSTART];
  if (r == 0) {
    const int64_t Ls = ctr[C_PLACED] - L0;
    if (SLOTS)
      for (int64_t t = (int64_t)blockIdx.x * KR_NT + tid; t < Ls; t += (int64_t)gridDim.x * KR_NT)
        level_slot<T>(indptr, indices, data, ws, t, ws.order[L0 + t]);
    return;
  }
  const int64_t f0 = ctr[C_LPREV], f1 = L0;
  __shared__ int64_t s_q0[KR_NT];
  __shared__ int64_t s_pre[KR_NT + 1];
  __shared__ int64_t s_wsum[KR_NT / 64];
  __shared__ int32_t s_q[KR_QCAP];
  __shared__ int s_n;
  __shared__ int64_t s_base;
  for (int64_t base = f0 + (int64_t)blockIdx.x * KR_NT; base < f1; base += (int64_t)gridDim.x * KR_NT) {
    const int64_t i = base + tid;
    int64_t q0 = 0, cnt = 0;
    if (i < f1) {
      const int32_t u = ws.order[i];
      q0 = ws.pptr[u];
      cnt = ws.pptr[u + 1] - q0;
    }
    int64_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wsum[w] = incl;
    if (tid == 0) s_n = 0;
    __syncthreads();
    int64_t off = 0, total = 0;
#pragma unroll
    for (int k = 0; k < KR_NT / 64; k++) {
      off += k < w ? s_wsum[k] : 0;
      total += s_wsum[k];
    }
    s_pre[tid] = off + incl - cnt;
    s_q0[tid] = q0;
    if (tid == 0) s_pre[KR_NT] = total;
    __syncthreads();
    int k = 0;
    for (int64_t e = tid; e < total; e += KR_NT) {
      // owner of flattened edge e: the last k with s_pre[k] <= e (owners of
      // empty lists share their successor's prefix and are passed over)
      int lo = k, hi = KR_NT;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_pre[mid] <= e) lo = mid;
        else hi = mid;
      }
      k = lo;
      const int32_t p = ws.preds[s_q0[k] + (e - s_pre[k])];
      if (atomicSub(&ws.rem[0][p], 1) == 1) {
        ws.level[p] = r;
        const int slot = atomicAdd(&s_n, 1);
        if (slot < KR_QCAP) {
          s_q[slot] = p;
        } else {
          const int64_t g = (int64_t)atomicAdd((unsigned long long*)&ctr[C_PLACED], 1ull);
          ws.order[g] = p;
          if (SLOTS) level_slot<T>(indptr, indices, data, ws, g - L0, p);
        }
      }
    }
    __syncthreads();
    const int m = s_n < KR_QCAP ? s_n : KR_QCAP;
    if (tid == 0 && m) s_base = (int64_t)atomicAdd((unsigned long long*)&ctr[C_PLACED], (unsigned long long)m);
    __syncthreads();
    for (int j = tid; j < m; j += KR_NT) {
      const int32_t p = s_q[j];
      ws.order[s_base + j] = p;
      if (SLOTS) level_slot<T>(indptr, indices, data, ws, s_base + j - L0, p);
    }
    __syncthreads();  // s_n, s_pre and s_q are reused by the next pass
  }
}

// upper bound (first index with value > x) / lower bound over a sorted column list
__device__ inline int64_t ub_col(const int32_t* c, int64_t n, int32_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (c[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ inline int64_t lb_col(const int32_t* c, int64_t n, int32_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (c[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename T>
struct DagArena {
  int32_t* cols;
  T* vals;
  int64_t cap;
};

constexpr int LV_WAVES = 4;                    // waves per workgroup in the level kernel
constexpr int LV_NT = 64 * LV_WAVES;
constexpr int LV_SMALL = 512;                  // candidates a wave merges in its LDS share
constexpr int LV_BIG = LV_WAVES * LV_SMALL;    // candidates the workgroup merges in LDS

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// One closure row.  Candidate lists of v: list 0 = its direct entries, list k
// = A[v][w_k] (x) C[w_k] for its k-th non-self successor w_k, each sorted by
// column without repeats.  `lcol`, when non-null, holds the lists' columns
// concatenated (staged in LDS); otherwise they are read from the arena.
template <typename T>
struct RowJob {
  int32_t v;
  int64_t e0, e1, nd, cnt, dst;
};

template <typename T>
__device__ inline void stage_lists(const RowJob<T>& J, const int32_t* __restrict__ indices, const DagWork& ws,
                                   const DagArena<T>& ar, int32_t* lcol, int lane, int nl) {
  for (int64_t q = lane; q < J.nd; q += nl) lcol[q] = indices[J.e0 + q];
  int64_t off = J.nd;
  for (int64_t e = J.e0; e < J.e1; e++) {
    const int32_t w = indices[e];
    if (w == J.v) continue;
    const int64_t lw = ws.row_len[w], ro = ws.row_off[w];
    for (int64_t q = lane; q < lw; q += nl) lcol[off + q] = ar.cols[ro + q];
    off += lw;
  }
}

// Place every candidate: its position in the merged order is its index in its
// own list plus, for every other list, the number of entries that come before
// it (ties go to the lower list).  A candidate whose column appears in a lower
// list is a follower (written as column -1); a run head takes the AVOS sum of
// the values every list holds for its column, and a zero sum is dropped too.
template <typename T>
__device__ inline void merge_lists(const RowJob<T>& J, const int32_t* __restrict__ indices,
                                   const T* __restrict__ data, const DagWork& ws, const DagArena<T>& ar,
                                   const int32_t* lcol, int32_t* mcol, T* mval, int lane, int nl, int width,
                                   T& cmax, uint32_t& evs) {
  const int32_t* l0 = lcol ? lcol : indices + J.e0;
  for (int64_t idx = lane; idx < J.cnt; idx += nl) {
    int li = 0;
    int64_t p = idx;
    int32_t c = 0;
    T x = 0;
    if (idx < J.nd) {
      c = l0[idx];
      x = data[J.e0 + idx];
    } else {
      int64_t local = idx - J.nd, off = J.nd;
      int k = 0;
      for (int64_t e = J.e0; e < J.e1; e++) {
        const int32_t w = indices[e];
        if (w == J.v) continue;
        k++;
        const int64_t lw = ws.row_len[w];
        if (local < lw) {
          const int64_t o = ws.row_off[w] + local;
          li = k;
          p = local;
          c = lcol ? lcol[off + local] : ar.cols[o];
          uint32_t ev = 0;
          x = avos_product<T>(data[e], ar.vals[o], width, false, ev);
          evs |= ev;
          break;
        }
        local -= lw;
        off += lw;
      }
    }
    {
      const T a = x < 0 ? (T)-x : x;
      cmax = a > cmax ? a : cmax;
    }
    int64_t pos = p;
    bool head = true;
    T comb = x;
    if (li != 0) {
      const int64_t r0 = ub_col(l0, J.nd, c);
      pos += r0;
      if (r0 > 0 && l0[r0 - 1] == c) head = false;
    }
    int64_t off = J.nd;
    int k = 0;
    for (int64_t e = J.e0; e < J.e1; e++) {
      const int32_t w = indices[e];
      if (w == J.v) continue;
      k++;
      const int64_t lw = ws.row_len[w];
      if (k != li) {
        const int32_t* wc = lcol ? lcol + off : ar.cols + ws.row_off[w];
        if (k < li) {
          const int64_t rk = ub_col(wc, lw, c);
          pos += rk;
          if (rk > 0 && wc[rk - 1] == c) head = false;
        } else {
          const int64_t rk = lb_col(wc, lw, c);
          pos += rk;
          if (head && rk < lw && wc[rk] == c) {
            uint32_t ev = 0;
            const T y = avos_product<T>(data[e], ar.vals[ws.row_off[w] + rk], width, false, ev);
            evs |= ev;
            comb = avos_sum<T>(comb, y);
          }
        }
      }
      off += lw;
    }
    mcol[pos] = (head && comb != 0) ? c : -1;
    mval[pos] = comb;
  }
}

// Per-row list metadata for the LDS merge: list 0 is the vertex's own entries
// (columns in `indices`, values in `data` from e0), list k >= 1 is
// A[v][w_k] (x) C[w_k] for its k-th non-self successor (arena row at ro[k],
// multiplier av[k]); lst[k] is the list's first candidate index (lst[nl] =
// the row's candidate count).
constexpr int LV_MAXD = DagWork::SLOT_D;  // direct entries a row may have for the LDS paths
template <typename T>
struct RowMeta {
  int32_t nl;
  int32_t lst[LV_MAXD + 2];
  int64_t ro[LV_MAXD + 1];
  T av[LV_MAXD + 1];
};

constexpr int LV_R = LV_SMALL / 64;  // candidates per thread on both LDS paths
static_assert(LV_BIG / LV_NT == LV_R, "a wave's and the workgroup's LDS paths hold as many candidates per thread");

// Merge one row entirely in LDS: `lane` of NL cooperating threads (a wave, or
// the workgroup), list metadata in M, merged order to mcol/mval[base ...].
// One round of global loads stages every candidate: its column to LDS, its
// value (A[v][w_k] (x) C[w_k][c] for lists k >= 1) to a register of the
// thread that owns it (candidates lane + NL r).  A candidate's place in the
// merged order is the number of entries before it over all lists: in its own
// list its index, in a lower list the entries with column <= c, in a higher
// list those with column < c.  Equal columns therefore sit next to each other
// in list order, and the compaction sums each such run (lds_run_head).
template <typename T, int NL>
__device__ inline void merge_row_lds(const int32_t* __restrict__ indices, const T* __restrict__ data,
                                     const DagArena<T>& ar, const RowMeta<T>& M, int64_t e0, int cnt, int base,
                                     int32_t* s_lcol, int32_t* s_mcol, T* s_mval, int lane, int width, T& cmax,
                                     uint32_t& evs) {
  const int nl = M.nl;
  int lst[LV_MAXD + 2];
#pragma unroll
  for (int q = 0; q < LV_MAXD + 2; q++) lst[q] = q <= nl ? M.lst[q] : cnt;
  T x[LV_R];
  int32_t cr[LV_R];
#pragma unroll
  for (int r = 0; r < LV_R; r++) {
    const int idx = lane + NL * r;
    x[r] = 0;
    cr[r] = 0;
    if (idx < cnt) {
      int k = 0;
#pragma unroll
      for (int q = 1; q <= LV_MAXD; q++) k += (q < nl && idx >= lst[q]) ? 1 : 0;
      const int p = idx - lst[k];
      int32_t c;
      T v;
      if (k == 0) {
        c = indices[e0 + p];
        v = data[e0 + p];
      } else {
        const int64_t o = M.ro[k] + p;
        c = ar.cols[o];
        uint32_t ev = 0;
        v = avos_product<T>(M.av[k], ar.vals[o], width, false, ev);
        evs |= ev;
      }
      s_lcol[base + idx] = c;
      x[r] = v;
      cr[r] = c;
      const T a = v < 0 ? (T)-v : v;
      cmax = a > cmax ? a : cmax;
    }
  }
  if (NL == 64) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
  for (int r = 0; r < LV_R; r++) {
    const int idx = lane + NL * r;
    if (idx >= cnt) break;
    const int32_t c = cr[r];
    int pos = 0;
    for (int k2 = 0; k2 < nl; k2++) {
      const int lo = base + lst[k2];
      const int hi = lst[k2 + 1];
      if (idx >= lst[k2] && idx < hi) {  // its own list: its index
        pos += idx - lst[k2];
        continue;
      }
      // entries before it: column <= c in a lower list, < c in a higher one
      const int32_t key = idx >= hi ? c + 1 : c;
      int a = 0, b = hi - lst[k2];
      while (a < b) {
        const int mid = (a + b) >> 1;
        if (s_lcol[lo + mid] < key) a = mid + 1;
        else b = mid;
      }
      pos += a;
    }
    s_mcol[base + pos] = c;
    s_mval[base + pos] = x[r];
  }
}

// Merged position i of an LDS-merged row: whether it starts a run of equal
// columns, and the AVOS sum of the run's values (the closure entry).
template <typename T>
__device__ inline bool lds_run_head(const int32_t* mcol, const T* mval, int i, int cnt, int32_t& c, T& sum) {
  c = mcol[i];
  if (i > 0 && mcol[i - 1] == c) return false;
  sum = mval[i];
  for (int j = i + 1; j < cnt && mcol[j] == c; j++) sum = avos_sum<T>(sum, mval[j]);
  return true;
}

// A row's slot data in registers (wave-uniform), and in lane k < SLOT_D the
// arena row, length and edge value of its list k + 1.
struct SlotRegs {
  int32_t v, nd, nl, len;
  int64_t e0, seg, cnt, ro, av;
};

__device__ inline int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline int64_t uni(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Issues a row's slot loads (t is wave-uniform; S.cnt holds the end of the
// row's candidate range until settle_slot).
__device__ inline void load_slot(const DagWork& ws, int64_t t, int64_t Ls, int lane, SlotRegs& S) {
  S.cnt = -1;
  S.seg = 0;
  if (t >= Ls) return;
  S.v = ws.slot_v[t];
  S.e0 = ws.slot_e0[t];
  S.nd = ws.slot_nd[t];
  S.nl = ws.slot_nl[t];
  S.seg = scan_at(ws.cand_loc, ws.sums_a, t);
  S.cnt = scan_at(ws.cand_loc, ws.sums_a, t + 1);
  if (lane < DagWork::SLOT_D) {
    const int64_t q = t * DagWork::SLOT_D + lane;
    S.ro = ws.slot_ro[q];
    S.len = ws.slot_len[q];
    S.av = ws.slot_av[q];
  }
}

// Once a row's slot loads are needed: its scalars as wave-uniform values, so
// that the row's loops branch on scalar conditions.
__device__ inline void settle_slot(SlotRegs& S) {
  if (S.cnt < 0) return;
  S.v = uni(S.v);
  S.e0 = uni(S.e0);
  S.nd = uni(S.nd);
  S.nl = uni(S.nl);
  S.seg = uni(S.seg);
  S.cnt = uni(S.cnt) - S.seg;
}


// The list metadata of a row that one wave merges (lanes 0..SLOT_D-1 hold its
// lists; list starts are a prefix sum over those lanes).
template <typename T>
__device__ inline void put_meta(const SlotRegs& S, RowMeta<T>& M, int lane) {
  const int len = lane + 1 < S.nl ? S.len : 0;
  int incl = len;
#pragma unroll
  for (int o = 1; o < 8; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  if (lane + 1 < S.nl) {
    M.ro[lane + 1] = S.ro;
    M.av[lane + 1] = (T)S.av;
    M.lst[lane + 2] = S.nd + incl;
  }
  if (lane == 0) {
    M.nl = S.nl;
    M.lst[0] = 0;
    M.lst[1] = S.nd;
  }
}

// Round r, step 3 (after the scan of the candidate counts): the level's rows
// merged by single waves, several at a time.  A unit is a run of consecutive slots (rows) whose candidates (at most
// LU_CAND, in at most LU_LISTS lists) one wave stages, merges and compacts
// together.  Chunks of C consecutive slots are the work items, C from the
// level's mean candidate count; a chunk that does not fit one unit is split
// into greedy sub-units of consecutive rows.  Rows with more than LU_CAND
// candidates or more than LV_MAXD direct entries go to the long-row list.
//
// In a unit, a candidate of row j's list g is placed at: the candidates of the
// unit's rows before j, plus its index in g, plus for every other list of row
// j the number of entries before it (column <= c in a lower list, < c in a
// higher one).  Equal columns of a row sit next to each other in list order;
// the compaction keeps each run's head with the run's AVOS sum (zero sums
// dropped) and writes the unit's rows one after another from the unit's first
// reserved arena entry; each row lands at or below its own reserved range.
constexpr int LK_UNITS = 1;  // level kinds: merged by k_lvl_units (else k_lvl_pairs)
constexpr int LK_LONG = 2;   // the level had rows for k_lvl_long
constexpr int LU_GRID = 4096;        // waves of k_lvl_pairs
constexpr int LU_UNITS_GRID = 2048;  // waves of k_lvl_units
constexpr int LU_ROWS_MEAN = 64;  // levels with a larger mean candidate count use k_lvl_pairs
constexpr int LU_CAND = 512;
constexpr int LU_ROWS = 64;
constexpr int LU_LISTS = 128;
static_assert(LU_CAND / 64 == LV_R, "candidates per lane");

// Lane j's value of x, read by every lane (j uniform).
__device__ __forceinline__ int32_t lane_i32(int32_t x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ __forceinline__ int64_t lane_i64(int64_t x, int j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)x, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)((uint64_t)x >> 32), j);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Appends the slots of the lanes with `mine` set to the long-row list: one
// atomic for the wave, so that a level with many long rows does not queue
// one same-address atomic per row.
__device__ __forceinline__ void push_long(const DagWork& ws, bool mine, int64_t t, int lane) {
  const uint64_t m = __ballot(mine);
  if (!m) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd((unsigned long long*)&ws.counters[C_LONG], (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (mine) ws.long_list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)t;
}

template <typename T>
struct UnitLds {
  int32_t lcol[LU_CAND];          // staged columns, list after list
  int32_t mcol[LU_CAND];          // merged columns
  T mval[LU_CAND];                // merged values (run heads: the run's sum)
  int32_t pref[LU_CAND + 1];      // kept entries before each merged position
  uint8_t mrow[LU_CAND];          // merged position -> row
  int32_t lstart[LU_LISTS + 1];   // list starts in the unit; lstart[nlists] = candidates
  int64_t lsrc[LU_LISTS];         // a row's own entries: its first entry in A; else the arena row
  T lav[LU_LISTS];                // edge value multiplying an arena list
  uint8_t lrow[LU_LISTS];         // list -> row
  uint8_t lown[LU_LISTS];         // list holds the row's own entries
  int32_t rlist[LU_ROWS + 1];     // row -> first list
  int32_t rbase[LU_ROWS + 1];     // row -> first candidate; rbase[rows] = candidates
  int32_t rv[LU_ROWS];            // row -> vertex
};

// Wave-wide exclusive prefix sum (64 lanes) and total.
__device__ inline int wave_excl_scan(int x, int lane, int& total) {
  int incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  total = __shfl(incl, 63);
  return incl - x;
}

// Merge the unit's rows j < m; lane j < m holds row j's slot t, its scalars
// (v, e0, nd, nl, candidate count) and, through slot_* arrays, its lists.  A
// row with nl = 0 is a place holder (a row left to the long-row step): no
// lists, no candidates, no output.  U and NL: the unit's candidates and lists
// (uniform); dst0: the unit's first reserved arena entry.
template <typename T>
__device__ inline void merge_unit(const int32_t* __restrict__ indices, const T* __restrict__ data,
                                  const DagWork& ws, const DagArena<T>& ar, UnitLds<T>& L, int lane, int m,
                                  int64_t t, int32_t v, int64_t e0, int32_t nd, int32_t nl, int64_t cntj, int U,
                                  int NL, int64_t dst0, int width, T& cmax, T& fmax, uint32_t& evs) {
  auto wsync = []() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
  };
  // row and list tables
  {
    int tot;
    const int rb = wave_excl_scan(lane < m ? (int)cntj : 0, lane, tot);
    const int rl = wave_excl_scan(lane < m ? nl : 0, lane, tot);
    if (lane < m) {
      L.rbase[lane] = rb;
      L.rlist[lane] = rl;
      L.rv[lane] = v;
    }
    if (lane < m && nl > 0) {
      L.lstart[rl] = rb;
      L.lsrc[rl] = e0;
      L.lav[rl] = 0;
      L.lrow[rl] = (uint8_t)lane;
      L.lown[rl] = 1;
      int64_t ro[DagWork::SLOT_D], av[DagWork::SLOT_D];
      int32_t len[DagWork::SLOT_D];
#pragma unroll
      for (int k = 0; k < DagWork::SLOT_D; k++) {
        if (k + 1 < nl) {
          const int64_t q = t * DagWork::SLOT_D + k;
          ro[k] = ws.slot_ro[q];
          len[k] = ws.slot_len[q];
          av[k] = ws.slot_av[q];
        }
      }
      int start = rb + nd;
#pragma unroll
      for (int k = 0; k < DagWork::SLOT_D; k++) {
        if (k + 1 < nl) {
          const int g = rl + 1 + k;
          L.lstart[g] = start;
          L.lsrc[g] = ro[k];
          L.lav[g] = (T)av[k];
          L.lrow[g] = (uint8_t)lane;
          L.lown[g] = 0;
          start += len[k];
        }
      }
    }
    if (lane == 0) {
      L.rbase[m] = U;
      L.rlist[m] = NL;
      L.lstart[NL] = U;
    }
  }
  wsync();
  // stage: columns to LDS, values to registers
  int32_t c[LV_R];
  T x[LV_R];
  int gq[LV_R];
#pragma unroll
  for (int r = 0; r < LV_R; r++) {
    if (64 * r >= U) break;
    const int u = lane + 64 * r;
    gq[r] = 0;
    if (u < U) {
      int lo = 0, hi = NL;  // the last list start <= u
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (L.lstart[mid] <= u) lo = mid;
        else hi = mid;
      }
      gq[r] = lo;
      const int64_t o = L.lsrc[lo] + (u - L.lstart[lo]);
      if (L.lown[lo]) {
        c[r] = indices[o];
        x[r] = data[o];
      } else {
        c[r] = ar.cols[o];
        x[r] = ar.vals[o];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < LV_R; r++) {
    if (64 * r >= U) break;
    const int u = lane + 64 * r;
    if (u < U) {
      const int g = gq[r];
      if (!L.lown[g]) {
        uint32_t ev = 0;
        x[r] = avos_product<T>(L.lav[g], x[r], width, false, ev);
        evs |= ev;
      }
      L.lcol[u] = c[r];
      const T a = x[r] < 0 ? (T)-x[r] : x[r];
      cmax = a > cmax ? a : cmax;
    }
  }
  wsync();
  // place
  for (int r = 0; r < LV_R; r++) {
    if (64 * r >= U) break;
    const int u = lane + 64 * r;
    if (u >= U) continue;
    const int g = gq[r];
    const int j = L.lrow[g];
    const int32_t cc = c[r];
    int pos = L.rbase[j] + (u - L.lstart[g]);
    const int g1 = L.rlist[j + 1];
    for (int g2 = L.rlist[j]; g2 < g1; g2++) {
      if (g2 == g) continue;
      const int lo = L.lstart[g2];
      const int32_t key = g2 < g ? cc + 1 : cc;  // entries before it: <= c below, < c above
      int a = 0, b = L.lstart[g2 + 1] - lo;
      while (a < b) {
        const int mid = (a + b) >> 1;
        if (L.lcol[lo + mid] < key) a = mid + 1;
        else b = mid;
      }
      pos += a;
    }
    L.mcol[pos] = cc;
    L.mval[pos] = x[r];
    L.mrow[pos] = (uint8_t)j;
  }
  wsync();
  // compaction: run heads and their sums, then the kept-entry prefix
  int run = 0;
  for (int b0 = 0; b0 < U; b0 += 64) {
    const int i = b0 + lane;
    bool keep = false;
    if (i < U) {
      const int32_t cc = L.mcol[i];
      const int j = L.mrow[i];
      const int end = L.rbase[j + 1];
      if (i == L.rbase[j] || L.mcol[i - 1] != cc) {
        T sum = L.mval[i];
        for (int i2 = i + 1; i2 < end && L.mcol[i2] == cc; i2++) sum = avos_sum<T>(sum, L.mval[i2]);
        keep = sum != 0;
        if (keep) L.mval[i] = sum;
      }
    }
    const uint64_t msk = __ballot(keep);
    if (i < U) L.pref[i] = run + __popcll(msk & ((1ull << lane) - 1ull));
    run += __popcll(msk);
  }
  if (lane == 0) L.pref[U] = run;
  wsync();
  for (int b0 = 0; b0 < U; b0 += 64) {
    const int i = b0 + lane;
    if (i < U && L.pref[i + 1] != L.pref[i]) {
      const int64_t o = dst0 + L.pref[i];
      const T val = L.mval[i];
      ar.cols[o] = L.mcol[i];
      ar.vals[o] = val;
      const T a = val < 0 ? (T)-val : val;
      if (a > 1 && a > fmax) fmax = a;
    }
  }
  if (lane < m && nl > 0) {
    const int p0 = L.pref[L.rbase[lane]], p1 = L.pref[L.rbase[lane + 1]];
    ws.row_off[L.rv[lane]] = dst0 + p0;
    ws.row_len[L.rv[lane]] = p1 - p0;
  }
  wsync();  // the tables are rewritten by the next unit
}

template <typename T>
__global__ __launch_bounds__(64) void k_lvl_units(const int32_t* __restrict__ indices, const T* __restrict__ data,
                                                  DagWork ws, DagArena<T> ar, int width, rbg_status* st) {
  __shared__ UnitLds<T> L;
  int64_t* ctr = ws.counters;
  if (dag_halted(ctr)) return;
  const int64_t Ls = ctr[C_PLACED] - ctr[C_LSTART], M = ctr[C_M], abase = ctr[C_ARENA];
  if (abase + M > ar.cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[C_CAP] = 1;
    return;
  }
  const int lane = threadIdx.x;
  const T seen_c = (T)ctr[C_CMAX], seen_f = (T)ctr[C_FMAX];  // lower bounds of the values at exit
  T cmax = 0, fmax = 0;
  uint32_t evs = 0;
  // rows per chunk: the largest power of two (at most LU_ROWS) whose mean
  // candidate count is at most tgt / 2.  A unit's fixed costs (its tables,
  // its dependent loads, the wait for its stores) dominate levels of short
  // rows, so their chunks are larger.
  const int64_t mean = Ls > 0 ? (M + Ls - 1) / Ls : 1;
  if (mean > LU_ROWS_MEAN) return;  // a level of long rows: k_lvl_pairs takes it
  const int64_t tgt = mean <= 16 ? 1024 : (mean <= 32 ? 512 : 256);
  int C = 1;
  while (C < LU_ROWS && (int64_t)(2 * C) * mean <= tgt) C <<= 1;
  const int64_t nch = (Ls + C - 1) / C;
  for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    const int64_t t0 = ch * C;
    const int m = (int)(Ls - t0 < C ? Ls - t0 : C);
    // lane j < m: row t0 + j
    const int64_t t = t0 + (lane < m ? lane : 0);
    const int32_t v = ws.slot_v[t], nd = ws.slot_nd[t], nl = ws.slot_nl[t];
    const int64_t e0 = ws.slot_e0[t];
    const int64_t seg = scan_at(ws.cand_loc, ws.sums_a, t);
    const int64_t cntj = scan_at(ws.cand_loc, ws.sums_a, t + 1) - seg;
    // rows no unit takes go to the long-row list
    const bool lng = lane < m && (nd > LV_MAXD || cntj > LU_CAND);
    push_long(ws, lng, t, lane);
    // greedy sub-units of consecutive rows within LU_CAND candidates and
    // LU_LISTS lists (a long row counts as none)
    const int cw = lane < m && !lng ? (int)cntj : 0, lw = lane < m && !lng ? nl : 0;
    int totc, totl;
    const int pc = wave_excl_scan(cw, lane, totc) + cw;  // inclusive prefixes
    const int pl = wave_excl_scan(lw, lane, totl) + lw;
    int s0 = 0;
    while (s0 < m) {
      const int bc = s0 > 0 ? __shfl(pc, s0 - 1) : 0, bl = s0 > 0 ? __shfl(pl, s0 - 1) : 0;
      const uint64_t fit = __ballot(lane >= s0 && lane < m && pc - bc <= LU_CAND && pl - bl <= LU_LISTS);
      const int e = fit ? 63 - __builtin_clzll(fit) : s0;  // last row of the sub-unit (a row always fits)
      const int mu = e - s0 + 1;
      // rows s0..e to lanes 0..mu-1
      const int src = s0 + lane < 64 ? s0 + lane : 63;
      const int64_t tu = __shfl(t, src), e0u = __shfl(e0, src), cu = __shfl((int64_t)cw, src);
      const int32_t vu = __shfl(v, src), ndu = __shfl(nd, src), nlu = __shfl(lw, src);
      const int64_t segu = __shfl(seg, s0);
      merge_unit<T>(indices, data, ws, ar, L, lane, mu, tu, vu, e0u, ndu, nlu, cu, __shfl(pc, e) - bc,
                    __shfl(pl, e) - bl, abase + segu, width, cmax, fmax, evs);
      s0 = e + 1;
    }
  }
  const int64_t wc = wave_max_i64((int64_t)cmax), wf = wave_max_i64((int64_t)fmax);
  if (lane == 0) {
    if (wc > (int64_t)seen_c) atomicMax((long long*)&ctr[C_CMAX], (long long)wc);
    if (wf > (int64_t)seen_f) atomicMax((long long*)&ctr[C_FMAX], (long long)wf);
  }
  if (evs) atomicOr(&st->events, evs);
}

// The merge of a pair row (k_lvl_pairs): lists A = s_c/s_v[0, a_len) and B =
// [a_len, a_len + b_len) staged, the direct entries in lanes 0..nd-1 (l0c,
// l0v).  RQ bounds the merged positions per lane (R <= RQ), so that a short
// row runs loops unrolled for its own size.  Leaves the row in s_c / s_v and
// returns its length.
template <typename T, int RQ>
__device__ __forceinline__ int pair_merge(int32_t* s_c, T* s_v, int a_len, int b_len, int nd, int32_t l0c, T l0v,
                                          int lane, T& fmax) {
  auto wave_sync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  const int m = a_len + b_len, R = (m + 63) >> 6;
  const int32_t* Bc = s_c + a_len;
  const T* Bv = s_v + a_len;
  const int d = lane * R < m ? lane * R : m;
  int lo = d > b_len ? d - b_len : 0, hi = d < a_len ? d : a_len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s_c[mid] <= Bc[d - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  int i = lo, j = d - lo;
  int32_t ha = i < a_len ? s_c[i] : INT32_MAX, hb = j < b_len ? Bc[j] : INT32_MAX;
  const int nv = m - d < R ? m - d : R;  // this lane's merged positions
  int32_t oc[RQ];
  T ov[RQ];
#pragma unroll
  for (int q = 0; q < RQ; q++) {
    oc[q] = INT32_MAX;
    ov[q] = 0;
    if (q < nv) {
      if (ha <= hb) {
        oc[q] = ha;
        ov[q] = s_v[i];
        i++;
        ha = i < a_len ? s_c[i] : INT32_MAX;
      } else {
        oc[q] = hb;
        ov[q] = Bv[j];
        j++;
        hb = j < b_len ? Bc[j] : INT32_MAX;
      }
    }
  }
  // pairs of equal columns: the first takes the AVOS sum, the second drops
  int32_t lastc = INT32_MIN;
#pragma unroll
  for (int q = 0; q < RQ; q++)
    if (q < nv) lastc = oc[q];
  const int32_t prev0 = __shfl_up(lastc, 1);
  const int32_t next0c = __shfl_down(oc[0], 1);
  const T next0v = __shfl_down(ov[0], 1);
  bool keep[RQ];
#pragma unroll
  for (int q = 0; q < RQ; q++) {
    const int32_t pc = q == 0 ? (lane > 0 ? prev0 : INT32_MIN) : oc[q == 0 ? 0 : q - 1];
    const bool nx_in = q + 1 < nv;
    const int32_t nc = nx_in ? oc[q + 1 < RQ ? q + 1 : q] : (lane < 63 ? next0c : INT32_MAX);
    const T nvv = nx_in ? ov[q + 1 < RQ ? q + 1 : q] : next0v;
    keep[q] = q < nv && pc != oc[q];
    if (keep[q] && nc == oc[q]) ov[q] = avos_sum<T>(ov[q], nvv);
  }
  // direct entries: absorbed by a merged head of the same column, else kept
  bool l0keep = lane < nd;
  for (int k = 0; k < nd; k++) {
    const int32_t ck = __builtin_amdgcn_readlane(l0c, k);
    const T vk = __shfl(l0v, k);
    bool hit = false;
#pragma unroll
    for (int q = 0; q < RQ; q++)
      if (keep[q] && oc[q] == ck) {
        ov[q] = avos_sum<T>(vk, ov[q]);
        hit = true;
      }
    if (__ballot(hit) && lane == k) l0keep = false;
  }
#pragma unroll
  for (int q = 0; q < RQ; q++) keep[q] = keep[q] && ov[q] != 0;
  l0keep = l0keep && l0v != 0;
  const uint64_t l0mask = __ballot(l0keep);
  int mine = 0;
#pragma unroll
  for (int q = 0; q < RQ; q++) mine += keep[q] ? 1 : 0;
  int totS;
  const int baseS = wave_excl_scan(mine, lane, totS);
  wave_sync();  // the staged lists are read: the buffers take the merged row
  int run = baseS;
#pragma unroll
  for (int q = 0; q < RQ; q++) {
    if (keep[q]) {
      int pos = run;
      for (int k = 0; k < nd; k++)
        if (((l0mask >> k) & 1ull) && __builtin_amdgcn_readlane(l0c, k) < oc[q]) pos++;
      s_c[pos] = oc[q];
      s_v[pos] = ov[q];
      const T ax = ov[q] < 0 ? (T)-ov[q] : ov[q];
      if (ax > 1 && ax > fmax) fmax = ax;
      run++;
    }
  }
  for (int k = 0; k < nd; k++) {
    if (!((l0mask >> k) & 1ull)) continue;
    const int32_t ck = __builtin_amdgcn_readlane(l0c, k);
    int below = 0;
#pragma unroll
    for (int q = 0; q < RQ; q++) below += __popcll(__ballot(keep[q] && oc[q] < ck));
    if (lane == k) {
      const int pos = __popcll(l0mask & ((1ull << k) - 1ull)) + below;
      s_c[pos] = l0c;
      s_v[pos] = l0v;
      const T ax = l0v < 0 ? (T)-l0v : l0v;
      if (ax > 1 && ax > fmax) fmax = ax;
    }
  }
  return totS + __popcll(l0mask);
}

// Round r, step 3 for levels of long rows (mean candidate count above
// LU_ROWS_MEAN), the rows with one or two successor lists (at most LV_SMALL
// candidates, LV_MAXD direct entries): one row per wave, no state carried
// between rows, so that many waves fit a CU.  The two closure lists A and B
// (sorted columns, and their products with the row's edge values) are staged
// in LDS and merged by a merge path: lane j produces merged positions j R ..
// j R + R - 1 (R = (|A| + |B|) / 64 rounded up) after one binary search for
// its start on the merge diagonal.  A goes first on equal columns, so a column
// both lists hold sits in adjacent positions and the first of the pair takes
// the AVOS sum.  The direct entries (sorted, in lanes 0..nd-1) are matched
// against the merged heads and placed by counting (a direct entry goes before
// a merged candidate of the same column).  The row is written through LDS so
// that its stores are coalesced.  Other rows go to the long-row list.
template <typename T>
__global__ __launch_bounds__(64) void k_lvl_pairs(const int32_t* __restrict__ indices, const T* __restrict__ data,
                                                  DagWork ws, DagArena<T> ar, int width, rbg_status* st) {
  __shared__ int32_t s_c[LV_SMALL];
  __shared__ T s_v[LV_SMALL];
  int64_t* ctr = ws.counters;
  if (dag_halted(ctr)) return;
  const int64_t Ls = ctr[C_PLACED] - ctr[C_LSTART], M = ctr[C_M], abase = ctr[C_ARENA];
  if (abase + M > ar.cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[C_CAP] = 1;
    return;
  }
  if ((Ls > 0 ? (M + Ls - 1) / Ls : 1) <= LU_ROWS_MEAN) return;  // k_lvl_units takes the level
  const int lane = threadIdx.x;
  const T seen_c = (T)ctr[C_CMAX], seen_f = (T)ctr[C_FMAX];  // lower bounds of the values at exit
  T cmax = 0, fmax = 0;
  uint32_t evs = 0;
  auto wsync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // The wave's rows are t = blockIdx.x + G j.  Each batch of 64 of them is
  // read ahead, one row per lane, so that every short row's slot data is a
  // lane read when its turn comes and a row's first dependent loads are its
  // two lists.  The batch's long rows go to the long-row list after its short
  // rows: waves finish at different times, so their atomics do not queue.
  const int64_t G = gridDim.x;
  for (int64_t tb = blockIdx.x; tb < Ls; tb += 64 * G) {
    const int64_t tl = tb + G * lane;
    int32_t p_nd = 0, p_nl = 0, p_cnt = 0, p_v = 0, p_alen = 0, p_blen = 0;
    int64_t p_seg = 0, p_e0 = 0, p_ro1 = 0, p_ro2 = 0, p_av1 = 0, p_av2 = 0;
    bool lng = false;
    if (tl < Ls) {
      p_nd = ws.slot_nd[tl];
      p_nl = ws.slot_nl[tl];
      p_seg = scan_at(ws.cand_loc, ws.sums_a, tl);
      const int64_t c = scan_at(ws.cand_loc, ws.sums_a, tl + 1) - p_seg;
      lng = p_nl < 2 || p_nl > 3 || c > LV_SMALL || p_nd > LV_MAXD;
      p_cnt = (int32_t)(c < LV_SMALL ? c : LV_SMALL);
      p_v = ws.slot_v[tl];
      p_e0 = ws.slot_e0[tl];
      const int64_t q0 = tl * DagWork::SLOT_D;  // the slot data of a long row is not used
      p_ro1 = ws.slot_ro[q0];
      p_alen = ws.slot_len[q0];
      p_av1 = ws.slot_av[q0];
      if (p_nl > 2) {
        p_ro2 = ws.slot_ro[q0 + 1];
        p_blen = ws.slot_len[q0 + 1];
        p_av2 = ws.slot_av[q0 + 1];
      }
    }
    const uint64_t short_rows = __ballot(tl < Ls && !lng);
  for (int rj = 0; rj < 64; rj++) {
    if (!((short_rows >> rj) & 1)) continue;
    const int nd = lane_i32(p_nd, rj), nl = lane_i32(p_nl, rj);
    const int64_t seg = lane_i64(p_seg, rj);
    const int cnt = lane_i32(p_cnt, rj);
    const int32_t v = lane_i32(p_v, rj);
    const int64_t e0 = lane_i64(p_e0, rj);
    const int64_t ro1 = lane_i64(p_ro1, rj);
    const int a_len = lane_i32(p_alen, rj);
    const T av1 = (T)lane_i64(p_av1, rj);
    const int64_t ro2 = lane_i64(p_ro2, rj);
    const int b_len = lane_i32(p_blen, rj);
    const T av2 = (T)lane_i64(p_av2, rj);
    // stage both lists, each with its own (uniform) edge value
    for (int q = lane; q < a_len; q += 64) {
      uint32_t ev = 0;
      const T x = avos_product<T>(av1, ar.vals[ro1 + q], width, false, ev);
      evs |= ev;
      s_c[q] = ar.cols[ro1 + q];
      s_v[q] = x;
      const T ax = x < 0 ? (T)-x : x;
      cmax = ax > cmax ? ax : cmax;
    }
    for (int q = lane; q < b_len; q += 64) {
      uint32_t ev = 0;
      const T x = avos_product<T>(av2, ar.vals[ro2 + q], width, false, ev);
      evs |= ev;
      s_c[a_len + q] = ar.cols[ro2 + q];
      s_v[a_len + q] = x;
      const T ax = x < 0 ? (T)-x : x;
      cmax = ax > cmax ? ax : cmax;
    }
    int32_t l0c = INT32_MAX;
    T l0v = 0;
    if (lane < nd) {
      l0c = indices[e0 + lane];
      l0v = data[e0 + lane];
      const T ax = l0v < 0 ? (T)-l0v : l0v;
      cmax = ax > cmax ? ax : cmax;
    }
    wsync();
    const int R = (a_len + b_len + 63) >> 6;
    int out;
    if (R <= 1) out = pair_merge<T, 1>(s_c, s_v, a_len, b_len, nd, l0c, l0v, lane, fmax);
    else if (R <= 2) out = pair_merge<T, 2>(s_c, s_v, a_len, b_len, nd, l0c, l0v, lane, fmax);
    else if (R <= 4) out = pair_merge<T, 4>(s_c, s_v, a_len, b_len, nd, l0c, l0v, lane, fmax);
    else out = pair_merge<T, LV_R>(s_c, s_v, a_len, b_len, nd, l0c, l0v, lane, fmax);
    wsync();
    const int64_t dst = abase + seg;
    for (int q = lane; q < out; q += 64) {
      ar.cols[dst + q] = s_c[q];
      ar.vals[dst + q] = s_v[q];
    }
    if (lane == 0) {
      ws.row_off[v] = dst;
      ws.row_len[v] = out;
    }
    wsync();  // the next row restages the buffers
  }
    push_long(ws, lng, tl, lane);
  }
  const int64_t wc = wave_max_i64((int64_t)cmax), wf = wave_max_i64((int64_t)fmax);
  if (lane == 0) {
    if (wc > (int64_t)seen_c) atomicMax((long long*)&ctr[C_CMAX], (long long)wc);
    if (wf > (int64_t)seen_f) atomicMax((long long*)&ctr[C_FMAX], (long long)wf);
  }
  if (evs) atomicOr(&st->events, evs);
}

// Round r, step 4: the long rows step 3 left, one per workgroup at a time:
// merged in LDS when they fit (LV_BIG candidates, LV_MAXD direct entries),
// otherwise in their arena slot from the lists in global memory.
template <typename T>
__global__ __launch_bounds__(LV_NT) void k_lvl_long(const int32_t* __restrict__ indices, const T* __restrict__ data,
                                                    DagWork ws, DagArena<T> ar, int width, rbg_status* st) {
  __shared__ int32_t s_lcol[LV_BIG];
  __shared__ int32_t s_mcol[LV_BIG];
  __shared__ T s_mval[LV_BIG];
  __shared__ RowMeta<T> s_meta;
  __shared__ int s_wsum[LV_WAVES];
  __shared__ int64_t s_cmax, s_fmax;
  int64_t* ctr = ws.counters;
  if (dag_halted(ctr) || ctr[C_CAP]) return;
  const int64_t nlong = ctr[C_LONG], Ls = ctr[C_PLACED] - ctr[C_LSTART], abase = ctr[C_ARENA];
  if (nlong == 0) return;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) s_cmax = s_fmax = 0;
  T cmax = 0, fmax = 0;
  uint32_t evs = 0;
  for (int64_t i = blockIdx.x; i < nlong; i += gridDim.x) {
    const int64_t t = ws.long_list[i];
    SlotRegs S;
    load_slot(ws, t, Ls, lane, S);
    settle_slot(S);
    const int64_t cnt = S.cnt;
    const int64_t dst = abase + S.seg;
    const bool in_lds = cnt <= LV_BIG && S.nd <= LV_MAXD;
    if (in_lds && wv == 0) put_meta<T>(S, s_meta, lane);
    __syncthreads();
    int32_t* mcol = in_lds ? s_mcol : ar.cols + dst;
    T* mval = in_lds ? s_mval : ar.vals + dst;
    if (in_lds) {
      merge_row_lds<T, LV_NT>(indices, data, ar, s_meta, S.e0, (int)cnt, 0, s_lcol, s_mcol, s_mval,
                              threadIdx.x, width, cmax, evs);
    } else {
      RowJob<T> J;
      J.v = S.v;
      J.e0 = S.e0;
      J.e1 = S.e0 + S.nd;
      J.nd = S.nd;
      J.cnt = cnt;
      J.dst = dst;
      merge_lists<T>(J, indices, data, ws, ar, nullptr, mcol, mval, threadIdx.x, LV_NT, width, cmax, evs);
    }
    __syncthreads();
    int64_t out = 0;
    for (int64_t b0 = 0; b0 < cnt; b0 += LV_NT) {
      const int64_t i2 = b0 + threadIdx.x;
      int32_t c = -1;
      T val = 0;
      bool keep = false;
      if (i2 < cnt) {
        if (in_lds) {
          keep = lds_run_head<T>(s_mcol, s_mval, (int)i2, (int)cnt, c, val) && val != 0;
        } else {
          c = mcol[i2];
          val = mval[i2];
          keep = c >= 0;
        }
      }
      const uint64_t msk = __ballot(keep);
      if (lane == 0) s_wsum[wv] = __popcll(msk);
      __syncthreads();
      int64_t before = 0, total = 0;
      for (int q = 0; q < LV_WAVES; q++) {
        before += q < wv ? s_wsum[q] : 0;
        total += s_wsum[q];
      }
      if (keep) {
        const int64_t o = dst + out + before + __popcll(msk & ((1ull << lane) - 1ull));
        ar.cols[o] = c;
        ar.vals[o] = val;
        const T a = val < 0 ? (T)-val : val;
        if (a > 1 && a > fmax) fmax = a;
      }
      out += total;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      ws.row_off[S.v] = dst;
      ws.row_len[S.v] = out;
    }
  }
  if (cmax) atomicMax((long long*)&s_cmax, (long long)cmax);
  if (fmax) atomicMax((long long*)&s_fmax, (long long)fmax);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_cmax) atomicMax((long long*)&ctr[C_CMAX], (long long)s_cmax);
    if (s_fmax) atomicMax((long long*)&ctr[C_FMAX], (long long)s_fmax);
  }
  if (evs) atomicOr(&st->events, evs);
}

// Round r, step 5 (single lane): advance the arena, count the level, detect a
// cycle (nothing placed while vertices remain) and completion, and set up
// round r + 1.
__global__ void k_round_end(DagWork ws, int32_t r, int32_t n, int32_t hint) {
  int64_t* ctr = ws.counters;
  if (dag_halted(ctr)) return;
  const int64_t placed_now = ctr[C_PLACED] - ctr[C_LSTART];
  // the level's kind, recorded for the next closure of the same graph; a hint
  // that left out a merge kernel the level needed stops the rounds
  const int64_t mean = placed_now > 0 ? (ctr[C_M] + placed_now - 1) / placed_now : 1;
  const int kind = (mean <= LU_ROWS_MEAN ? LK_UNITS : 0) | (ctr[C_LONG] > 0 ? LK_LONG : 0);
  ws.kinds[r < n ? r : n - 1] = (uint8_t)kind;
  if (hint >= 0 && placed_now > 0 && ((kind & LK_UNITS) != (hint & LK_UNITS) || (kind & ~hint & LK_LONG))) {
    ctr[C_HINT] = 1;
    return;
  }
  ctr[C_ARENA] += ctr[C_M];  // the level's reserved slots
  if (placed_now > 0) ctr[C_NLEV] = r + 1;
  if (ctr[C_PLACED] == n) {
    ctr[C_DONE] = 1;
  } else if (placed_now == 0) {
    ctr[C_STUCK] = 1;  // nothing placed while vertices remain: a cycle
  }
  ctr[C_LPREV] = ctr[C_LSTART];
  ctr[C_LSTART] = ctr[C_PLACED];
  ctr[C_LONG] = 0;
  ctr[C_ROUND] = r + 1;
}

// A vertex on a cycle among the unplaced ones: start at the smallest unplaced
// vertex and keep following its first unplaced successor; the first vertex
// met twice lies on a cycle.  Single lane; launched after every closure and
// returns at once unless a round found no vertex to place.
__global__ void k_dag_cycle_vertex(const int64_t* indptr, const int32_t* indices, DagWork ws, int32_t n,
                                   int32_t* stamp, rbg_status* st) {
  if (!ws.counters[C_STUCK]) return;
  int32_t v = -1;
  for (int32_t u = 0; u < n; u++)
    if (ws.level[u] == INT32_MAX) {
      v = u;
      break;
    }
  if (v < 0) return;
  for (int32_t u = 0; u < n; u++) stamp[u] = 0;
  for (int64_t step = 0; step <= n; step++) {
    if (stamp[v]) break;
    stamp[v] = 1;
    int32_t nxt = -1;
    for (int64_t e = indptr[v]; e < indptr[v + 1]; e++) {
      int32_t w = indices[e];
      if (w != v && ws.level[w] == INT32_MAX) {
        nxt = w;
        break;
      }
    }
    if (nxt < 0) break;
    v = nxt;
  }
  st->aux[0] = v;
  atomicOr(&st->flags, (uint32_t)RBG_F_CYCLE);
}

// Round r of the topological sort, last step (single lane): cycle and
// completion tests, and the set-up of round r + 1.
__global__ void k_topo_round_end(DagWork ws, int32_t r, int32_t n) {
  int64_t* ctr = ws.counters;
  if (dag_halted(ctr)) return;
  const int64_t placed_now = ctr[C_PLACED] - ctr[C_LSTART];
  if (placed_now > 0) ctr[C_NLEV] = r + 1;
  if (ctr[C_PLACED] == n) ctr[C_DONE] = 1;
  else if (placed_now == 0) ctr[C_STUCK] = 1;
  ctr[C_LPREV] = ctr[C_LSTART];
  ctr[C_LSTART] = ctr[C_PLACED];
  ctr[C_ROUND] = r + 1;
}

// The sort's outputs: level[v], and the vertices by decreasing level (every
// vertex before its successors; `order` holds them by increasing level).
__global__ void k_topo_output(DagWork ws, int32_t n, int32_t* level_out, int32_t* order_out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || !ws.counters[C_DONE]) return;
  if (level_out) level_out[i] = ws.level[i];
  if (order_out) order_out[i] = ws.order[n - 1 - i];
}

// Entries below (aux[0]) and on (aux[1]) the diagonal of a CSR matrix.
__global__ void k_csr_triangles(const int64_t* __restrict__ indptr, const int32_t* __restrict__ idx, int32_t n,
                                rbg_status* st) {
  int64_t below = 0, diag = 0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    for (int64_t e = indptr[r]; e < indptr[r + 1]; e++) {
      below += idx[e] < r;
      diag += idx[e] == r;
    }
  for (int o = 32; o > 0; o >>= 1) {
    below += __shfl_xor(below, o);
    diag += __shfl_xor(diag, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (below) atomicAdd((unsigned long long*)&st->aux[0], (unsigned long long)below);
    if (diag) atomicAdd((unsigned long long*)&st->aux[1], (unsigned long long)diag);
  }
}

__global__ void k_dag_report(DagWork ws, rbg_status* st) {
  const int64_t* c = ws.counters;
  if (c[C_CAP]) atomicOr(&st->flags, (uint32_t)RBG_F_CAPACITY);
  st->aux[2] = c[C_CMAX];
  st->aux[3] = c[C_FMAX];
  st->aux[4] = c[C_ARENA];
  st->aux[6] = c[C_NLEV];
  st->aux[7] = c[C_DONE];
}

// ---------------------------------------------------------------------------
// wave-level helpers
// ---------------------------------------------------------------------------
__device__ inline int wave_lane() { return threadIdx.x & 63; }

// Make this wave's LDS / global writes visible to its other lanes before they
// read them (the merge buffers may live in either address space).

template <typename V>
__device__ inline V wave_bcast(V v, int src) {
  return __shfl(v, src, 64);
}

__device__ inline int64_t wave_sum64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Stable merge of two column-sorted lists (ties: list A first) into out.
template <typename T>
__device__ void wave_merge(const int32_t* ac, const T* av, int na, const int32_t* bc, const T* bv, int nb,
                           int32_t* oc, T* ov) {
  const int lane = wave_lane();
  const int total = na + nb;
  const int per = (total + 63) >> 6;
  int d0 = lane * per;
  if (d0 > total) d0 = total;
  int d1 = d0 + per;
  if (d1 > total) d1 = total;
  // merge-path split for diagonal d0: smallest i with i + j = d0 and A[i] > B[j-1]
  int lo = d0 > nb ? d0 - nb : 0, hi = d0 < na ? d0 : na;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    // take A[mid] before B[d0-1-mid]?  A wins ties.
    if (ac[mid] <= bc[d0 - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  int i = lo, j = d0 - lo;
  for (int d = d0; d < d1; d++) {
    bool takeA = (j >= nb) || (i < na && ac[i] <= bc[j]);
    if (takeA) {
      oc[d] = ac[i];
      ov[d] = av[i];
      i++;
    } else {
      oc[d] = bc[j];
      ov[d] = bv[j];
      j++;
    }
  }
}

// Combine adjacent equal columns with the AVOS sum, drop zero values, compact
// into dst.  Returns the new length (wave-uniform).
template <typename T>
__device__ int wave_dedup(const int32_t* sc, const T* sv, int n, int32_t* dc, T* dv) {
  const int lane = wave_lane();
  int out = 0;
  for (int base = 0; base < n; base += 64) {
    int e = base + lane;
    bool keep = false;
    int32_t c = 0;
    T v = 0;
    if (e < n) {
      c = sc[e];
      v = sv[e];
      bool first = (e == 0) || (sc[e - 1] != c);
      if (first) {
        int f = e + 1;
        while (f < n && sc[f] == c) {
          v = avos_sum<T>(v, sv[f]);
          f++;
        }
        keep = v != 0;
      }
    }
    uint64_t mask = __ballot(keep);
    int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (keep) {
      dc[out + before] = c;
      dv[out + before] = v;
    }
    out += __popcll(mask);
  }
  return out;
}

template <typename T>
struct RowArena {
  int32_t* cols;
  T* vals;
  int64_t cap;   // entries in the row arena
  int32_t* tcols;
  T* tvals;
  int64_t tcap;  // entries in the temp arena (rows longer than DAG_CAP)
};

// ---------------------------------------------------------------------------
// AVOS SpGEMM C = A (x) B, one wave per row of A: the row is the AVOS sum of
// A[i][k] (x) B[k][:] over the entries of A's row, folded one B row at a time
// with the same merge-path merge.  With mask_upper, columns < i are dropped
// (the j >= i mask of spgemm_upper_triangular, gpu/spgemm.py:122-149).
// Reference: sparse/sparsetools/rbm_math.h:164-222 (values),
//            gpu/spgemm_numeric.py (GPU driver).  Zero sums are dropped.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64 * DAG_WAVES) void k_spgemm_rows(
    const int64_t* __restrict__ ap, const int32_t* __restrict__ aj, const T* __restrict__ ax, int32_t m,
    const int64_t* __restrict__ bp, const int32_t* __restrict__ bj, const T* __restrict__ bx, int mask_upper,
    DagWork ws, RowArena<T> ar, int width, int is_uns, rbg_status* st) {
  __shared__ int32_t lc[DAG_WAVES][3][DAG_CAP];
  __shared__ T lv[DAG_WAVES][3][DAG_CAP];
  const int wv = threadIdx.x >> 6;
  const int lane = wave_lane();
  const int64_t i = (int64_t)blockIdx.x * DAG_WAVES + wv;
  if (i >= m) return;
  const int64_t e0 = ap[i], e1 = ap[i + 1];
  int64_t ub = 0;
  for (int64_t e = e0 + lane; e < e1; e += 64) {
    int32_t k = aj[e];
    ub += bp[k + 1] - bp[k];
  }
  ub = wave_sum64(ub);
  int64_t off = 0;
  if (lane == 0) off = (int64_t)atomicAdd((unsigned long long*)&ws.counters[C_ARENA], (unsigned long long)ub);
  off = wave_bcast(off, 0);
  if (off + ub > ar.cap) {
    if (lane == 0) {
      atomicOr(&st->flags, (uint32_t)RBG_F_CAPACITY);
      ws.row_len[i] = 0;
      ws.row_off[i] = 0;
    }
    return;
  }
  const bool big = ub > DAG_CAP;
  int32_t *c0p, *c1p, *c2p;
  T *v0p, *v1p, *v2p;
  if (!big) {
    c0p = lc[wv][0]; c1p = lc[wv][1]; c2p = lc[wv][2];
    v0p = lv[wv][0]; v1p = lv[wv][1]; v2p = lv[wv][2];
  } else {
    int64_t toff = 0;
    if (lane == 0) toff = (int64_t)atomicAdd((unsigned long long*)&ws.counters[C_TEMP], (unsigned long long)(3 * ub));
    toff = wave_bcast(toff, 0);
    if (toff + 3 * ub > ar.tcap) {
      if (lane == 0) {
        atomicOr(&st->flags, (uint32_t)RBG_F_CAPACITY);
        ws.row_len[i] = 0;
        ws.row_off[i] = 0;
      }
      return;
    }
    c0p = ar.tcols + toff; c1p = c0p + ub; c2p = c1p + ub;
    v0p = ar.tvals + toff; v1p = v0p + ub; v2p = v1p + ub;
  }
  int na = 0;
  uint32_t evs = 0;
  for (int64_t e = e0; e < e1; e++) {
    const int32_t k = aj[e];
    const T a = ax[e];
    const int64_t bo = bp[k];
    const int nb = (int)(bp[k + 1] - bo);
    int nw = 0;
    // stage a (x) B[k] (masked) into buffer 1
    for (int base = 0; base < nb; base += 64) {
      int q = base + lane;
      bool keep = false;
      int32_t c = 0;
      T p = 0;
      if (q < nb) {
        c = bj[bo + q];
        uint32_t ev = 0;
        p = avos_product<T>(a, bx[bo + q], width, is_uns != 0, ev);
        evs |= ev;
        keep = !(mask_upper && c < i);
      }
      uint64_t msk = __ballot(keep);
      int before = __popcll(msk & ((1ull << lane) - 1ull));
      if (keep) {
        c1p[nw + before] = c;
        v1p[nw + before] = p;
      }
      nw += __popcll(msk);
    }
    wave_sync();
    wave_merge<T>(c0p, v0p, na, c1p, v1p, nw, c2p, v2p);
    wave_sync();
    na = wave_dedup<T>(c2p, v2p, na + nw, c0p, v0p);
    wave_sync();
  }
  for (int q = lane; q < na; q += 64) {
    ar.cols[off + q] = c0p[q];
    ar.vals[off + q] = v0p[q];
  }
  if (lane == 0) {
    ws.row_off[i] = off;
    ws.row_len[i] = na;
  }
  uint32_t evall = evs;
  for (int o = 32; o > 0; o >>= 1) evall |= (uint32_t)__shfl_xor((int)evall, o, 64);
  if (lane == 0 && evall) atomicOr(&st->events, evall);
}

__global__ void k_rows_init(int64_t* row_len, int64_t* counters, int32_t n) {
  int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v < n) row_len[v] = 0;
  if (v == 0)
    for (int q = 0; q < C_NCTR; q++) counters[q] = 0;
}

// Segmented copy into CSR order: one wave moves 64 consecutive output rows,
// which are contiguous in the output from out_ptr[r0].  Lane j holds row
// r0 + j's source offset and length; the prefix of the lengths (in LDS) maps
// every output entry of the group to its row, so stores are fully coalesced
// and each lane keeps GC entries in flight.  src_of(r) / len_of(r) give a
// row's source offset and length.
constexpr int GC = 8;
template <typename T, typename SrcF, typename LenF>
__device__ inline void seg_copy_rows(int64_t r0, int32_t n, const int64_t* __restrict__ out_ptr, SrcF src_of,
                                     LenF len_of, const int32_t* __restrict__ in_idx, const T* __restrict__ in_val,
                                     int32_t* __restrict__ out_idx, T* __restrict__ out_val, int64_t* s_pre,
                                     int64_t* s_src) {
  const int lane = threadIdx.x & 63;
  const int64_t r = r0 + lane;
  int64_t len = 0, src = 0;
  if (r < n) {
    len = len_of(r);
    src = src_of(r);
  }
  int64_t incl = len;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  s_pre[lane + 1] = incl;
  if (lane == 0) s_pre[0] = 0;
  s_src[lane] = src;
  const int64_t total = __shfl(incl, 63);
  const int64_t dst0 = out_ptr[r0];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int k = 0;
  for (int64_t e0 = 0; e0 < total; e0 += 64 * GC) {
    int64_t sp[GC];
#pragma unroll
    for (int u = 0; u < GC; u++) {
      const int64_t e = e0 + u * 64 + lane;
      sp[u] = -1;
      if (e < total) {
        while (s_pre[k + 1] <= e) k++;
        sp[u] = s_src[k] + (e - s_pre[k]);
      }
    }
    int32_t c[GC];
    T v[GC];
#pragma unroll
    for (int u = 0; u < GC; u++)
      if (sp[u] >= 0) {
        c[u] = in_idx[sp[u]];
        v[u] = in_val[sp[u]];
      }
#pragma unroll
    for (int u = 0; u < GC; u++)
      if (sp[u] >= 0) {
        const int64_t e = e0 + u * 64 + lane;
        out_idx[dst0 + e] = c[u];
        out_val[dst0 + e] = v[u];
      }
  }
  __builtin_amdgcn_wave_barrier();  // s_pre / s_src are rewritten by the next group
}

constexpr int GATHER_NT = 256;
template <typename T>
__global__ __launch_bounds__(GATHER_NT) void k_dag_gather(const DagWork ws, const DagArena<T> ar,
                                                          const int64_t* __restrict__ out_ptr, int32_t n,
                                                          int32_t* __restrict__ out_idx, T* __restrict__ out_val,
                                                          int64_t out_cap, rbg_status* st) {
  __shared__ int64_t s_pre[GATHER_NT / 64][65];
  __shared__ int64_t s_src[GATHER_NT / 64][64];
  if (ws.counters[C_NNZ] > out_cap) {  // the output buffers are too small: nothing is written
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&st->flags, (uint32_t)RBG_F_CAPACITY);
    return;
  }
  const int w = threadIdx.x >> 6;
  const int64_t r0 = ((int64_t)blockIdx.x * (GATHER_NT / 64) + w) * 64;
  if (r0 >= n) return;
  seg_copy_rows<T>(
      r0, n, out_ptr, [&](int64_t r) { return ws.row_off[r]; }, [&](int64_t r) { return ws.row_len[r]; }, ar.cols,
      ar.vals, out_idx, out_val, s_pre[w], s_src[w]);
}

__global__ void k_dag_finish_status(const DagWork ws, rbg_status* st) {
  st->aux[1] = ws.counters[C_NNZ];    // nnz of the gathered CSR
  st->aux[2] = ws.counters[C_CMAX];   // max |candidate|
  st->aux[3] = ws.counters[C_FMAX];   // max |final value| > 1
  st->aux[4] = ws.counters[C_ARENA];  // arena entries used
  st->aux[5] = ws.counters[C_TEMP];   // temp arena entries used
}

__global__ void k_copy_i64(const int64_t* src, int64_t* dst, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_set_last(int64_t* p, int64_t idx, const int64_t* val) { p[idx] = *val; }

// ---------------------------------------------------------------------------
// CSR canonicalisation: sort each row by column, combine duplicate columns
// with the AVOS sum and drop explicit zeros (one lane per row; insertion sort
// in place).  Row lengths go to len[], then a scan + gather compacts.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_csr_sort_rows(const int64_t* __restrict__ indptr, int32_t* __restrict__ idx, T* __restrict__ val,
                                int32_t n, int64_t* __restrict__ len) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t a = indptr[r], b = indptr[r + 1];
  for (int64_t x = a + 1; x < b; x++) {
    int32_t c = idx[x];
    T v = val[x];
    int64_t y = x - 1;
    while (y >= a && idx[y] > c) {
      idx[y + 1] = idx[y];
      val[y + 1] = val[y];
      y--;
    }
    idx[y + 1] = c;
    val[y + 1] = v;
  }
  int64_t out = a;
  for (int64_t x = a; x < b;) {
    int32_t c = idx[x];
    T v = val[x];
    int64_t y = x + 1;
    while (y < b && idx[y] == c) {
      v = avos_sum<T>(v, val[y]);
      y++;
    }
    if (v != 0) {
      idx[out] = c;
      val[out] = v;
      out++;
    }
    x = y;
  }
  len[r] = out - a;
}

template <typename T>
__global__ __launch_bounds__(GATHER_NT) void k_csr_compact(const int64_t* __restrict__ indptr,
                                                           const int32_t* __restrict__ idx, const T* __restrict__ val,
                                                           const int64_t* __restrict__ len,
                                                           const int64_t* __restrict__ optr, int32_t n,
                                                           int32_t* __restrict__ oidx, T* __restrict__ oval) {
  __shared__ int64_t s_pre[GATHER_NT / 64][65];
  __shared__ int64_t s_src[GATHER_NT / 64][64];
  const int w = threadIdx.x >> 6;
  const int64_t r0 = ((int64_t)blockIdx.x * (GATHER_NT / 64) + w) * 64;
  if (r0 >= n) return;
  seg_copy_rows<T>(
      r0, n, optr, [&](int64_t r) { return indptr[r]; }, [&](int64_t r) { return len[r]; }, idx, val, oidx, oval,
      s_pre[w], s_src[w]);
}

// Rows that are not canonical (columns not strictly increasing, or an
// explicit zero), counted into st->aux[0] with one atomic per wave.
template <typename T>
__global__ void k_csr_check_canonical(const int64_t* __restrict__ indptr, const int32_t* __restrict__ idx,
                                      const T* __restrict__ val, int32_t n, rbg_status* st) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool bad = false;
  if (r < n) {
    const int64_t a = indptr[r], b = indptr[r + 1];
    for (int64_t x = a; x < b && !bad; x++) bad = val[x] == 0 || (x > a && idx[x] <= idx[x - 1]);
  }
  const uint64_t msk = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && msk) atomicAdd((unsigned long long*)&st->aux[0], (unsigned long long)__popcll(msk));
}

__global__ void k_csr_check(const int64_t* __restrict__ indptr, const int32_t* __restrict__ idx, int32_t n,
                            int32_t ncols, rbg_status* st) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t a = indptr[r], b = indptr[r + 1];
  bool bad = b < a;
  for (int64_t x = a; x < b && !bad; x++)
    if (idx[x] < 0 || idx[x] >= ncols) bad = true;
  if (bad) atomicOr(&st->flags, (uint32_t)RBG_F_CAPACITY);
}

// Element-wise sum of two canonical CSR matrices, one lane per row: a
// sequential merge of the two sorted rows into the concatenated slot
// (cat_ptr = a_ptr + b_ptr), equal columns combined by `combine`
// (0: AVOS sum, 1: smallest |value|, first operand on ties), zeros dropped.
template <typename T>
__global__ void k_csr_add_rows(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj, const T* __restrict__ ax,
                               const int64_t* __restrict__ bp, const int32_t* __restrict__ bj, const T* __restrict__ bx,
                               int32_t n, int combine, const int64_t* __restrict__ cp, int32_t* __restrict__ cj,
                               T* __restrict__ cx, int64_t* __restrict__ len) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t i = ap[r], ie = ap[r + 1], j = bp[r], je = bp[r + 1], o = cp[r], o0 = o;
  while (i < ie || j < je) {
    int32_t c;
    T v;
    if (j >= je || (i < ie && aj[i] < bj[j])) {
      c = aj[i];
      v = ax[i++];
    } else if (i >= ie || bj[j] < aj[i]) {
      c = bj[j];
      v = bx[j++];
    } else {
      c = aj[i];
      T x = ax[i++], y = bx[j++];
      if (combine == 0) {
        v = avos_sum<T>(x, y);
      } else {
        T ax_ = x < 0 ? -x : x, ay_ = y < 0 ? -y : y;
        v = ay_ < ax_ ? y : x;
      }
    }
    if (v != 0) {
      cj[o] = c;
      cx[o] = v;
      o++;
    }
  }
  len[r] = o - o0;
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int64_t rbg_dag_workspace_bytes(int32_t n, int64_t nnz) { return DagWork::bytes(n > 0 ? n : 1, nnz); }

#define RBG_DAG_ENTRY(SUF, T)                                                                                   \
  int rbg_dag_closure_##SUF(const int64_t* indptr, const int32_t* indices, const T* data, int32_t n,            \
                            void* workspace, int64_t ws_bytes, int32_t* arena_cols, T* arena_vals,              \
                            int64_t arena_cap, int32_t batch, const uint8_t* level_kinds, int32_t width,        \
                            rbg_status* st, rbg_stream_t stream) {                                              \
    hipStream_t s = (hipStream_t)stream;                                                                        \
    if (n <= 0 || !workspace || !arena_cols || !arena_vals || arena_cap < 0) return RBG_E_ARG;                  \
    if (ws_bytes < DagWork::bytes(n)) return RBG_E_WORKSPACE;                                                   \
    const bool exact = batch < 0; /* -batch rounds and no synchronisation */                                    \
    if (batch == 0) batch = 8;                                                                                  \
    DagWork ws = DagWork::carve(workspace, n);                                                                  \
    ws.ecap = ((char*)workspace + ws_bytes - (char*)ws.preds) / 4;                                              \
    DagArena<T> ar{arena_cols, arena_vals, arena_cap};                                                          \
    hipLaunchKernelGGL(k_dag_init, dim3((n + 255) / 256), dim3(256), 0, s, ws, n);                              \
    RBG_CHECK_LAUNCH();                                                                                         \
    const int gv = (int)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);                                      \
    const int gr = (int)(n < LU_GRID ? n : LU_GRID);                                                          \
    const int gu = (int)(n < LU_UNITS_GRID ? n : LU_UNITS_GRID);                                              \
    const int gl = (int)(n < 1024 ? n : 1024);                                                                \
    int64_t host[C_NCTR];                                                                                       \
    hipLaunchKernelGGL(k_kahn_count, dim3(gv), dim3(256), 0, s, indptr, indices, ws, n);                        \
    RBG_CHECK_LAUNCH();                                                                                         \
    {                                                                                                           \
      int rc = scan_exclusive(ws.cnt, ws.pptr, n, ws.sums_a, ws.counters + C_NNZ, s);                           \
      if (rc) return rc;                                                                                        \
      hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, ws.pptr, (int64_t)n,                               \
                         (const int64_t*)(ws.counters + C_NNZ));                                                \
      RBG_CHECK_LAUNCH();                                                                                       \
    }                                                                                                           \
    hipLaunchKernelGGL(k_kahn_fill, dim3(gv), dim3(256), 0, s, indptr, indices, ws, n);                         \
    RBG_CHECK_LAUNCH();                                                                                         \
    for (int32_t r = 0; r <= n; r++) {                                                                          \
      hipLaunchKernelGGL((k_kahn_round<T>), dim3(gv), dim3(KR_NT), 0, s, indptr, indices, data, ws, r);        \
      RBG_CHECK_LAUNCH();                                                                                       \
      int rc = scan_dev(ws.cnt, ws.cand_loc, ws.sums_a, ws.counters + C_PLACED, ws.counters + C_LSTART, 1, n,   \
                        ws.counters + C_M, s);                                                                  \
      if (rc) return rc;                                                                                        \
      /* with a hint (exact rounds only), the merge kernels the level needs */                                  \
      const int hint = exact && level_kinds ? (int)level_kinds[r] : -1;                                         \
      int pi = prof_start(s);                                                                                   \
      if (hint < 0 || (hint & LK_UNITS)) {                                                                      \
        hipLaunchKernelGGL((k_lvl_units<T>), dim3(gu), dim3(64), 0, s, indices, data, ws, ar, width, st);       \
        RBG_CHECK_LAUNCH();                                                                                     \
      }                                                                                                         \
      if (hint < 0 || !(hint & LK_UNITS)) {                                                                     \
        hipLaunchKernelGGL((k_lvl_pairs<T>), dim3(gr), dim3(64), 0, s, indices, data, ws, ar, width, st);       \
        RBG_CHECK_LAUNCH();                                                                                     \
      }                                                                                                         \
      if (hint < 0 || (hint & LK_LONG)) {                                                                       \
        hipLaunchKernelGGL((k_lvl_long<T>), dim3(gl), dim3(LV_NT), 0, s, indices, data, ws, ar, width, st);      \
        RBG_CHECK_LAUNCH();                                                                                     \
      }                                                                                                         \
      prof_stop(2, pi, s);                                                                                      \
      hipLaunchKernelGGL(k_round_end, dim3(1), dim3(1), 0, s, ws, r, n, hint);                                  \
      RBG_CHECK_LAUNCH();                                                                                       \
      if (exact) {                                                                                              \
        if (r + 1 == -batch) break;                                                                             \
      } else if ((r + 1) % batch == 0) {                                                                        \
        hipError_t e = hipMemcpyAsync(host, ws.counters, sizeof(host), hipMemcpyDeviceToHost, s);              \
        if (e != hipSuccess) return (int)e;                                                                     \
        e = hipStreamSynchronize(s);                                                                            \
        if (e != hipSuccess) return (int)e;                                                                     \
        if (host[C_DONE] || host[C_STUCK] || host[C_CAP]) break;                                                \
      }                                                                                                         \
    }                                                                                                           \
    hipLaunchKernelGGL(k_dag_report, dim3(1), dim3(1), 0, s, ws, st);                                           \
    RBG_CHECK_LAUNCH();                                                                                         \
    hipLaunchKernelGGL(k_dag_cycle_vertex, dim3(1), dim3(1), 0, s, indptr, indices, ws, n, ws.rem[0], st);      \
    RBG_CHECK_LAUNCH();                                                                                         \
    return 0;                                                                                                   \
  }                                                                                                             \
  int rbg_dag_gather_##SUF(void* workspace, int32_t n, const int32_t* arena_cols, const T* arena_vals,          \
                           int64_t* out_indptr, int32_t* out_indices, T* out_data, int64_t out_cap,             \
                           rbg_status* st, rbg_stream_t stream) {                                               \
    hipStream_t s = (hipStream_t)stream;                                                                        \
    if (n <= 0 || !workspace || !out_indptr) return RBG_E_ARG;                                                  \
    DagWork ws = DagWork::carve(workspace, n);                                                                  \
    DagArena<T> ar{(int32_t*)arena_cols, (T*)arena_vals, 0};                                                    \
    int rc = scan_exclusive(ws.row_len, out_indptr, n, ws.sums_a, ws.counters + C_NNZ, s);                      \
    if (rc) return rc;                                                                                          \
    hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, out_indptr, (int64_t)n,                              \
                       (const int64_t*)(ws.counters + C_NNZ));                                                  \
    RBG_CHECK_LAUNCH();                                                                                         \
    if (out_indices && out_data) {                                                                              \
      hipLaunchKernelGGL((k_dag_gather<T>), dim3((n + GATHER_NT - 1) / GATHER_NT), dim3(GATHER_NT), 0, s, ws, ar, \
                         (const int64_t*)out_indptr, n, out_indices, out_data, out_cap, st);                    \
      RBG_CHECK_LAUNCH();                                                                                       \
    }                                                                                                           \
    hipLaunchKernelGGL(k_dag_finish_status, dim3(1), dim3(1), 0, s, ws, st);                                     \
    RBG_CHECK_LAUNCH();                                                                                         \
    return 0;                                                                                                   \
  }

RBG_DAG_ENTRY(i32, int32_t)
RBG_DAG_ENTRY(i64, int64_t)

int rbg_topological_sort(const int64_t* indptr, const int32_t* indices, int32_t n, void* workspace,
                         int64_t ws_bytes, int32_t* level_out, int32_t* order_out, rbg_status* st,
                         rbg_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0 || !indptr || !indices || !workspace || !st) return RBG_E_ARG;
  if (ws_bytes < DagWork::bytes(n)) return RBG_E_WORKSPACE;
  DagWork ws = DagWork::carve(workspace, n);
  ws.ecap = ((char*)workspace + ws_bytes - (char*)ws.preds) / 4;
  const int gv = (int)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);
  hipLaunchKernelGGL(k_dag_init, dim3((n + 255) / 256), dim3(256), 0, s, ws, n);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_kahn_count, dim3(gv), dim3(256), 0, s, indptr, indices, ws, n);
  RBG_CHECK_LAUNCH();
  int rc = scan_exclusive(ws.cnt, ws.pptr, n, ws.sums_a, ws.counters + C_NNZ, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, ws.pptr, (int64_t)n, (const int64_t*)(ws.counters + C_NNZ));
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_kahn_fill, dim3(gv), dim3(256), 0, s, indptr, indices, ws, n);
  RBG_CHECK_LAUNCH();
  int64_t host[C_NCTR];
  for (int32_t r = 0; r <= n; r++) {
    if (r > 0) {
      hipLaunchKernelGGL((k_kahn_round<int32_t, false>), dim3(gv), dim3(KR_NT), 0, s, indptr, indices,
                         (const int32_t*)nullptr, ws, r);
      RBG_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_topo_round_end, dim3(1), dim3(1), 0, s, ws, r, n);
    RBG_CHECK_LAUNCH();
    if ((r + 1) % 8 == 0 || r == n) {
      hipError_t e = hipMemcpyAsync(host, ws.counters, sizeof(host), hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return (int)e;
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return (int)e;
      if (host[C_DONE] || host[C_STUCK] || host[C_CAP]) break;
    }
  }
  hipLaunchKernelGGL(k_dag_report, dim3(1), dim3(1), 0, s, ws, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_dag_cycle_vertex, dim3(1), dim3(1), 0, s, indptr, indices, ws, n, ws.rem[0], st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_topo_output, dim3((n + 255) / 256), dim3(256), 0, s, ws, n, level_out, order_out);
  RBG_CHECK_LAUNCH();
  return 0;
}

int rbg_csr_triangles(const int64_t* indptr, const int32_t* indices, int32_t n, rbg_status* st,
                      rbg_stream_t stream) {
  if (n < 0 || !st || (n > 0 && (!indptr || !indices))) return RBG_E_ARG;
  if (n == 0) return 0;
  const int grid = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_csr_triangles, dim3(grid), dim3(256), 0, (hipStream_t)stream, indptr, indices, n, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

int rbg_dag_level_kinds(void* workspace, int32_t n, uint8_t* host_dst, int32_t count, rbg_stream_t stream) {
  if (n <= 0 || !workspace || !host_dst || count < 0 || count > n) return RBG_E_ARG;
  if (count == 0) return 0;
  DagWork ws = DagWork::carve(workspace, n);
  hipError_t e = hipMemcpyAsync(host_dst, ws.kinds, (size_t)count, hipMemcpyDeviceToHost, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"

extern "C" {
// Sort rows, combine duplicates with the AVOS sum, drop explicit zeros, in
// place on (idx, val); the canonical matrix is written to (out_indptr,
// out_idx, out_val).  tmp needs rbg_csr_tmp_bytes(n).
int64_t rbg_csr_tmp_bytes(int32_t n) { return 8 * ((int64_t)n + 1) + 8 * ((int64_t)n / 256 + 8) + 8 * 4 + 1024; }

#define RBG_CSR_ENTRY(SUF, T)                                                                                   \
  int rbg_csr_canonicalize_##SUF(const int64_t* indptr, int32_t* idx, T* val, int32_t n, void* tmp,            \
                                 int64_t* out_indptr, int32_t* out_idx, T* out_val, rbg_stream_t stream) {     \
    hipStream_t s = (hipStream_t)stream;                                                                        \
    if (n < 0 || !indptr || !tmp || !out_indptr) return RBG_E_ARG;                                              \
    if (n == 0) return 0;                                                                                       \
    int64_t* len = (int64_t*)(((uintptr_t)tmp + 255) & ~(uintptr_t)255);                                        \
    int64_t* stmp = len + ((n + 1 + 31) & ~31);                                                                 \
    int64_t* total = stmp + (n / 256 + 8);                                                                      \
    hipLaunchKernelGGL((k_csr_sort_rows<T>), dim3((n + 255) / 256), dim3(256), 0, s, indptr, idx, val, n, len); \
    RBG_CHECK_LAUNCH();                                                                                         \
    int rc = scan_exclusive(len, out_indptr, n, stmp, total, s);                                                \
    if (rc) return rc;                                                                                          \
    hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, out_indptr, (int64_t)n, (const int64_t*)total);      \
    RBG_CHECK_LAUNCH();                                                                                         \
    if (out_idx && out_val) {                                                                                   \
      hipLaunchKernelGGL((k_csr_compact<T>), dim3((n + GATHER_NT - 1) / GATHER_NT), dim3(GATHER_NT), 0, s, indptr, (const int32_t*)idx,     \
                         (const T*)val, (const int64_t*)len, (const int64_t*)out_indptr, n, out_idx, out_val);   \
      RBG_CHECK_LAUNCH();                                                                                       \
    }                                                                                                           \
    return 0;                                                                                                   \
  }
RBG_CSR_ENTRY(i32, int32_t)
RBG_CSR_ENTRY(i64, int64_t)

#define RBG_CSR_CANON_ENTRY(SUF, T)                                                                             \
  int rbg_csr_check_canonical_##SUF(const int64_t* indptr, const int32_t* idx, const T* val, int32_t n,        \
                                    rbg_status* st, rbg_stream_t stream) {                                      \
    if (n < 0 || !indptr || !st) return RBG_E_ARG;                                                             \
    if (n == 0) return 0;                                                                                       \
    hipLaunchKernelGGL((k_csr_check_canonical<T>), dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,   \
                       indptr, idx, val, n, st);                                                                \
    RBG_CHECK_LAUNCH();                                                                                         \
    return 0;                                                                                                   \
  }
RBG_CSR_CANON_ENTRY(i32, int32_t)
RBG_CSR_CANON_ENTRY(i64, int64_t)

// Validate a CSR structure: monotone indptr, column indices in [0, ncols).
// Sets RBG_F_CAPACITY in st->flags on a violation.
int rbg_csr_validate(const int64_t* indptr, const int32_t* idx, int32_t n, int32_t ncols, rbg_status* st,
                     rbg_stream_t stream) {
  if (n < 0 || !indptr || !st) return RBG_E_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_csr_check, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, indptr, idx, n, ncols,
                     st);
  RBG_CHECK_LAUNCH();
  return 0;
}
}  // extern "C"

extern "C" {
#define RBG_SPGEMM_ENTRY(SUF, T)                                                                                \
  int rbg_spgemm_rows_##SUF(const int64_t* ap, const int32_t* aj, const T* ax, int32_t m, const int64_t* bp,     \
                            const int32_t* bj, const T* bx, int32_t mask_upper, void* workspace,                 \
                            int32_t* arena_cols, T* arena_vals, int64_t arena_cap, int32_t* tmp_cols,            \
                            T* tmp_vals, int64_t tmp_cap, int32_t width, int32_t is_unsigned, rbg_status* st,    \
                            rbg_stream_t stream) {                                                              \
    hipStream_t s = (hipStream_t)stream;                                                                        \
    if (m <= 0 || !workspace || !arena_cols || !arena_vals) return RBG_E_ARG;                                   \
    DagWork ws = DagWork::carve(workspace, m);                                                                  \
    RowArena<T> ar{arena_cols, arena_vals, arena_cap, tmp_cols, tmp_vals, tmp_cap};                             \
    hipLaunchKernelGGL(k_rows_init, dim3((m + 255) / 256), dim3(256), 0, s, ws.row_len, ws.counters, m);        \
    RBG_CHECK_LAUNCH();                                                                                         \
    int pi = prof_start(s);                                                                                     \
    hipLaunchKernelGGL((k_spgemm_rows<T>), dim3((m + DAG_WAVES - 1) / DAG_WAVES), dim3(64 * DAG_WAVES), 0, s,   \
                       ap, aj, ax, m, bp, bj, bx, mask_upper, ws, ar, width, is_unsigned, st);                  \
    RBG_CHECK_LAUNCH();                                                                                         \
    prof_stop(3, pi, s);                                                                                        \
    return 0;                                                                                                   \
  }
RBG_SPGEMM_ENTRY(i32, int32_t)
RBG_SPGEMM_ENTRY(i64, int64_t)
}  // extern "C"

extern "C" {
#define RBG_CSR_ADD_ENTRY(SUF, T)                                                                               \
  int rbg_csr_add_##SUF(const int64_t* ap, const int32_t* aj, const T* ax, const int64_t* bp, const int32_t* bj, \
                        const T* bx, int32_t n, int32_t combine, const int64_t* cat_ptr, int32_t* cat_idx,      \
                        T* cat_val, void* tmp, int64_t* out_indptr, int32_t* out_idx, T* out_val,               \
                        rbg_stream_t stream) {                                                                  \
    hipStream_t s = (hipStream_t)stream;                                                                        \
    if (n < 0 || !tmp || !out_indptr) return RBG_E_ARG;                                                         \
    if (n == 0) return 0;                                                                                       \
    int64_t* len = (int64_t*)(((uintptr_t)tmp + 255) & ~(uintptr_t)255);                                        \
    int64_t* stmp = len + ((n + 1 + 31) & ~31);                                                                 \
    int64_t* total = stmp + (n / 256 + 8);                                                                      \
    hipLaunchKernelGGL((k_csr_add_rows<T>), dim3((n + 255) / 256), dim3(256), 0, s, ap, aj, ax, bp, bj, bx, n,   \
                       combine, cat_ptr, cat_idx, cat_val, len);                                                \
    RBG_CHECK_LAUNCH();                                                                                         \
    int rc = scan_exclusive(len, out_indptr, n, stmp, total, s);                                                \
    if (rc) return rc;                                                                                          \
    hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, out_indptr, (int64_t)n, (const int64_t*)total);      \
    RBG_CHECK_LAUNCH();                                                                                         \
    if (out_idx && out_val) {                                                                                   \
      hipLaunchKernelGGL((k_csr_compact<T>), dim3((n + GATHER_NT - 1) / GATHER_NT), dim3(GATHER_NT), 0, s,          \
                         cat_ptr, (const int32_t*)cat_idx,                                                      \
                         (const T*)cat_val, (const int64_t*)len, (const int64_t*)out_indptr, n, out_idx, out_val); \
      RBG_CHECK_LAUNCH();                                                                                       \
    }                                                                                                           \
    return 0;                                                                                                   \
  }
RBG_CSR_ADD_ENTRY(i32, int32_t)
RBG_CSR_ADD_ENTRY(i64, int64_t)
}  // extern "C"
