// Device helpers around the closures: CSR <-> dense, column and row edits of a
// closed CSR, the triangular filter, CSR reductions, and the component layout,
// extraction and merge of the component-wise closure.
//
// Reference:
//   sparse/csgraph/_components.pyx:180-251 (extract_submatrix), :254-323
//     (merge_component_matrices) and transitive_closure.py:85-195
//     (component_wise_closure: find, extract, close, merge);
//   sparse/csgraph/_permutation.pyx:138-157 (the triangular filter of
//     permute_sparse);
//   core/redblack.py:74-97 (edge composition: the composed row and column);
//   gpu/transitive_closure.py:43-50 (the squaring closure's final diameter);
//   gpu/spgemm_symbolic.py:124-183 (the SpGEMM candidate count).
// Row counts are turned into row pointers by rocPRIM's device scan; the
// component layout sorts vertices by component with rocPRIM's radix sort.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rbg_common.hpp"

namespace rbg {

constexpr int GL_NT = 256;

static inline int gl_grid(int64_t n, int per = GL_NT, int cap = 8192) {
  const int64_t g = (n + per - 1) / per;
  return (int)(g < 1 ? 1 : (g < cap ? g : cap));
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t scan_temp_bytes(int64_t n) {
  size_t b = 0;
  (void)rocprim::inclusive_scan((void*)nullptr, b, (const int64_t*)nullptr, (int64_t*)nullptr, (size_t)n,
                                rocprim::plus<int64_t>(), (hipStream_t)0);
  return b;
}

static size_t sort_temp_bytes(int64_t n) {
  size_t b = 0;
  (void)rocprim::radix_sort_pairs((void*)nullptr, b, (const int32_t*)nullptr, (int32_t*)nullptr,
                                  (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0u, 32u, (hipStream_t)0);
  return b;
}

__global__ void k_gl_zero_i64(int64_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// out[0] = 0, out[1..n] = inclusive scan of cnt[0..n): row pointers from counts
static int rows_to_ptr(const int64_t* cnt, int64_t n, int64_t* out, void* temp, size_t temp_bytes, hipStream_t s) {
  hipLaunchKernelGGL(k_gl_zero_i64, dim3(1), dim3(64), 0, s, out, (int64_t)1);
  RBG_CHECK_LAUNCH();
  if (n == 0) return 0;
  size_t need = scan_temp_bytes(n);
  if (temp_bytes < need) return RBG_E_WORKSPACE;
  hipError_t e = rocprim::inclusive_scan(temp, need, cnt, out + 1, (size_t)n, rocprim::plus<int64_t>(), s);
  return e == hipSuccess ? 0 : (int)e;
}

// ---- reductions -------------------------------------------------------------
// aux[0] += sum over A's entries (i, k) of |row k of B|: the candidate count
// of C = A (x) B (an exact bound on every row's output)
__global__ void k_gl_expand(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj, int32_t m,
                            const int64_t* __restrict__ bp, rbg_status* st) {
  const int64_t nnz = ap[m];
  int64_t acc = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k = aj[e];
    acc += bp[k + 1] - bp[k];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd((unsigned long long*)&st->aux[0], (unsigned long long)acc);
}

// aux[0] = max |v| over the stored values (atomicMax)
template <typename T>
__global__ void k_gl_max_abs(const int64_t* __restrict__ ap, const T* __restrict__ ax, int32_t n, rbg_status* st) {
  const int64_t nnz = ap[n];
  int64_t m = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = (int64_t)ax[e];
    const int64_t a = v < 0 ? -v : v;
    m = a > m ? a : m;
  }
  m = wave_max_i64(m);
  if ((threadIdx.x & 63) == 0 && m) atomicMax((long long*)&st->aux[0], (long long)m);
}

// ---- CSR <-> dense ------------------------------------------------------------
// One wave per row: the row cleared, then its entries scattered.
template <typename T>
__global__ __launch_bounds__(GL_NT) void k_gl_csr_to_dense(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj,
                                                           const T* __restrict__ ax, int32_t n, int64_t ncols,
                                                           T* __restrict__ D) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (GL_NT / 64);
  for (int64_t r = (int64_t)blockIdx.x * (GL_NT / 64) + (threadIdx.x >> 6); r < n; r += waves) {
    T* row = D + r * ncols;
    for (int64_t c = lane; c < ncols; c += 64) row[c] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int64_t e = ap[r] + lane; e < ap[r + 1]; e += 64) row[aj[e]] = ax[e];
  }
}

// One wave per row: the row's nonzero count (pass 1), or its nonzeros
// written in column order at out_ptr[r] (pass 2: a ballot per 64 columns).
template <typename T, bool FILL>
__global__ __launch_bounds__(GL_NT) void k_gl_dense_rows(const T* __restrict__ D, int32_t n, int64_t ncols,
                                                         int64_t* __restrict__ cnt, const int64_t* __restrict__ out_ptr,
                                                         int32_t* __restrict__ oj, T* __restrict__ ox) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (GL_NT / 64);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int64_t r = (int64_t)blockIdx.x * (GL_NT / 64) + (threadIdx.x >> 6); r < n; r += waves) {
    const T* row = D + r * ncols;
    int64_t base = FILL ? out_ptr[r] : 0;
    for (int64_t c0 = 0; c0 < ncols; c0 += 64) {
      const int64_t c = c0 + lane;
      const T v = c < ncols ? row[c] : (T)0;
      const uint64_t m = __ballot(v != 0);
      if (FILL && v != 0) {
        const int64_t q = base + __popcll(m & below);
        oj[q] = (int32_t)c;
        ox[q] = v;
      }
      base += __popcll(m);
    }
    if (!FILL && lane == 0) cnt[r] = base;
  }
}

// ---- column and row of a closed CSR -----------------------------------------------
// Row r's entry in column `col` (rows sorted): 1 or 0 (pass 1), written
// (pass 2) as row r of an n x 1 CSR.
template <typename T, bool FILL>
__global__ void k_gl_column(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj, const T* __restrict__ ax,
                            int32_t n, int32_t col, int64_t* __restrict__ cnt, const int64_t* __restrict__ out_ptr,
                            int32_t* __restrict__ oj, T* __restrict__ ox) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = ap[r], hi = ap[r + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (aj[mid] < col) lo = mid + 1;
      else hi = mid;
    }
    const bool hit = lo < ap[r + 1] && aj[lo] == col;
    if (!FILL) {
      cnt[r] = hit ? 1 : 0;
    } else if (hit) {
      oj[out_ptr[r]] = 0;
      ox[out_ptr[r]] = ax[lo];
    }
  }
}

// Row `row` of a sorted CSR with the entry in column `col` set to `val`
// (replaced, or inserted in column order); aux[0] = the row's new length.
// One wave.
template <typename T>
__global__ void k_gl_row_set(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj, const T* __restrict__ ax,
                             int64_t row, int32_t col, T val, int32_t* __restrict__ oj, T* __restrict__ ox,
                             rbg_status* st) {
  const int lane = threadIdx.x;
  const int64_t e0 = ap[row], len = ap[row + 1] - e0;
  // position of col: the entries before it
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (aj[e0 + mid] < col) lo = mid + 1;
    else hi = mid;
  }
  const bool present = lo < len && aj[e0 + lo] == col;
  for (int64_t k = lane; k < len; k += 64) {
    const int32_t c = aj[e0 + k];
    const int64_t q = (present || k < lo) ? k : k + 1;
    oj[q] = c;
    ox[q] = c == col ? val : ax[e0 + k];
  }
  if (lane == 0) {
    if (!present) {
      oj[lo] = col;
      ox[lo] = val;
    }
    st->aux[0] = present ? len : len + 1;
  }
}

// ---- triangular filter ----------------------------------------------------------
// Entries on or above the diagonal: per-row count (pass 1) and copy (pass 2).
// Rows are sorted, so they are a suffix of each row.
template <typename T, bool FILL>
__global__ void k_gl_upper(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj, const T* __restrict__ ax,
                           int32_t n, int64_t* __restrict__ cnt, const int64_t* __restrict__ out_ptr,
                           int32_t* __restrict__ oj, T* __restrict__ ox) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); r < n; r += waves) {
    int64_t lo = ap[r], hi = ap[r + 1];
    const int64_t e1 = hi;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (aj[mid] < r) lo = mid + 1;
      else hi = mid;
    }
    if (!FILL) {
      if (lane == 0) cnt[r] = e1 - lo;
    } else {
      const int64_t d = out_ptr[r] - lo;
      for (int64_t e = lo + lane; e < e1; e += 64) {
        oj[d + e] = aj[e];
        ox[d + e] = ax[e];
      }
    }
  }
}

// ---- components -------------------------------------------------------------------
// is_root[v] = (labels[v] == v), scanned into each root's component id
__global__ void k_gl_roots(const int32_t* __restrict__ lab, int32_t n, int64_t* __restrict__ flag) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    flag[v] = lab[v] == v ? 1 : 0;
}

// comp[v] = (inclusive root count at v's root) - 1
__global__ void k_gl_comp_ids(const int32_t* __restrict__ lab, int32_t n, const int64_t* __restrict__ incl,
                              int32_t* __restrict__ comp, int64_t* ncomp) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    comp[v] = (int32_t)(incl[lab[v]] - 1);
  if (blockIdx.x == 0 && threadIdx.x == 0) *ncomp = incl[n - 1];
}

// From the component ids in sorted order: starts[c] = the first place of
// component c, and n for c from the count up to n (no atomics: one giant
// component would serialise a per-vertex count on one word)
__global__ void k_gl_starts(const int32_t* __restrict__ keys, int32_t n, const int64_t* __restrict__ ncomp,
                            int64_t* __restrict__ starts) {
  const int64_t nc = *ncomp;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q <= n; q += (int64_t)gridDim.x * blockDim.x) {
    if (q < n) {
      const int32_t c = keys[q];
      if (q == 0 || keys[q - 1] != c) starts[c] = q;
    }
    if (q >= nc) starts[q] = n;
  }
}

__global__ void k_gl_sizes(const int64_t* __restrict__ starts, int32_t n, const int64_t* __restrict__ ncomp,
                           int32_t* __restrict__ sizes) {
  const int64_t nc = *ncomp;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    sizes[c] = c < nc ? (int32_t)(starts[c + 1] - starts[c]) : 0;
}

__global__ void k_gl_iota(int32_t* p, int32_t n) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    p[v] = (int32_t)v;
}

// members sorted by (component, vertex): local[v] = its place in its
// component; vsize[v] (when given) = its component's vertex count
__global__ void k_gl_local(const int32_t* __restrict__ members, const int32_t* __restrict__ comp,
                           const int64_t* __restrict__ starts, const int32_t* __restrict__ sizes, int32_t n,
                           int32_t* __restrict__ local, int32_t* __restrict__ vsize) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = members[q];
    const int32_t c = comp[v];
    local[v] = (int32_t)(q - starts[c]);
    if (vsize) vsize[v] = sizes[c];
  }
}

// Entry (v, w, x) of the graph to cell (local v, local w) of component
// comp[v]'s padded matrix in the pool (one wave per row).
__global__ __launch_bounds__(GL_NT) void k_gl_scatter(const int64_t* __restrict__ ap, const int32_t* __restrict__ aj,
                                                      const int32_t* __restrict__ ax, int32_t n,
                                                      const int32_t* __restrict__ comp, const int32_t* __restrict__ local,
                                                      const int64_t* __restrict__ off, const int32_t* __restrict__ npad,
                                                      int32_t* __restrict__ pool) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (GL_NT / 64);
  for (int64_t v = (int64_t)blockIdx.x * (GL_NT / 64) + (threadIdx.x >> 6); v < n; v += waves) {
    const int32_t c = comp[v];
    int32_t* row = pool + off[c] + (int64_t)local[v] * npad[c];
    for (int64_t e = ap[v] + lane; e < ap[v + 1]; e += 64) row[local[aj[e]]] = ax[e];
  }
}

// Global row v of the merged closure = row local[v] of component comp[v]'s
// closed matrix, its columns renamed to the component's members (which keep
// the column order): count (pass 1) and write (pass 2), one wave per row.
template <bool FILL>
__global__ __launch_bounds__(GL_NT) void k_gl_merge(const int32_t* __restrict__ pool, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ npad, const int32_t* __restrict__ comp,
                                                    const int32_t* __restrict__ local, const int32_t* __restrict__ members,
                                                    const int64_t* __restrict__ starts, const int32_t* __restrict__ sizes,
                                                    int32_t n, int64_t* __restrict__ cnt,
                                                    const int64_t* __restrict__ out_ptr, int32_t* __restrict__ oj,
                                                    int32_t* __restrict__ ox) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (GL_NT / 64);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int64_t v = (int64_t)blockIdx.x * (GL_NT / 64) + (threadIdx.x >> 6); v < n; v += waves) {
    const int32_t c = comp[v], size = sizes[c];
    const int32_t* row = pool + off[c] + (int64_t)local[v] * npad[c];
    const int32_t* mem = members + starts[c];
    int64_t base = FILL ? out_ptr[v] : 0;
    for (int32_t j0 = 0; j0 < size; j0 += 64) {
      const int32_t j = j0 + lane;
      const int32_t x = j < size ? row[j] : 0;
      const uint64_t m = __ballot(x != 0);
      if (FILL && x != 0) {
        const int64_t q = base + __popcll(m & below);
        oj[q] = mem[j];
        ox[q] = x;
      }
      base += __popcll(m);
    }
    if (!FILL && lane == 0) cnt[v] = base;
  }
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int64_t rbg_glue_temp_bytes(int64_t n) {
  const int64_t m = n > 0 ? n : 1;
  size_t a = scan_temp_bytes(m), b = sort_temp_bytes(m);
  // scan / sort storage, then an int64 count per row and two int32 buffers of n
  return (int64_t)(al256(a > b ? a : b) + al256(8 * (size_t)(m + 1)) + 2 * al256(4 * (size_t)m) + 512);
}

// carve: [0] scan/sort storage, [1] int64 counts (n + 1), [2], [3] int32 scratch (n)
struct GlueTemp {
  void* store;
  size_t store_bytes;
  int64_t* cnt;
  int32_t* s0;
  int32_t* s1;
  static bool carve(void* base, int64_t bytes, int64_t n, GlueTemp* t) {
    if (!base || bytes < rbg_glue_temp_bytes(n)) return false;
    const int64_t m = n > 0 ? n : 1;
    size_t a = scan_temp_bytes(m), b = sort_temp_bytes(m);
    char* p = (char*)(((uintptr_t)base + 255) & ~(uintptr_t)255);
    t->store = p;
    t->store_bytes = al256(a > b ? a : b);
    p += t->store_bytes;
    t->cnt = (int64_t*)p;
    p += al256(8 * (size_t)(m + 1));
    t->s0 = (int32_t*)p;
    p += al256(4 * (size_t)m);
    t->s1 = (int32_t*)p;
    return true;
  }
};

int rbg_csr_expand_size(const int64_t* ap, const int32_t* aj, int32_t m, const int64_t* bp, rbg_status* st,
                        rbg_stream_t stream) {
  if (m < 0 || !ap || !bp || !st || (m > 0 && !aj)) return RBG_E_ARG;
  if (m == 0) return 0;
  hipLaunchKernelGGL(k_gl_expand, dim3(2048), dim3(GL_NT), 0, (hipStream_t)stream, ap, aj, m, bp, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

#define RBG_GLUE_TYPED(SUF, T)                                                                                    \
  int rbg_csr_max_abs_##SUF(const int64_t* ap, const T* ax, int32_t n, rbg_status* st, rbg_stream_t stream) {     \
    if (n < 0 || !ap || !st || (n > 0 && !ax)) return RBG_E_ARG;                                                  \
    if (n == 0) return 0;                                                                                         \
    hipLaunchKernelGGL((k_gl_max_abs<T>), dim3(2048), dim3(GL_NT), 0, (hipStream_t)stream, ap, ax, n, st);        \
    RBG_CHECK_LAUNCH();                                                                                           \
    return 0;                                                                                                     \
  }                                                                                                               \
  int rbg_csr_to_dense_##SUF(const int64_t* ap, const int32_t* aj, const T* ax, int32_t n, int64_t ncols, T* D,   \
                             rbg_stream_t stream) {                                                               \
    if (n < 0 || ncols < 0 || !ap || !D) return RBG_E_ARG;                                                        \
    if (n == 0) return 0;                                                                                         \
    hipLaunchKernelGGL((k_gl_csr_to_dense<T>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, (hipStream_t)stream, ap, aj,  \
                       ax, n, ncols, D);                                                                          \
    RBG_CHECK_LAUNCH();                                                                                           \
    return 0;                                                                                                     \
  }                                                                                                               \
  int rbg_dense_to_csr_##SUF(const T* D, int32_t n, int64_t ncols, int64_t* out_indptr, int32_t* out_indices,     \
                             T* out_data, void* temp, int64_t temp_bytes, rbg_stream_t stream) {                  \
    hipStream_t s = (hipStream_t)stream;                                                                          \
    if (n < 0 || ncols < 0 || !D || !out_indptr) return RBG_E_ARG;                                                \
    if (!out_indices || !out_data) {                                                                              \
      GlueTemp t;                                                                                                 \
      if (!GlueTemp::carve(temp, temp_bytes, n, &t)) return RBG_E_WORKSPACE;                                     \
      if (n > 0) {                                                                                                \
        hipLaunchKernelGGL((k_gl_dense_rows<T, false>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, s, D, n, ncols,     \
                           t.cnt, nullptr, nullptr, nullptr);                                                     \
        RBG_CHECK_LAUNCH();                                                                                       \
      }                                                                                                           \
      return rows_to_ptr(t.cnt, n, out_indptr, t.store, t.store_bytes, s);                                        \
    }                                                                                                             \
    if (n == 0) return 0;                                                                                         \
    hipLaunchKernelGGL((k_gl_dense_rows<T, true>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, s, D, n, ncols, nullptr,  \
                       (const int64_t*)out_indptr, out_indices, out_data);                                        \
    RBG_CHECK_LAUNCH();                                                                                           \
    return 0;                                                                                                     \
  }                                                                                                               \
  int rbg_csr_column_##SUF(const int64_t* ap, const int32_t* aj, const T* ax, int32_t n, int32_t col,             \
                           int64_t* out_indptr, int32_t* out_indices, T* out_data, void* temp, int64_t temp_bytes,\
                           rbg_stream_t stream) {                                                                 \
    hipStream_t s = (hipStream_t)stream;                                                                          \
    if (n <= 0 || col < 0 || !ap || !aj || !ax || !out_indptr || !out_indices || !out_data) return RBG_E_ARG;     \
    GlueTemp t;                                                                                                   \
    if (!GlueTemp::carve(temp, temp_bytes, n, &t)) return RBG_E_WORKSPACE;                                       \
    hipLaunchKernelGGL((k_gl_column<T, false>), dim3(gl_grid(n)), dim3(GL_NT), 0, s, ap, aj, ax, n, col, t.cnt,   \
                       nullptr, nullptr, nullptr);                                                                \
    RBG_CHECK_LAUNCH();                                                                                           \
    int rc = rows_to_ptr(t.cnt, n, out_indptr, t.store, t.store_bytes, s);                                        \
    if (rc) return rc;                                                                                            \
    hipLaunchKernelGGL((k_gl_column<T, true>), dim3(gl_grid(n)), dim3(GL_NT), 0, s, ap, aj, ax, n, col, nullptr,  \
                       (const int64_t*)out_indptr, out_indices, out_data);                                        \
    RBG_CHECK_LAUNCH();                                                                                           \
    return 0;                                                                                                     \
  }                                                                                                               \
  int rbg_csr_row_set_##SUF(const int64_t* ap, const int32_t* aj, const T* ax, int64_t row, int32_t col, T val,   \
                            int32_t* out_indices, T* out_data, rbg_status* st, rbg_stream_t stream) {             \
    if (!ap || !aj || !ax || row < 0 || col < 0 || !out_indices || !out_data || !st) return RBG_E_ARG;            \
    hipLaunchKernelGGL((k_gl_row_set<T>), dim3(1), dim3(64), 0, (hipStream_t)stream, ap, aj, ax, row, col, val,   \
                       out_indices, out_data, st);                                                                \
    RBG_CHECK_LAUNCH();                                                                                           \
    return 0;                                                                                                     \
  }                                                                                                               \
  int rbg_csr_upper_##SUF(const int64_t* ap, const int32_t* aj, const T* ax, int32_t n, int64_t* out_indptr,      \
                          int32_t* out_indices, T* out_data, void* temp, int64_t temp_bytes, rbg_stream_t stream) { \
    hipStream_t s = (hipStream_t)stream;                                                                          \
    if (n <= 0 || !ap || !out_indptr) return RBG_E_ARG;                                                           \
    GlueTemp t;                                                                                                   \
    if (!GlueTemp::carve(temp, temp_bytes, n, &t)) return RBG_E_WORKSPACE;                                       \
    hipLaunchKernelGGL((k_gl_upper<T, false>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, s, ap, aj, ax, n, t.cnt,      \
                       nullptr, nullptr, nullptr);                                                                \
    RBG_CHECK_LAUNCH();                                                                                           \
    int rc = rows_to_ptr(t.cnt, n, out_indptr, t.store, t.store_bytes, s);                                        \
    if (rc) return rc;                                                                                            \
    if (out_indices && out_data) {                                                                                \
      hipLaunchKernelGGL((k_gl_upper<T, true>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, s, ap, aj, ax, n, nullptr,   \
                         (const int64_t*)out_indptr, out_indices, out_data);                                      \
      RBG_CHECK_LAUNCH();                                                                                         \
    }                                                                                                             \
    return 0;                                                                                                     \
  }

RBG_GLUE_TYPED(i32, int32_t)
RBG_GLUE_TYPED(i64, int64_t)

int rbg_components_layout(const int32_t* labels, int32_t n, int32_t* comp, int32_t* local, int32_t* members,
                          int32_t* sizes, int64_t* starts, int64_t* ncomp, int32_t* vertex_size, void* temp,
                          int64_t temp_bytes, rbg_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0 || !labels || !comp || !local || !members || !sizes || !starts || !ncomp) return RBG_E_ARG;
  GlueTemp t;
  if (!GlueTemp::carve(temp, temp_bytes, n, &t)) return RBG_E_WORKSPACE;
  const int g = gl_grid(n);
  hipLaunchKernelGGL(k_gl_roots, dim3(g), dim3(GL_NT), 0, s, labels, n, t.cnt);
  RBG_CHECK_LAUNCH();
  // inclusive root counts into `starts` (n entries of int64) as scratch
  hipError_t e = rocprim::inclusive_scan(t.store, t.store_bytes, t.cnt, starts, (size_t)n, rocprim::plus<int64_t>(), s);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_gl_comp_ids, dim3(g), dim3(GL_NT), 0, s, labels, n, (const int64_t*)starts, comp, ncomp);
  RBG_CHECK_LAUNCH();
  // members: the vertices sorted by component (stable: increasing within one)
  hipLaunchKernelGGL(k_gl_iota, dim3(g), dim3(GL_NT), 0, s, t.s0, n);
  RBG_CHECK_LAUNCH();
  unsigned bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) < (int64_t)n) bits++;
  e = rocprim::radix_sort_pairs(t.store, t.store_bytes, (const int32_t*)comp, t.s1, (const int32_t*)t.s0, members,
                                (size_t)n, 0u, bits, s);
  if (e != hipSuccess) return (int)e;
  // component starts from the sorted ids (starts[c] = n past the count),
  // then the sizes as their differences
  hipLaunchKernelGGL(k_gl_starts, dim3(g), dim3(GL_NT), 0, s, (const int32_t*)t.s1, n, (const int64_t*)ncomp, starts);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_gl_sizes, dim3(g), dim3(GL_NT), 0, s, (const int64_t*)starts, n, (const int64_t*)ncomp, sizes);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_gl_local, dim3(g), dim3(GL_NT), 0, s, (const int32_t*)members, (const int32_t*)comp,
                     (const int64_t*)starts, (const int32_t*)sizes, n, local, vertex_size);
  RBG_CHECK_LAUNCH();
  return 0;
}

int rbg_components_scatter(const int64_t* indptr, const int32_t* indices, const int32_t* data, int32_t n,
                           const int32_t* comp, const int32_t* local, const int64_t* comp_offset,
                           const int32_t* comp_npad, int32_t* pool, rbg_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !data || !comp || !local || !comp_offset || !comp_npad || !pool)
    return RBG_E_ARG;
  hipLaunchKernelGGL(k_gl_scatter, dim3(gl_grid(n, 4)), dim3(GL_NT), 0, (hipStream_t)stream, indptr, indices, data, n,
                     comp, local, comp_offset, comp_npad, pool);
  RBG_CHECK_LAUNCH();
  return 0;
}

int rbg_components_merge(const int32_t* pool, const int64_t* comp_offset, const int32_t* comp_npad,
                         const int32_t* comp, const int32_t* local, const int32_t* members, const int64_t* starts,
                         const int32_t* sizes, int32_t n, int64_t* out_indptr, int32_t* out_indices,
                         int32_t* out_data, void* temp, int64_t temp_bytes, rbg_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0 || !pool || !comp_offset || !comp_npad || !comp || !local || !members || !starts || !sizes ||
      !out_indptr)
    return RBG_E_ARG;
  if (!out_indices || !out_data) {
    GlueTemp t;
    if (!GlueTemp::carve(temp, temp_bytes, n, &t)) return RBG_E_WORKSPACE;
    hipLaunchKernelGGL((k_gl_merge<false>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, s, pool, comp_offset, comp_npad,
                       comp, local, members, starts, sizes, n, t.cnt, nullptr, nullptr, nullptr);
    RBG_CHECK_LAUNCH();
    return rows_to_ptr(t.cnt, n, out_indptr, t.store, t.store_bytes, s);
  }
  hipLaunchKernelGGL((k_gl_merge<true>), dim3(gl_grid(n, 4)), dim3(GL_NT), 0, s, pool, comp_offset, comp_npad, comp,
                     local, members, starts, sizes, n, nullptr, (const int64_t*)out_indptr, out_indices, out_data);
  RBG_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
