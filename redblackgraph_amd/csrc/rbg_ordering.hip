// Canonical ordering of a closed graph: the vertices' sort keys, the device
// sort of the vertices, and the permutation of the matrix (dense or CSR).
//
// Reference: sparse/csgraph/_ordering.pyx:17-47 (_get_permutation, dense),
// :50-157 (_get_permutation_sparse and _compute_ordering_metrics_sparse) and
// reference/ordering.py:13-53; sparse/csgraph/_permutation.pyx:38-66 (dense
// permute) and :67-136 (permute_sparse without the triangular filter).
// Vertices sort on
//   (-component size, component id, -ancestor count, max relationship,
//    -diagonal, vertex id)
// where a vertex's ancestor count is the sum of MSB(v) over the nonzero
// entries v of its row (MSB(v) = floor(log2 v) for v > 1, else 0: the
// _rbg_math.pxi loop that shifts while v > 1), its max relationship is the
// largest entry of its column and at least 0, and its diagonal is its own
// entry (0 when a CSR row holds none).  Sums and maxima are kept in int64.
#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "rbg_common.hpp"

namespace rbg {

template <typename T>
__device__ __forceinline__ int64_t ord_msb(T x) {
  if (x <= 1) return 0;
  if constexpr (sizeof(T) == 8) return 63 - __clzll((long long)x);
  else return 31 - __clz((int)x);
}

constexpr int ORD_NT = 256;
constexpr int ORD_COL_ROWS = 128;  // rows per workgroup of the column maxima

__global__ void k_ord_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_ord_fill_i64(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// One workgroup per row: its ancestor count and diagonal.
template <typename T>
__global__ __launch_bounds__(ORD_NT) void k_ord_rows_dense(const T* __restrict__ A, int64_t n,
                                                           int64_t* __restrict__ anc, int64_t* __restrict__ diag) {
  __shared__ int64_t red[ORD_NT];
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const T* row = A + i * n;
    int64_t s = 0;
    for (int64_t j = threadIdx.x; j < n; j += ORD_NT) {
      const T v = row[j];
      if (v != 0) s += ord_msb<T>(v);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = ORD_NT / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      anc[i] = red[0];
      diag[i] = (int64_t)row[i];
    }
    __syncthreads();
  }
}

// Workgroup (x, y): columns [256 x, 256 x + 256) over rows [128 y, 128 y + 128);
// each lane folds its column's maximum into maxrel (filled with 0 first).
template <typename T>
__global__ __launch_bounds__(ORD_NT) void k_ord_cols_dense(const T* __restrict__ A, int64_t n,
                                                           int64_t* __restrict__ maxrel) {
  const int64_t j = (int64_t)blockIdx.x * ORD_NT + threadIdx.x;
  if (j >= n) return;
  const int64_t r0 = (int64_t)blockIdx.y * ORD_COL_ROWS;
  const int64_t r1 = r0 + ORD_COL_ROWS < n ? r0 + ORD_COL_ROWS : n;
  int64_t m = 0;
  for (int64_t i = r0; i < r1; i++) {
    const int64_t v = (int64_t)A[i * n + j];
    m = v > m ? v : m;
  }
  if (m > 0) atomicMax((long long*)&maxrel[j], (long long)m);
}

// One wave per row: its ancestor count (a wave sum) and diagonal, and its
// entries folded into their columns' maxima (filled with 0 first).  The max
// only grows, so an entry no larger than the stored maximum needs no atomic.
template <typename T>
__global__ __launch_bounds__(ORD_NT) void k_ord_csr(const int64_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices, const T* __restrict__ data,
                                                    int32_t n, int64_t* __restrict__ anc, int64_t* maxrel,
                                                    int64_t* __restrict__ diag) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (ORD_NT / 64);
  for (int64_t i = (int64_t)blockIdx.x * (ORD_NT / 64) + (threadIdx.x >> 6); i < n; i += waves) {
    const int64_t e0 = indptr[i], e1 = indptr[i + 1];
    int64_t s = 0, de = -1;  // de: the row's last diagonal entry, as in the reference's loop
    for (int64_t e = e0 + lane; e < e1; e += 64) {
      const int32_t c = indices[e];
      const T v = data[e];
      if (v != 0) {
        s += ord_msb<T>(v);
        if (v > 0 && (int64_t)v > __atomic_load_n(&maxrel[c], __ATOMIC_RELAXED))
          atomicMax((long long*)&maxrel[c], (long long)v);
      }
      if (c == i) de = e;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_xor(s, o);
      const int64_t od = __shfl_xor(de, o);
      de = od > de ? od : de;
    }
    if (lane == 0) {
      anc[i] = s;
      diag[i] = de >= 0 ? (int64_t)data[de] : 0;
    }
  }
}

// The canonical order of two vertices.
struct OrdLess {
  const int32_t* comp;
  const int32_t* csize;
  const int64_t* anc;
  const int64_t* maxrel;
  const int64_t* diag;
  __device__ bool operator()(const int32_t a, const int32_t b) const {
    if (csize[a] != csize[b]) return csize[a] > csize[b];
    if (comp[a] != comp[b]) return comp[a] < comp[b];
    if (anc[a] != anc[b]) return anc[a] > anc[b];
    if (maxrel[a] != maxrel[b]) return maxrel[a] < maxrel[b];
    if (diag[a] != diag[b]) return diag[a] > diag[b];
    return a < b;
  }
};

// B[i][j] = A[p[i]][p[j]] (0 below the diagonal when `upper`)
template <typename T>
__global__ __launch_bounds__(ORD_NT) void k_permute_dense(const T* __restrict__ A, int64_t n,
                                                          const int32_t* __restrict__ p, T* __restrict__ B,
                                                          int upper) {
  const int64_t j = (int64_t)blockIdx.x * ORD_NT + threadIdx.x;
  if (j >= n) return;
  const int64_t pj = p[j];
  for (int64_t i = blockIdx.y; i < n; i += gridDim.y)
    B[i * n + j] = (upper && j < i) ? (T)0 : A[(int64_t)p[i] * n + pj];
}

__global__ void k_perm_inverse(const int32_t* __restrict__ p, int32_t n, int32_t* __restrict__ pinv,
                               int64_t* __restrict__ lens, const int64_t* __restrict__ indptr) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t old = p[k];
    pinv[old] = (int32_t)k;
    lens[k] = indptr[old + 1] - indptr[old];
  }
}

// One wave per new row i, one pass: row p[i]'s columns renamed through pinv
// into LDS, then every entry written straight to its sorted place (its rank:
// the row's smaller columns, and equal ones earlier in the row, so ranks are
// a permutation of the row even with repeated columns).  A row longer than
// PS_MAX is copied renamed to the staging buffer and listed (its output range)
// for the segmented sort.
constexpr int PS_MAX = 256;
template <typename T>
__global__ __launch_bounds__(ORD_NT) void k_permute_csr_fused(const int64_t* __restrict__ indptr,
                                                              const int32_t* __restrict__ indices,
                                                              const T* __restrict__ data, int32_t n,
                                                              const int32_t* __restrict__ p,
                                                              const int32_t* __restrict__ pinv,
                                                              const int64_t* __restrict__ out_ptr,
                                                              int32_t* __restrict__ out_idx, T* __restrict__ out_val,
                                                              int32_t* __restrict__ tmp_idx, T* __restrict__ tmp_val,
                                                              int64_t* __restrict__ long_begin,
                                                              int64_t* __restrict__ long_end, int32_t* long_count) {
  __shared__ int32_t keys[ORD_NT / 64][PS_MAX];
  auto wsync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t* kw = keys[wv];
  const int64_t waves = (int64_t)gridDim.x * (ORD_NT / 64);
  for (int64_t i = (int64_t)blockIdx.x * (ORD_NT / 64) + wv; i < n; i += waves) {
    const int64_t s = indptr[p[i]], len = indptr[p[i] + 1] - s, d = out_ptr[i];
    if (len > PS_MAX) {
      for (int64_t k = lane; k < len; k += 64) {
        tmp_idx[d + k] = pinv[indices[s + k]];
        tmp_val[d + k] = data[s + k];
      }
      if (lane == 0) {
        const int32_t j = atomicAdd(long_count, 1);
        long_begin[j] = d;
        long_end[j] = d + len;
      }
      continue;
    }
    const int L = (int)len;
    for (int k = lane; k < L; k += 64) kw[k] = pinv[indices[s + k]];
    wsync();
    for (int k = lane; k < L; k += 64) {
      const int32_t key = kw[k];
      int r = 0;
      for (int m = 0; m < L; m++) {
        const int32_t o = kw[m];
        r += (o < key || (o == key && m < k)) ? 1 : 0;
      }
      out_idx[d + r] = key;
      out_val[d + r] = data[s + k];
    }
    wsync();  // the next row restages the keys
  }
}

// Stable LSD radix sort of the long rows (one workgroup per row, 4-bit
// digits): thread t owns the contiguous chunk t of its row, so a pass counts
// each thread's digits in its own LDS column, scans the 16 x 256 counts in
// (digit, thread) order, and scatters each chunk in order -- stable, so equal
// columns keep their row order, as the one-pass ranking does.  Passes
// alternate between the staging buffer and the output; after an even number
// the row is copied to the output.  64-bit offsets throughout: no size limit.
constexpr int LR_BITS = 4;
constexpr int LR_BINS = 1 << LR_BITS;
template <typename T>
__global__ __launch_bounds__(ORD_NT) void k_sort_long_rows(int32_t* __restrict__ sk, T* __restrict__ sv,
                                                           int32_t* __restrict__ ok, T* __restrict__ ov,
                                                           const int64_t* __restrict__ begin,
                                                           const int64_t* __restrict__ end,
                                                           const int32_t* __restrict__ count, int bits) {
  __shared__ int64_t cnt[LR_BINS][ORD_NT];
  __shared__ int64_t wsum[ORD_NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int32_t nrows = *count;
  const int passes = (bits + LR_BITS - 1) / LR_BITS;
  // the scan's share of this thread: digit dd, threads [tt0, tt0 + 16)
  const int dd = tid / (ORD_NT / LR_BINS), tt0 = (tid % (ORD_NT / LR_BINS)) * LR_BINS;
  for (int64_t rj = blockIdx.x; rj < nrows; rj += gridDim.x) {
    const int64_t b0 = begin[rj], len = end[rj] - b0;
    const int64_t chunk = (len + ORD_NT - 1) / ORD_NT;
    const int64_t c0 = (int64_t)tid * chunk < len ? (int64_t)tid * chunk : len;
    const int64_t c1 = c0 + chunk < len ? c0 + chunk : len;
    int32_t* srck = sk + b0;
    T* srcv = sv + b0;
    int32_t* dstk = ok + b0;
    T* dstv = ov + b0;
    for (int pass = 0; pass < passes; pass++) {
      const int sh = pass * LR_BITS;
#pragma unroll
      for (int d = 0; d < LR_BINS; d++) cnt[d][tid] = 0;
      for (int64_t e = c0; e < c1; e++) cnt[(srck[e] >> sh) & (LR_BINS - 1)][tid]++;
      __syncthreads();
      // exclusive scan over (digit, thread): this thread's 16 counts ...
      int64_t run = 0;
#pragma unroll
      for (int q = 0; q < LR_BINS; q++) {
        const int64_t v = cnt[dd][tt0 + q];
        cnt[dd][tt0 + q] = run;
        run += v;
      }
      // ... offset by the sum of the threads before it (a block scan)
      int64_t incl = run;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
      }
      if (lane == 63) wsum[wv] = incl;
      __syncthreads();
      int64_t base = incl - run;
      for (int w = 0; w < wv; w++) base += wsum[w];
#pragma unroll
      for (int q = 0; q < LR_BINS; q++) cnt[dd][tt0 + q] += base;
      __syncthreads();
      for (int64_t e = c0; e < c1; e++) {
        const int32_t k = srck[e];
        const int64_t pos = cnt[(k >> sh) & (LR_BINS - 1)][tid]++;
        dstk[pos] = k;
        dstv[pos] = srcv[e];
      }
      __syncthreads();  // the pass's writes are done before the next reads them
      int32_t* tk = srck;
      srck = dstk;
      dstk = tk;
      T* tv = srcv;
      srcv = dstv;
      dstv = tv;
    }
    if (srck != ok + b0) {
      for (int64_t e = tid; e < len; e += ORD_NT) {
        ok[b0 + e] = srck[e];
        ov[b0 + e] = srcv[e];
      }
    }
    __syncthreads();  // the next row reuses the counts
  }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static unsigned key_bits(int32_t n) {
  unsigned b = 1;
  while (b < 31 && ((int64_t)1 << b) < (int64_t)n) b++;
  return b;
}

template <typename T>
static size_t permute_csr_sort_bytes(int32_t n, int64_t nnz) {
  size_t sort_bytes = 0;
  (void)rocprim::segmented_radix_sort_pairs((void*)nullptr, sort_bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                            (const T*)nullptr, (T*)nullptr, (unsigned)nnz, (unsigned)n,
                                            (const int64_t*)nullptr, (const int64_t*)nullptr, 0u, key_bits(n),
                                            (hipStream_t)0);
  return sort_bytes;
}

static size_t scan_bytes(int32_t n) {
  size_t b = 0;
  (void)rocprim::inclusive_scan((void*)nullptr, b, (const int64_t*)nullptr, (int64_t*)nullptr, (size_t)n,
                                rocprim::plus<int64_t>(), (hipStream_t)0);
  return b;
}

// Long rows go to rocPRIM's segmented radix sort while its 32-bit sizes hold
// the matrix, to k_sort_long_rows beyond.
constexpr int64_t ROCPRIM_SORT_BELOW = (int64_t)1 << 32;

// workspace: pinv (n), row lengths (n), the long rows' output ranges (2 n)
// and their count, staging for the long rows (nnz), scan and sort storage
template <typename T>
static int64_t permute_csr_ws_bytes(int32_t n, int64_t nnz) {
  return (int64_t)(align256(4 * (size_t)n) + align256(8 * (size_t)n) + 2 * align256(8 * (size_t)n) + 256 +
                   align256(4 * (size_t)nnz) + align256(sizeof(T) * (size_t)nnz) + align256(scan_bytes(n)) +
                   (nnz < ROCPRIM_SORT_BELOW ? align256(permute_csr_sort_bytes<T>(n, nnz)) : 0) + 256);
}

template <typename T>
static int permute_csr(const int64_t* indptr, const int32_t* indices, const T* data, int32_t n, int64_t nnz,
                       const int32_t* p, int64_t* out_indptr, int32_t* out_indices, T* out_data, void* ws,
                       int64_t ws_bytes, hipStream_t s) {
  if (n <= 0 || nnz < 0 || !indptr || !p || !out_indptr || !ws) return RBG_E_ARG;
  if (nnz > 0 && (!indices || !data || !out_indices || !out_data)) return RBG_E_ARG;
  if (ws_bytes < permute_csr_ws_bytes<T>(n, nnz)) return RBG_E_WORKSPACE;
  char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  int32_t* pinv = (int32_t*)w;
  w += align256(4 * (size_t)n);
  int64_t* lens = (int64_t*)w;
  w += align256(8 * (size_t)n);
  int64_t* long_begin = (int64_t*)w;
  w += align256(8 * (size_t)n);
  int64_t* long_end = (int64_t*)w;
  w += align256(8 * (size_t)n);
  int32_t* long_count = (int32_t*)w;
  w += 256;
  int32_t* tmp_idx = (int32_t*)w;
  w += align256(4 * (size_t)nnz);
  T* tmp_val = (T*)w;
  w += align256(sizeof(T) * (size_t)nnz);
  void* scan_tmp = w;
  size_t scan_sz = scan_bytes(n);
  w += align256(scan_sz);
  void* sort_tmp = w;
  const bool rp = nnz < ROCPRIM_SORT_BELOW;
  size_t sort_sz = rp ? permute_csr_sort_bytes<T>(n, nnz) : 0;
  const int g = (int)((n + ORD_NT - 1) / ORD_NT < 4096 ? (n + ORD_NT - 1) / ORD_NT : 4096);
  hipLaunchKernelGGL(k_perm_inverse, dim3(g), dim3(ORD_NT), 0, s, p, n, pinv, lens, indptr);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ord_fill_i64, dim3(1), dim3(64), 0, s, out_indptr, (int64_t)1, (int64_t)0);
  RBG_CHECK_LAUNCH();
  hipError_t e = rocprim::inclusive_scan(scan_tmp, scan_sz, lens, out_indptr + 1, (size_t)n,
                                         rocprim::plus<int64_t>(), s);
  if (e != hipSuccess) return (int)e;
  if (nnz == 0) return 0;
  const int gw = (int)((n + 3) / 4 < 8192 ? (n + 3) / 4 : 8192);
  // Rows longer than PS_MAX go to rocPRIM's segmented sort from the staging
  // buffer.  Their count is known only on the device, so the sort runs over
  // the most there can be (each holds more than PS_MAX of the nnz entries),
  // the unused segments cleared to empty ones: no host round trip, and the
  // call stays capturable.
  const int64_t nseg = nnz / (PS_MAX + 1) < n ? nnz / (PS_MAX + 1) : n;
  hipLaunchKernelGGL(k_ord_fill_i32, dim3(1), dim3(64), 0, s, long_count, 1, 0);
  RBG_CHECK_LAUNCH();
  if (nseg > 0) {
    const int gf = (int)((nseg + ORD_NT - 1) / ORD_NT < 4096 ? (nseg + ORD_NT - 1) / ORD_NT : 4096);
    hipLaunchKernelGGL(k_ord_fill_i64, dim3(gf), dim3(ORD_NT), 0, s, long_begin, nseg, (int64_t)0);
    RBG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ord_fill_i64, dim3(gf), dim3(ORD_NT), 0, s, long_end, nseg, (int64_t)0);
    RBG_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL((k_permute_csr_fused<T>), dim3(gw), dim3(ORD_NT), 0, s, indptr, indices, data, n, p, pinv,
                     (const int64_t*)out_indptr, out_indices, out_data, tmp_idx, tmp_val, long_begin, long_end,
                     long_count);
  RBG_CHECK_LAUNCH();
  if (nseg == 0) return 0;
  if (!rp) {
    const int gl = (int)(nseg < 8192 ? nseg : 8192);
    hipLaunchKernelGGL((k_sort_long_rows<T>), dim3(gl), dim3(ORD_NT), 0, s, tmp_idx, tmp_val, out_indices, out_data,
                       (const int64_t*)long_begin, (const int64_t*)long_end, (const int32_t*)long_count,
                       (int)key_bits(n));
    RBG_CHECK_LAUNCH();
    return 0;
  }
  e = rocprim::segmented_radix_sort_pairs(sort_tmp, sort_sz, (const int32_t*)tmp_idx, out_indices,
                                          (const T*)tmp_val, out_data, (unsigned)nnz, (unsigned)nseg,
                                          (const int64_t*)long_begin, (const int64_t*)long_end, 0u, key_bits(n), s);
  return e == hipSuccess ? 0 : (int)e;
}

template <typename T>
static int order_metrics_dense(const T* A, int64_t n, int64_t* anc, int64_t* maxrel, int64_t* diag, hipStream_t s) {
  if (!A || n <= 0 || !anc || !maxrel || !diag) return RBG_E_ARG;
  hipLaunchKernelGGL(k_ord_fill_i64, dim3(64), dim3(ORD_NT), 0, s, maxrel, n, (int64_t)0);
  RBG_CHECK_LAUNCH();
  const int rows = (int)(n < 8192 ? n : 8192);
  hipLaunchKernelGGL((k_ord_rows_dense<T>), dim3(rows), dim3(ORD_NT), 0, s, A, n, anc, diag);
  RBG_CHECK_LAUNCH();
  const unsigned cx = (unsigned)((n + ORD_NT - 1) / ORD_NT), cy = (unsigned)((n + ORD_COL_ROWS - 1) / ORD_COL_ROWS);
  if (cy > 65535) return RBG_E_ARG;
  hipLaunchKernelGGL((k_ord_cols_dense<T>), dim3(cx, cy), dim3(ORD_NT), 0, s, A, n, maxrel);
  RBG_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int order_metrics_csr(const int64_t* indptr, const int32_t* indices, const T* data, int32_t n, int64_t* anc,
                             int64_t* maxrel, int64_t* diag, hipStream_t s) {
  if (!indptr || n <= 0 || !anc || !maxrel || !diag) return RBG_E_ARG;
  hipLaunchKernelGGL(k_ord_fill_i64, dim3(64), dim3(ORD_NT), 0, s, maxrel, (int64_t)n, (int64_t)0);
  RBG_CHECK_LAUNCH();
  const int g = (int)((n + 3) / 4 < 8192 ? (n + 3) / 4 : 8192);  // four rows (waves) per workgroup
  hipLaunchKernelGGL((k_ord_csr<T>), dim3(g), dim3(ORD_NT), 0, s, indptr, indices, data, n, anc, maxrel, diag);
  RBG_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int permute_dense(const T* A, int64_t n, const int32_t* p, T* B, int upper, hipStream_t s) {
  if (!A || !B || !p || n <= 0 || A == B) return RBG_E_ARG;
  const unsigned gx = (unsigned)((n + ORD_NT - 1) / ORD_NT);
  const unsigned gy = (unsigned)(n < 16384 ? n : 16384);
  hipLaunchKernelGGL((k_permute_dense<T>), dim3(gx, gy), dim3(ORD_NT), 0, s, A, n, p, B, upper);
  RBG_CHECK_LAUNCH();
  return 0;
}

static size_t order_sort_bytes(int32_t n) {
  size_t b = 0;
  (void)rocprim::merge_sort((void*)nullptr, b, rocprim::counting_iterator<int32_t>(0), (int32_t*)nullptr,
                            (size_t)n, OrdLess{}, (hipStream_t)0);
  return b;
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int rbg_order_metrics_dense_i32(const int32_t* A, int64_t n, int64_t* anc, int64_t* maxrel, int64_t* diag,
                                rbg_stream_t stream) {
  return order_metrics_dense<int32_t>(A, n, anc, maxrel, diag, (hipStream_t)stream);
}
int rbg_order_metrics_dense_i64(const int64_t* A, int64_t n, int64_t* anc, int64_t* maxrel, int64_t* diag,
                                rbg_stream_t stream) {
  return order_metrics_dense<int64_t>(A, n, anc, maxrel, diag, (hipStream_t)stream);
}
int rbg_order_metrics_csr_i32(const int64_t* indptr, const int32_t* indices, const int32_t* data, int32_t n,
                              int64_t* anc, int64_t* maxrel, int64_t* diag, rbg_stream_t stream) {
  return order_metrics_csr<int32_t>(indptr, indices, data, n, anc, maxrel, diag, (hipStream_t)stream);
}
int rbg_order_metrics_csr_i64(const int64_t* indptr, const int32_t* indices, const int64_t* data, int32_t n,
                              int64_t* anc, int64_t* maxrel, int64_t* diag, rbg_stream_t stream) {
  return order_metrics_csr<int64_t>(indptr, indices, data, n, anc, maxrel, diag, (hipStream_t)stream);
}

int64_t rbg_order_sort_temp_bytes(int32_t n) { return n <= 0 ? 0 : (int64_t)order_sort_bytes(n); }

int rbg_order_sort(int32_t n, const int32_t* comp, const int32_t* comp_size, const int64_t* anc,
                   const int64_t* maxrel, const int64_t* diag, int32_t* perm, void* temp, int64_t temp_bytes,
                   rbg_stream_t stream) {
  if (n <= 0 || !comp || !comp_size || !anc || !maxrel || !diag || !perm || !temp) return RBG_E_ARG;
  size_t b = order_sort_bytes(n);
  if (temp_bytes < (int64_t)b) return RBG_E_WORKSPACE;
  hipError_t e = rocprim::merge_sort(temp, b, rocprim::counting_iterator<int32_t>(0), perm, (size_t)n,
                                     OrdLess{comp, comp_size, anc, maxrel, diag}, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (int)e;
}

int rbg_permute_dense_i32(const int32_t* A, int64_t n, const int32_t* p, int32_t* B, int32_t upper,
                          rbg_stream_t stream) {
  return permute_dense<int32_t>(A, n, p, B, upper, (hipStream_t)stream);
}
int rbg_permute_dense_i64(const int64_t* A, int64_t n, const int32_t* p, int64_t* B, int32_t upper,
                          rbg_stream_t stream) {
  return permute_dense<int64_t>(A, n, p, B, upper, (hipStream_t)stream);
}

int64_t rbg_permute_csr_workspace_bytes_i32(int32_t n, int64_t nnz) { return permute_csr_ws_bytes<int32_t>(n, nnz); }
int64_t rbg_permute_csr_workspace_bytes_i64(int32_t n, int64_t nnz) { return permute_csr_ws_bytes<int64_t>(n, nnz); }

int rbg_permute_csr_i32(const int64_t* indptr, const int32_t* indices, const int32_t* data, int32_t n, int64_t nnz,
                        const int32_t* p, int64_t* out_indptr, int32_t* out_indices, int32_t* out_data,
                        void* workspace, int64_t ws_bytes, rbg_stream_t stream) {
  return permute_csr<int32_t>(indptr, indices, data, n, nnz, p, out_indptr, out_indices, out_data, workspace,
                              ws_bytes, (hipStream_t)stream);
}
int rbg_permute_csr_i64(const int64_t* indptr, const int32_t* indices, const int64_t* data, int32_t n, int64_t nnz,
                        const int32_t* p, int64_t* out_indptr, int32_t* out_indices, int64_t* out_data,
                        void* workspace, int64_t ws_bytes, rbg_stream_t stream) {
  return permute_csr<int64_t>(indptr, indices, data, n, nnz, p, out_indptr, out_indices, out_data, workspace,
                              ws_bytes, (hipStream_t)stream);
}

}  // extern "C"
