// Element-wise AVOS kernels, the dense AVOS matrix product and the
// relational-composition updates.
//
// References:
//   element-wise   redblackgraph/gpu/avos_kernels.py:150-173 (one value per lane)
//   dense product  redblackgraph/core/src/redblackgraph/redblack.c.src:554-700 ('ij,jk')
//   composition    redblackgraph/core/src/redblackgraph/relational_composition.c.in:51-267
#include "rbg_common.hpp"

namespace rbg {

constexpr int EW_NT = 256;

template <typename T>
__global__ void k_avos_sum(const T* __restrict__ x, const T* __restrict__ y, T* __restrict__ z, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    z[i] = avos_sum<T>(x[i], y[i]);
}

template <typename T>
__global__ void k_avos_product(const T* __restrict__ x, const T* __restrict__ y, T* __restrict__ z, int64_t n,
                               rbg_status* st) {
  uint32_t evs = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t ev = 0;
    z[i] = avos_product<T>(x[i], y[i], ev);
    if (ev && st) record_overflow(st, (uint64_t)i, ev);
    evs |= ev;
  }
  (void)evs;
}

static inline int ew_grid(int64_t n) {
  int64_t g = (n + EW_NT * 4 - 1) / (EW_NT * 4);
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

// ---------------------------------------------------------------------------
// Dense AVOS product, general arithmetic.  64x64 output tile per workgroup,
// 4x4 cells per lane, K staged through LDS in chunks of 32.
// ---------------------------------------------------------------------------
constexpr int MM_T = 64;
constexpr int MM_K = 32;

template <typename T>
__global__ __launch_bounds__(256) void k_matmul(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                int64_t ldb, T* __restrict__ C, int64_t ldc, int64_t M,
                                                int64_t K, int64_t N, int width, int is_uns, rbg_status* st) {
  __shared__ T As[MM_K][MM_T + 1];
  __shared__ T Bs[MM_K][MM_T + 1];
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.y * MM_T, k0 = (int64_t)blockIdx.x * MM_T;
  const int tr = tid >> 4, tc = tid & 15;
  T acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = 0;
  uint32_t evacc = 0;
  uint64_t evkey = 0;
  for (int64_t j0 = 0; j0 < K; j0 += MM_K) {
    for (int e = tid; e < MM_K * MM_T; e += 256) {
      int jj = e % MM_K, rr = e / MM_K;  // A: row rr, col jj
      int64_t gi = i0 + rr, gj = j0 + jj;
      As[jj][rr] = (gi < M && gj < K) ? A[gi * lda + gj] : (T)0;
      int cc = e % MM_T, jb = e / MM_T;  // B: row jb, col cc
      int64_t gk = k0 + cc, gj2 = j0 + jb;
      Bs[jb][cc] = (gj2 < K && gk < N) ? B[gj2 * ldb + gk] : (T)0;
    }
    __syncthreads();
    for (int jj = 0; jj < MM_K; jj++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        T a = As[jj][tr * 4 + r];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          uint32_t ev = 0;
          T p = avos_product<T>(a, Bs[jj][tc * 4 + c], width, is_uns != 0, ev);
          if (ev) {
            evacc |= ev;
            uint64_t key = (((uint64_t)(i0 + tr * 4 + r) * (uint64_t)N) + (uint64_t)(k0 + tc * 4 + c)) *
                               (uint64_t)K + (uint64_t)(j0 + jj);
            evkey = key > evkey ? key : evkey;
          }
          acc[r][c] = avos_sum<T>(acc[r][c], p);
        }
      }
    }
    __syncthreads();
  }
  if (evacc) record_overflow(st, evkey, evacc);
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int64_t gi = i0 + tr * 4 + r, gk = k0 + tc * 4 + c;
      if (gi < M && gk < N) C[gi * ldc + gk] = acc[r][c];
    }
}

// ---------------------------------------------------------------------------
// Relational composition.
// ---------------------------------------------------------------------------
// u[p] = (+)_i lhs_i (x) R[i][p], lhs = row `arow` of R with column `beta`
// replaced by `np` when np != 0 (vector_avos, relational_composition.c.in:52-69).
// With arow < 0 the left vector is `uvec`.
template <typename T>
__global__ void k_vecmat(const T* __restrict__ R, int64_t n, int64_t arow, int64_t beta, T np,
                         const T* __restrict__ uvec, T* __restrict__ out, int width, int is_uns, rbg_status* st) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  T r = 0;
  uint32_t evs = 0;
  for (int64_t i = 0; i < n; i++) {
    T lhs = arow >= 0 ? R[arow * n + i] : uvec[i];
    if (np != 0 && i == beta) lhs = np;
    uint32_t ev = 0;
    r = avos_sum<T>(r, avos_product<T>(lhs, R[i * n + p], width, is_uns != 0, ev));
    evs |= ev;
  }
  if (evs) atomicOr(&st->events, evs);
  out[p] = r;
}

// v'[a] = (+)_i R[a][i] (x) v[i]: one wave per row, lanes stride the columns.
template <typename T>
__global__ void k_matvec(const T* __restrict__ R, int64_t n, const T* __restrict__ v, T* __restrict__ out,
                         int64_t out_stride, int width, int is_uns, rbg_status* st) {
  const int lane = threadIdx.x & 63;
  const int64_t a = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  if (a >= n) return;
  T r = 0;
  uint32_t evs = 0;
  for (int64_t i = lane; i < n; i += 64) {
    uint32_t ev = 0;
    r = avos_sum<T>(r, avos_product<T>(R[a * n + i], v[i], width, is_uns != 0, ev));
    evs |= ev;
  }
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(r, off, 64);
    r = avos_sum<T>(r, o);
  }
  if (evs) atomicOr(&st->events, evs);
  if (lane == 0) out[a * out_stride] = r;
}

// Keys for combining AVOS sums with an unsigned atomicMin: RED_ONE (-1) maps
// to 0 (absorbs every sum), "absent" (0) to all ones (loses every sum), any
// other value to itself read as unsigned (the sum of two such values is their
// unsigned minimum).
template <typename T>
__device__ inline typename unsigned_of<T>::type sum_key(T v) {
  using U = typename unsigned_of<T>::type;
  return v == (T)-1 ? (U)0 : (v == 0 ? ~(U)0 : (U)v);
}
template <typename T>
__device__ inline T sum_unkey(typename unsigned_of<T>::type k) {
  using U = typename unsigned_of<T>::type;
  return k == 0 ? (T)-1 : (k == ~(U)0 ? (T)0 : (T)k);
}

// Edge composition, step 1 (relational_composition.c.in:52-69, 92-101): the
// composed row u = lhs (x) R, lhs = row alpha of R with entry beta replaced by
// the relationship.  u is accumulated as sum keys in row alpha of the output.
// k_edge_begin also flags a cycle when R[beta][alpha] != 0 (the loop check of
// core/redblack.py:93-95).
template <typename T>
__global__ void k_edge_begin(const T* __restrict__ R, int64_t n, int64_t alpha, int64_t beta,
                             typename unsigned_of<T>::type* __restrict__ ukey, rbg_status* st) {
  using U = typename unsigned_of<T>::type;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    ukey[p] = ~(U)0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && R[beta * n + alpha] != 0) atomicOr(&st->flags, (uint32_t)RBG_F_CYCLE);
}

constexpr int EC_ROWS = 256;  // rows of R per workgroup in the composed-row product

// Workgroup (x, y): columns [256x, 256x + 256), rows [256y, 256y + 256) of R.
// The nonzero left operands of the row chunk are gathered in LDS (pedigree
// rows are sparse: few ancestors), then every lane folds its column.
template <typename T>
__global__ __launch_bounds__(256) void k_edge_row(const T* __restrict__ R, int64_t n, int64_t alpha, int64_t beta,
                                                  T np, typename unsigned_of<T>::type* __restrict__ ukey, int width,
                                                  int is_uns, rbg_status* st) {
  __shared__ int32_t s_idx[EC_ROWS];
  __shared__ T s_lhs[EC_ROWS];
  __shared__ int s_cnt;
  const int tid = threadIdx.x;
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.y * EC_ROWS + tid;
  if (i < n) {
    T lhs = R[alpha * n + i];
    if (np != 0 && i == beta) lhs = np;
    if (lhs != 0) {
      int q = atomicAdd(&s_cnt, 1);
      s_idx[q] = (int32_t)i;
      s_lhs[q] = lhs;
    }
  }
  __syncthreads();
  const int cnt = s_cnt;
  if (cnt == 0) return;
  const int64_t p = (int64_t)blockIdx.x * 256 + tid;
  if (p >= n) return;
  T acc = 0;
  uint32_t evs = 0;
  for (int k = 0; k < cnt; k++) {
    uint32_t ev = 0;
    acc = avos_sum<T>(acc, avos_product<T>(s_lhs[k], R[(int64_t)s_idx[k] * n + p], width, is_uns != 0, ev));
    evs |= ev;
  }
  if (acc != 0) atomicMin(&ukey[p], sum_key<T>(acc));
  if (evs) atomicOr(&st->events, evs);
}

template <typename T>
__global__ void k_edge_row_end(typename unsigned_of<T>::type* __restrict__ ukey, int64_t n) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    ukey[p] = (typename unsigned_of<T>::type)sum_unkey<T>(ukey[p]);
}

// Edge composition, step 2 (relational_composition.c.in:110-137):
// out[m][p] = u[p] ? R[m][p] (+) R[m][alpha] (x) u[p] : R[m][p] for m != alpha
// (row alpha already holds u).  One workgroup row per matrix row, four
// consecutive columns per lane.
template <typename T>
__global__ __launch_bounds__(256) void k_edge_update(const T* __restrict__ R, int64_t n, int64_t alpha,
                                                     T* __restrict__ out, int width, int is_uns, rbg_status* st) {
  const T* u = out + alpha * n;
  uint32_t evs = 0;
  for (int64_t m = blockIdx.y; m < n; m += gridDim.y) {
    if (m == alpha) continue;
    const T rma = R[m * n + alpha];
    const T* Rm = R + m * n;
    T* Om = out + m * n;
    for (int64_t p0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4; p0 < n; p0 += (int64_t)gridDim.x * 256 * 4) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int64_t p = p0 + q;
        if (p < n) {
          const T rv = Rm[p];
          T o = rv;
          if (rma != 0) {
            const T up = u[p];
            if (up != 0) {
              uint32_t ev = 0;
              o = avos_sum<T>(rv, avos_product<T>(rma, up, width, is_uns != 0, ev));
              evs |= ev;
            }
          }
          Om[p] = o;
        }
      }
    }
  }
  if (evs) atomicOr(&st->events, evs);
}

__global__ void k_zero_i64(int64_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// In-place edge composition, step 2: the rows m with W[m][alpha] != 0 (the
// vertices that reach alpha) are the only rows the update changes, and within
// them only the columns where u is present.  The first kernel lists those
// rows (one read of column alpha), the second updates them in place; every
// other entry is left untouched, so a composition into a pedigree closure
// costs its affected rows rather than the whole matrix.
template <typename T>
__global__ __launch_bounds__(256) void k_edge_rows_hit(const T* __restrict__ W, int64_t n, int64_t alpha,
                                                       int64_t* __restrict__ hit, int64_t* __restrict__ nhit) {
  __shared__ int s_n;
  __shared__ int64_t s_base;
  __shared__ int64_t s_m[256];
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m < n && m != alpha && W[m * n + alpha] != 0) s_m[atomicAdd(&s_n, 1)] = m;
  __syncthreads();
  if (threadIdx.x == 0 && s_n) s_base = (int64_t)atomicAdd((unsigned long long*)nhit, (unsigned long long)s_n);
  __syncthreads();
  if (threadIdx.x < s_n) hit[s_base + threadIdx.x] = s_m[threadIdx.x];
}

// One workgroup per listed row, four consecutive columns per lane; row alpha
// takes u from the scratch row once every product is formed from it.
template <typename T>
__global__ __launch_bounds__(256) void k_edge_update_rows(T* __restrict__ W, int64_t n, int64_t alpha,
                                                          const T* __restrict__ u, const int64_t* __restrict__ hit,
                                                          const int64_t* __restrict__ nhit, int width, int is_uns,
                                                          rbg_status* st) {
  uint32_t evs = 0;
  const int64_t nh = *nhit;
  for (int64_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const int64_t m = hit[h];
    T* Wm = W + m * n;
    const T rma = Wm[alpha];  // unchanged by the update: u[alpha] is alpha's own diagonal
    for (int64_t p0 = (int64_t)threadIdx.x * 4; p0 < n; p0 += 256 * 4) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int64_t p = p0 + q;
        if (p < n) {
          const T up = u[p];
          if (up != 0) {
            uint32_t ev = 0;
            Wm[p] = avos_sum<T>(Wm[p], avos_product<T>(rma, up, width, is_uns != 0, ev));
            evs |= ev;
          }
        }
      }
    }
  }
  if (evs) atomicOr(&st->events, evs);
}

template <typename T>
__global__ void k_edge_put_row(T* __restrict__ W, int64_t n, int64_t alpha, const T* __restrict__ u) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    W[alpha * n + p] = u[p];
}

// Vertex composition, steps 3-4 (relational_composition.c.in:223-266).
template <typename T>
__global__ void k_vertex_update(const T* __restrict__ R, int64_t n, T color, T* __restrict__ out, int width,
                                int is_uns, rbg_status* st) {
  const int64_t m1 = n + 1;
  int64_t total = m1 * m1;
  uint32_t evs = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t a = e / m1, b = e % m1;
    if (a == n && b == n) {
      out[e] = color;
    } else if (a < n && b < n) {
      uint32_t ev = 0;
      T vv = out[a * m1 + n], uu = out[n * m1 + b];
      out[e] = avos_sum<T>(R[a * n + b], avos_product<T>(vv, uu, width, is_uns != 0, ev));
      evs |= ev;
    }
  }
  if (evs) atomicOr(&st->events, evs);
}

template <typename T>
__global__ void k_vertex_check(const T* __restrict__ out, int64_t n, rbg_status* st) {
  const int64_t m1 = n + 1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (out[i * m1 + n] != 0 && out[n * m1 + i] != 0) {
      atomicOr(&st->flags, (uint32_t)RBG_F_CYCLE);
      atomicMin((unsigned long long*)&st->cycle_first_key, (unsigned long long)i);
    }
  }
}

template <typename T>
__global__ void k_vertex_fill_aux(const T* __restrict__ out, int64_t n, rbg_status* st) {
  if (st->flags & RBG_F_CYCLE) {
    int64_t i = (int64_t)st->cycle_first_key;
    st->aux[0] = i;
    st->aux[1] = (int64_t)out[n * (n + 1) + i];
    st->aux[2] = (int64_t)out[i * (n + 1) + n];
  }
}

template <typename T>
static int matmul(const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t M, int64_t K,
                  int64_t N, int width, int is_uns, rbg_status* st, hipStream_t s) {
  if (!A || !B || !C || !st || M < 0 || K < 0 || N < 0) return RBG_E_ARG;
  if (M == 0 || N == 0) return 0;
  dim3 grid((unsigned)((N + MM_T - 1) / MM_T), (unsigned)((M + MM_T - 1) / MM_T));
  hipLaunchKernelGGL((k_matmul<T>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, K, N, width, is_uns, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int edge_compose(const T* R, int64_t n, int64_t alpha, int64_t beta, T rel, T* out, int width, int is_uns,
                        rbg_status* st, hipStream_t s) {
  if (!R || !out || !st || n <= 0 || alpha < 0 || alpha >= n || beta < 0 || beta >= n || R == out) return RBG_E_ARG;
  using U = typename unsigned_of<T>::type;
  U* ukey = (U*)(out + alpha * n);  // row alpha of the output holds u
  const unsigned cols = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL((k_edge_begin<T>), dim3(cols), dim3(256), 0, s, R, n, alpha, beta, ukey, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_edge_row<T>), dim3(cols, (unsigned)((n + EC_ROWS - 1) / EC_ROWS)), dim3(256), 0, s, R, n,
                     alpha, beta, rel, ukey, width, is_uns, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_edge_row_end<T>), dim3(cols), dim3(256), 0, s, ukey, n);
  RBG_CHECK_LAUNCH();
  const unsigned gx = (unsigned)((n + 1023) / 1024);
  const unsigned gy = (unsigned)(n < 65535 ? n : 65535);
  hipLaunchKernelGGL((k_edge_update<T>), dim3(gx, gy), dim3(256), 0, s, R, n, alpha, out, width, is_uns, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

// scratch: rbg_edge_compose_scratch_bytes(n) bytes (the composed row and the
// list of affected rows)
template <typename T>
static int edge_compose_inplace(T* W, int64_t n, int64_t alpha, int64_t beta, T rel, void* scratch, int width,
                                int is_uns, rbg_status* st, hipStream_t s) {
  if (!W || !scratch || !st || n <= 0 || alpha < 0 || alpha >= n || beta < 0 || beta >= n) return RBG_E_ARG;
  using U = typename unsigned_of<T>::type;
  char* base = (char*)(((uintptr_t)scratch + 255) & ~(uintptr_t)255);
  U* ukey = (U*)base;
  int64_t* nhit = (int64_t*)(base + ((n * (int64_t)sizeof(T) + 255) & ~(int64_t)255));
  int64_t* hit = nhit + 32;
  const unsigned cols = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL((k_edge_begin<T>), dim3(cols), dim3(256), 0, s, W, n, alpha, beta, ukey, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(32), 0, s, nhit, (int64_t)1);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_edge_row<T>), dim3(cols, (unsigned)((n + EC_ROWS - 1) / EC_ROWS)), dim3(256), 0, s, W, n,
                     alpha, beta, rel, ukey, width, is_uns, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_edge_row_end<T>), dim3(cols), dim3(256), 0, s, ukey, n);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_edge_rows_hit<T>), dim3(cols), dim3(256), 0, s, W, n, alpha, hit, nhit);
  RBG_CHECK_LAUNCH();
  const unsigned grid = (unsigned)(n < 2048 ? n : 2048);
  hipLaunchKernelGGL((k_edge_update_rows<T>), dim3(grid), dim3(256), 0, s, W, n, alpha, (const T*)ukey, hit, nhit,
                     width, is_uns, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_edge_put_row<T>), dim3(cols), dim3(256), 0, s, W, n, alpha, (const T*)ukey);
  RBG_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int vertex_compose(const T* u, const T* R, const T* v, T color, int64_t n, T* out, int width, int is_uns,
                          rbg_status* st, hipStream_t s) {
  if (!u || !R || !v || !out || !st || n < 0) return RBG_E_ARG;
  const int64_t m1 = n + 1;
  if (n > 0) {
    // new row n: u (x) R
    hipLaunchKernelGGL((k_vecmat<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R, n, (int64_t)-1,
                       (int64_t)-1, (T)0, u, out + n * m1, width, is_uns, st);
    RBG_CHECK_LAUNCH();
    // new column n: R (x) v
    hipLaunchKernelGGL((k_matvec<T>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, R, n, v, out + n, m1, width,
                       is_uns, st);
    RBG_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_vertex_check<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const T*)out, n, st);
    RBG_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL((k_vertex_update<T>), dim3((unsigned)ew_grid(m1 * m1)), dim3(EW_NT), 0, s, R, n, color, out,
                     width, is_uns, st);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_vertex_fill_aux<T>), dim3(1), dim3(1), 0, s, (const T*)out, n, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int rbg_avos_sum_i32(const int32_t* x, const int32_t* y, int32_t* z, int64_t n, rbg_stream_t stream) {
  if (n < 0 || (n && (!x || !y || !z))) return RBG_E_ARG;
  if (!n) return 0;
  hipLaunchKernelGGL((k_avos_sum<int32_t>), dim3(ew_grid(n)), dim3(EW_NT), 0, (hipStream_t)stream, x, y, z, n);
  RBG_CHECK_LAUNCH();
  return 0;
}
int rbg_avos_sum_i64(const int64_t* x, const int64_t* y, int64_t* z, int64_t n, rbg_stream_t stream) {
  if (n < 0 || (n && (!x || !y || !z))) return RBG_E_ARG;
  if (!n) return 0;
  hipLaunchKernelGGL((k_avos_sum<int64_t>), dim3(ew_grid(n)), dim3(EW_NT), 0, (hipStream_t)stream, x, y, z, n);
  RBG_CHECK_LAUNCH();
  return 0;
}
int rbg_avos_product_i32(const int32_t* x, const int32_t* y, int32_t* z, int64_t n, rbg_status* st,
                         rbg_stream_t stream) {
  if (n < 0 || (n && (!x || !y || !z))) return RBG_E_ARG;
  if (!n) return 0;
  hipLaunchKernelGGL((k_avos_product<int32_t>), dim3(ew_grid(n)), dim3(EW_NT), 0, (hipStream_t)stream, x, y, z, n,
                     st);
  RBG_CHECK_LAUNCH();
  return 0;
}
int rbg_avos_product_i64(const int64_t* x, const int64_t* y, int64_t* z, int64_t n, rbg_status* st,
                         rbg_stream_t stream) {
  if (n < 0 || (n && (!x || !y || !z))) return RBG_E_ARG;
  if (!n) return 0;
  hipLaunchKernelGGL((k_avos_product<int64_t>), dim3(ew_grid(n)), dim3(EW_NT), 0, (hipStream_t)stream, x, y, z, n,
                     st);
  RBG_CHECK_LAUNCH();
  return 0;
}

int64_t rbg_host_avos_sum_i64(int64_t a, int64_t b) { return avos_sum<int64_t>(a, b); }

uint64_t rbg_host_avos_product_u64(int64_t a, int64_t b, uint32_t* events) {
  // c_avos_product_impl: 64-bit unsigned rules; values >= 2^63 are legal here
  uint32_t ev = 0;
  uint64_t x = (uint64_t)a, y = (uint64_t)b;
  uint64_t r;
  if (x == 0 || y == 0) r = 0;
  else if (a == -1 && b == -1) r = x;
  else if (a == 1 && b == 1) r = 1;
  else if ((a == -1 && b == 1) || (a == 1 && b == -1)) r = 0;
  else {
    if (a == -1) x = 1;
    if (b == -1) r = (a & 1) ? 0 : x;
    else if (b == 1) r = (a & 1) ? x : 0;
    else {
      int m = msb_u(y), mx = msb_u(x);
      if (mx + m >= 64) ev |= EV_OVERFLOW_WIDTH;
      r = (y & ((1ull << m) - 1)) | (x << m);
      if (r == ~0ull) ev |= EV_OVERFLOW_RED;
    }
  }
  if (events) *events = ev;
  return r;
}

int rbg_matmul_i32(const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb, int32_t* C, int64_t ldc, int64_t M,
                   int64_t K, int64_t N, int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return matmul<int32_t>(A, lda, B, ldb, C, ldc, M, K, N, width, is_unsigned, st, (hipStream_t)stream);
}
int rbg_matmul_i64(const int64_t* A, int64_t lda, const int64_t* B, int64_t ldb, int64_t* C, int64_t ldc, int64_t M,
                   int64_t K, int64_t N, int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return matmul<int64_t>(A, lda, B, ldb, C, ldc, M, K, N, width, is_unsigned, st, (hipStream_t)stream);
}
int rbg_edge_compose_i32(const int32_t* R, int64_t n, int64_t alpha, int64_t beta, int32_t rel, int32_t* out,
                         int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return edge_compose<int32_t>(R, n, alpha, beta, rel, out, width, is_unsigned, st, (hipStream_t)stream);
}
int rbg_edge_compose_i64(const int64_t* R, int64_t n, int64_t alpha, int64_t beta, int64_t rel, int64_t* out,
                         int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return edge_compose<int64_t>(R, n, alpha, beta, rel, out, width, is_unsigned, st, (hipStream_t)stream);
}
int64_t rbg_edge_compose_scratch_bytes(int64_t n) {
  return n <= 0 ? 0 : 256 + ((n * 8 + 255) & ~(int64_t)255) + 8 * 32 + 8 * n;
}
int rbg_edge_compose_inplace_i32(int32_t* W, int64_t n, int64_t alpha, int64_t beta, int32_t rel, void* scratch,
                                 int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return edge_compose_inplace<int32_t>(W, n, alpha, beta, rel, scratch, width, is_unsigned, st, (hipStream_t)stream);
}
int rbg_edge_compose_inplace_i64(int64_t* W, int64_t n, int64_t alpha, int64_t beta, int64_t rel, void* scratch,
                                 int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return edge_compose_inplace<int64_t>(W, n, alpha, beta, rel, scratch, width, is_unsigned, st, (hipStream_t)stream);
}
int rbg_vertex_compose_i32(const int32_t* u, const int32_t* R, const int32_t* v, int32_t color, int64_t n,
                           int32_t* out, int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return vertex_compose<int32_t>(u, R, v, color, n, out, width, is_unsigned, st, (hipStream_t)stream);
}
int rbg_vertex_compose_i64(const int64_t* u, const int64_t* R, const int64_t* v, int64_t color, int64_t n,
                           int64_t* out, int32_t width, int32_t is_unsigned, rbg_status* st, rbg_stream_t stream) {
  return vertex_compose<int64_t>(u, R, v, color, n, out, width, is_unsigned, st, (hipStream_t)stream);
}

}  // extern "C"
