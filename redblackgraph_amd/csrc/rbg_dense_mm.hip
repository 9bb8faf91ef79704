// Dense AVOS product C = A (x) B, fast path: the phase-3 tile of the
// Floyd-Warshall kernels (fw_kernels.hpp) run over every k of the product.
//
// Reference: redblack.c.src:554-700 (the 'ij,jk' einsum loop: C[i][k] =
// (+)_j A[i][j] (x) B[j][k] with the C core's avos_product / avos_sum,
// rbg_math.c.in:30-121).
//
// Domain (rbg_matmul_precheck_*): every value >= -1, no off-diagonal +-1 and
// every diagonal value in {-1, 0, 1}, in A and in B -- closures of pedigrees
// are.  Then every product of two off-diagonal entries is either absent or
// ((x - 1) << MSB(y)) + y, and in the E domain (rbg_common.hpp) a term is one
// shift-or and the sum one unsigned min, as in phase 3.  The terms with a
// diagonal operand, A[i][i] (x) B[i][k] and A[i][k] (x) B[k][k], are left out
// of the tile loop (the diagonals are staged as absent) and added by the
// epilogue with the full arithmetic.  A per-k MSB guard (largest left-operand
// MSB plus largest right-operand shift <= width - 2) proves that no term can
// overflow; otherwise RBG_F_NEED_GENERAL asks for the general kernel.
#include "fw_kernels.hpp"

namespace rbg {

// E-encoded operands: AT = A transposed (Kp x Mp), BE = B (Kp x Np); diagonal
// cells and the padding beyond M, K, N are absent (all ones).
template <typename T, bool TRANSPOSE>
__global__ void k_mm_prep(const T* __restrict__ X, int64_t ld, int64_t rows, int64_t cols, T* __restrict__ out,
                          int64_t out_rows, int64_t out_cols) {
  using U = typename unsigned_of<T>::type;
  const int64_t total = out_rows * out_cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / out_cols, c = e % out_cols;  // out[r][c] = X[c][r] (transposed) or X[r][c]
    const int64_t i = TRANSPOSE ? c : r, j = TRANSPOSE ? r : c;
    U v = ~(U)0;
    if (i < rows && j < cols && i != j) v = (U)e_enc<T>(X[i * ld + j]);
    out[e] = (T)v;
  }
}

// Precheck of one operand (rules of k_fw_precheck on a rows x cols matrix).
template <typename T>
__global__ void k_mm_precheck(const T* __restrict__ X, int64_t ld, int64_t rows, int64_t cols, rbg_status* st) {
  const int64_t total = rows * cols;
  bool bad = false;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / cols, j = e % cols;
    const T v = X[i * ld + j];
    if (v < (T)-1) bad = true;
    if (i != j && (v == (T)1 || v == (T)-1)) bad = true;
    if (i == j && v > (T)1) bad = true;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&st->flags, (uint32_t)RBG_F_PRECOND_FAIL);
}

// One (16 RC) x (16 RC) tile of C per workgroup, RC x RC cells per thread in
// VGPRs across all of K; per 32-k block the operands are staged in LDS as
// phase 3 stages its snapshots, the next block's loads in flight during the
// current block's relaxations.
template <typename T, int RC, bool SHIFT_BYTES>
__global__ __launch_bounds__(NT, 2) void k_mm_fast(const T* __restrict__ AT, const T* __restrict__ BE, int64_t Kp,
                                                   int64_t Mp, int64_t Np, const T* __restrict__ A, int64_t lda,
                                                   const T* __restrict__ B, int64_t ldb, T* __restrict__ C,
                                                   int64_t ldc, int64_t M, int64_t K, int64_t N, int width,
                                                   rbg_status* st) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr U ALL = ~(U)0;
  constexpr int TL = 16 * RC;
  constexpr int NCH = RC / 4;
  constexpr int GL = TL / 4;          // lanes staging one k
  constexpr int SPT = BLK * GL / NT;  // staged chunks per thread
  __shared__ U As[BLK][TL];
  using MsT = typename std::conditional<SHIFT_BYTES, uint8_t, U>::type;
  __shared__ MsT Ms[BLK][TL];
  __shared__ U Ys[BLK][TL];
  __shared__ int sGa[BLK], sGb[BLK];
  const int tid = threadIdx.x;
  const int nj = (int)(Np / TL);
  const int I = blockIdx.x / nj, J = blockIdx.x % nj;
  const int tr = tid >> 4, tc = tid & 15;
  const U* ATu = (const U*)AT;
  const U* BEu = (const U*)BE;

  U w[RC][RC];
#pragma unroll
  for (int r = 0; r < RC; r++)
#pragma unroll
    for (int c = 0; c < RC; c++) w[r][c] = ALL;
  U sa[SPT][4], sb[SPT][4];
  auto fetch = [&](int64_t kb) {
#pragma unroll
    for (int q = 0; q < SPT; q++) {
      const int e = tid + q * NT, kk = e / GL, x = (e % GL) * 4;
      ld4(ATu + (kb * BLK + kk) * Mp + (int64_t)I * TL + x, sa[q]);
      ld4(BEu + (kb * BLK + kk) * Np + (int64_t)J * TL + x, sb[q]);
    }
  };
  const int64_t nkb = Kp / BLK;
  bool bad = false;
  constexpr uint32_t SHM = 8 * sizeof(U) - 1;
  fetch(0);
  for (int64_t kb = 0; kb < nkb; kb++) {
    __syncthreads();  // the previous block's relaxations are done with the LDS
#pragma unroll
    for (int q = 0; q < SPT; q++) {
      const int e = tid + q * NT, kk = e / GL, x = (e % GL) * 4;
      U m[4], y[4];
      int ga = 0, gb = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const U b = sb[q][c];
        if ((T)b < 0) {  // absent
          m[c] = 0;
          y[c] = TOP;
        } else {
          m[c] = (U)msb_u(b);
          y[c] = b & (((U)1 << m[c]) - 1);
        }
        const int mc = (int)m[c], ac = (T)sa[q][c] > 1 ? msb_u(sa[q][c]) : 0;
        gb = mc > gb ? mc : gb;
        ga = ac > ga ? ac : ga;
      }
      if (e % GL == 0) {
        sGa[kk] = 0;
        sGb[kk] = 0;
      }
      if (ga) atomicMax(&sGa[kk], ga);
      if (gb) atomicMax(&sGb[kk], gb);
      st4(&As[kk][x], sa[q]);
      if (SHIFT_BYTES)
        *(uint32_t*)&Ms[kk][x] =
            (uint32_t)m[0] | ((uint32_t)m[1] << 8) | ((uint32_t)m[2] << 16) | ((uint32_t)m[3] << 24);
      else
        st4((U*)&Ms[kk][x], m);
      st4(&Ys[kk][x], y);
    }
    __syncthreads();
    if (kb + 1 < nkb) fetch(kb + 1);
    if (tid < BLK && sGa[tid] + sGb[tid] > width - 2) bad = true;
    for (int kk = 0; kk < BLK; kk += 2) {
      U a0[RC], a1[RC], y0[RC], y1[RC];
      uint32_t m0[RC], m1[RC];
#pragma unroll
      for (int r = 0; r < RC; r++) {
        a0[r] = As[kk][tr * RC + r];
        a1[r] = As[kk + 1][tr * RC + r];
      }
#pragma unroll
      for (int c = 0; c < RC; c++) {
        const int x = 64 * (c >> 2) + tc * 4 + (c & 3);
        if (SHIFT_BYTES) {
          m0[c] = (uint32_t)Ms[kk][x];
          m1[c] = (uint32_t)Ms[kk + 1][x];
        } else {
          m0[c] = (uint32_t)Ms[kk][x];
          m1[c] = (uint32_t)Ms[kk + 1][x];
        }
        y0[c] = Ys[kk][x];
        y1[c] = Ys[kk + 1][x];
      }
#pragma unroll
      for (int r = 0; r < RC; r++)
#pragma unroll
        for (int c = 0; c < RC; c++) {
          const U p0 = (a0[r] << (m0[c] & SHM)) | y0[c];
          const U p1 = (a1[r] << (m1[c] & SHM)) | y1[c];
          const U v1 = w[r][c] < p0 ? w[r][c] : p0;
          w[r][c] = v1 < p1 ? v1 : p1;
        }
    }
  }
  if (bad) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
  // epilogue: back to natural values, plus the two diagonal terms
  // A[i][i] (x) B[i][k] and A[i][k] (x) B[k][k] with the full arithmetic
  const int64_t r0 = (int64_t)I * TL + tr * RC;
#pragma unroll
  for (int r = 0; r < RC; r++) {
    const int64_t i = r0 + r;
    if (i >= M) continue;
    const T aii = i < K ? A[i * lda + i] : (T)0;
#pragma unroll
    for (int c = 0; c < RC; c++) {
      const int64_t k = (int64_t)J * TL + 64 * (c >> 2) + tc * 4 + (c & 3);
      if (k >= N) continue;
      T v = (T)w[r][c] < 0 ? (T)0 : (T)w[r][c];
      uint32_t ev = 0;
      if (i < K && aii != 0) v = avos_sum<T>(v, avos_product<T>(aii, B[i * ldb + k], width, false, ev));
      if (k < K) {
        const T bkk = B[k * ldb + k];
        if (bkk != 0) v = avos_sum<T>(v, avos_product<T>(A[i * lda + k], bkk, width, false, ev));
      }
      if (ev) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);  // not expected in the domain
      C[i * ldc + k] = v;
    }
  }
}

template <typename T>
struct MmTile {
  static constexpr int RC = sizeof(T) == 4 ? 8 : 4;
  static constexpr int EDGE = 16 * RC;
};

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

template <typename T>
static int64_t mm_ws_bytes(int64_t M, int64_t K, int64_t N) {
  const int64_t E = MmTile<T>::EDGE;
  const int64_t Mp = round_up(M, E), Np = round_up(N, E), Kp = round_up(K, BLK);
  return (Kp * Mp + Kp * Np) * (int64_t)sizeof(T) + 512;
}

template <typename T>
static int mm_precheck(const T* X, int64_t ld, int64_t rows, int64_t cols, rbg_status* st, hipStream_t s) {
  if (!X || !st || rows < 0 || cols < 0 || ld < cols) return RBG_E_ARG;
  if (rows == 0 || cols == 0) return 0;
  int64_t g = (rows * cols + NT * 8 - 1) / (NT * 8);
  hipLaunchKernelGGL((k_mm_precheck<T>), dim3((unsigned)(g < 4096 ? g : 4096)), dim3(NT), 0, s, X, ld, rows, cols, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int mm_fast(const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t M, int64_t K,
                   int64_t N, void* ws, int64_t ws_bytes, rbg_status* st, hipStream_t s) {
  if (!A || !B || !C || !ws || !st || M <= 0 || K <= 0 || N <= 0 || lda < K || ldb < N || ldc < N) return RBG_E_ARG;
  if (ws_bytes < mm_ws_bytes<T>(M, K, N)) return RBG_E_WORKSPACE;
  constexpr int E = MmTile<T>::EDGE;
  const int64_t Mp = round_up(M, E), Np = round_up(N, E), Kp = round_up(K, BLK);
  if ((Mp / E) * (Np / E) > INT32_MAX) return RBG_E_ARG;
  T* AT = (T*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  T* BE = AT + Kp * Mp;
  auto grid = [](int64_t total) { int64_t g = (total + NT * 4 - 1) / (NT * 4); return (unsigned)(g < 8192 ? g : 8192); };
  hipLaunchKernelGGL((k_mm_prep<T, true>), dim3(grid(Kp * Mp)), dim3(NT), 0, s, A, lda, M, K, AT, Kp, Mp);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_mm_prep<T, false>), dim3(grid(Kp * Np)), dim3(NT), 0, s, B, ldb, K, N, BE, Kp, Np);
  RBG_CHECK_LAUNCH();
  const int64_t tiles = (Mp / E) * (Np / E);
  const int width = (int)sizeof(T) * 8;
  hipLaunchKernelGGL((k_mm_fast<T, MmTile<T>::RC, sizeof(T) == 8>), dim3((unsigned)tiles), dim3(NT), 0, s,
                     (const T*)AT, (const T*)BE, Kp, Mp, Np, A, lda, B, ldb, C, ldc, M, K, N, width, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int64_t rbg_matmul_fast_workspace_bytes(int64_t M, int64_t K, int64_t N, int32_t elem_bytes) {
  if (M < 0 || K < 0 || N < 0) return 0;
  return elem_bytes == 8 ? mm_ws_bytes<int64_t>(M, K, N) : mm_ws_bytes<int32_t>(M, K, N);
}
int rbg_matmul_precheck_i32(const int32_t* X, int64_t ld, int64_t rows, int64_t cols, rbg_status* st,
                            rbg_stream_t stream) {
  return mm_precheck<int32_t>(X, ld, rows, cols, st, (hipStream_t)stream);
}
int rbg_matmul_precheck_i64(const int64_t* X, int64_t ld, int64_t rows, int64_t cols, rbg_status* st,
                            rbg_stream_t stream) {
  return mm_precheck<int64_t>(X, ld, rows, cols, st, (hipStream_t)stream);
}
int rbg_matmul_fast_i32(const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb, int32_t* C, int64_t ldc,
                        int64_t M, int64_t K, int64_t N, void* workspace, int64_t ws_bytes, rbg_status* st,
                        rbg_stream_t stream) {
  return mm_fast<int32_t>(A, lda, B, ldb, C, ldc, M, K, N, workspace, ws_bytes, st, (hipStream_t)stream);
}
int rbg_matmul_fast_i64(const int64_t* A, int64_t lda, const int64_t* B, int64_t ldb, int64_t* C, int64_t ldc,
                        int64_t M, int64_t K, int64_t N, void* workspace, int64_t ws_bytes, rbg_status* st,
                        rbg_stream_t stream) {
  return mm_fast<int64_t>(A, lda, B, ldb, C, ldc, M, K, N, workspace, ws_bytes, st, (hipStream_t)stream);
}

}  // extern "C"
