#pragma once
// Dense AVOS transitive closure: snapshot-blocked Floyd-Warshall kernels for
// gfx950, shared by the single-matrix and the batched launchers.
//
// Reference: redblackgraph/core/src/redblackgraph/warshall.c.in:50-107 (naive
// in-place k-i-j loop, diameter = MSB of the largest value ever written, cycle
// detection on i == j).  The blocked schedule reproduces that loop's
// trajectory cell by cell:
//
//   * At step k the naive loop reads W[i][k] and W[k][j]; for inputs >= -1 and
//     no overflow neither row k nor column k changes during step k (x (x) y >= y
//     for x >= 2, the identities are neutral).  So step k is a pure function of
//     the step k-1 values of row k and column k.
//   * The panel kernel walks the pivot block's k sequentially and records, for
//     every k, column k and row k as they are at step k ("snapshots").  Phase 3
//     applies W[i][j] = W[i][j] (+) col_k[i] (x) row_k[j] for k ascending, which
//     is exactly the naive sequence of values of (i, j).
//   * Hence W, the diameter (max over every written value) and the (k, i) order
//     of cycle detections equal the naive loop's.
//
// Two arithmetic modes:
//   GENERAL  natural values, full AVOS product with the C core's overflow rules,
//            every event recorded with its naive-order key.
//   FAST     matrix kept in the "E domain" (rbg_common.hpp): a relaxation is
//            v_lshl_or_b32 + v_min_u32 and the trajectory max one v_max3 per
//            two steps.  Valid when no off-diagonal entry is +-1 or < -1 (checked
//            by rbg_fw_precheck) and the per-block MSB guard rules out overflow;
//            otherwise the kernels raise RBG_F_NEED_GENERAL.
#include <type_traits>

#include <atomic>

#include "rbg_common.hpp"

namespace rbg {

constexpr int BLK = 32;    // pivot block (k values per step)
constexpr int TILE = 128;  // phase-3 tile edge
constexpr int NT = 256;    // threads per workgroup
constexpr int RPT = 8;     // phase-3 rows per thread
constexpr int CPT = 8;     // phase-3 cols per thread

struct FwWork {
  // `ring` snapshot buffers (pivot block K uses buffer K mod ring; two for
  // the launch-per-step schedules, four for the persistent one), each BLK x n
  // for columns and rows, then the guard words of the current pivot block:
  // per 32-block b and k, gblk[b][0][k] = the largest MSB among the left
  // operands of rows of block b (column snapshot k) and gblk[b][1][k] = among
  // the right operands of columns of block b (row snapshot k).
  void* base;
  int64_t n;
  int elem;
  int ring = 2;  // a power of two
  // lower triangular schedules with the quad panel: every diagonal block as
  // the closure found it, block K at [K][32][32] (k_fw_diag_copy).  The
  // panel's workgroups read block K's diagonal block here, not in W, where
  // the workgroup that owns it stores the closed block when done -- possibly
  // before another workgroup of the same panel has started.
  void* dcopy = nullptr;
  // the persistent closure: each diagonal block's step snapshots, block K's
  // at [K][2][32][32]: column snapshots [kk][x] = D[x][kk] at step kk, then
  // row snapshots [kk][x] = D[kk][x] (k_fw_diag_walk)
  void* dsnap = nullptr;
  __host__ __device__ int slot(int64_t K) const { return (int)(K & (ring - 1)); }
  template <typename T>
  __host__ __device__ T* col(int s) const {
    return (T*)base + (size_t)s * 2 * BLK * n;
  }
  template <typename T>
  __host__ __device__ T* row(int s) const {
    return (T*)base + (size_t)s * 2 * BLK * n + (size_t)BLK * n;
  }
  __host__ __device__ int* gblk() const { return (int*)((char*)base + (size_t)2 * ring * BLK * n * elem); }
  static int64_t bytes(int64_t n, int elem, int ring = 2) {
    return (int64_t)2 * ring * BLK * n * elem + (n / BLK) * 2 * BLK * 4 + 256;
  }
};

__device__ inline uint64_t naive_key(int64_t k, int64_t i, int64_t j, int64_t n) {
  return ((uint64_t)k * (uint64_t)n + (uint64_t)i) * (uint64_t)n + (uint64_t)j;
}

template <typename T, bool FAST>
__device__ inline T snap_enc(T v) {
  return FAST ? e_enc<T>(v) : v;
}
template <typename T, bool FAST>
__device__ inline T snap_dec(T v) {
  return FAST ? e_dec<T>(v) : v;
}

// Copy the operands of detections that happened in block [k0, k0+BLK) out of
// that block's snapshots (single thread).
template <typename T, bool FAST>
__device__ void capture_block(rbg_status* st, const T* colT, const T* rowS, int64_t k0, int64_t n) {
  const uint64_t nn = (uint64_t)n * (uint64_t)n;
  if (st->cycle_count) {
    uint64_t fk = st->cycle_first_key;
    int64_t k = (int64_t)(fk / nn), i = (int64_t)((fk / n) % n);
    if (k >= k0 && k < k0 + BLK) {
      st->cycle_first_a = (int64_t)snap_dec<T, FAST>(colT[(k - k0) * n + i]);
      st->cycle_first_b = (int64_t)snap_dec<T, FAST>(rowS[(k - k0) * n + i]);
    }
    uint64_t lk = st->cycle_last_key;
    k = (int64_t)(lk / nn);
    i = (int64_t)((lk / n) % n);
    if (k >= k0 && k < k0 + BLK) {
      st->cycle_last_a = (int64_t)snap_dec<T, FAST>(colT[(k - k0) * n + i]);
      st->cycle_last_b = (int64_t)snap_dec<T, FAST>(rowS[(k - k0) * n + i]);
    }
  }
  if (st->ov_first_key != ~0ull) {
    uint64_t lk = st->ov_last_key;
    int64_t k = (int64_t)(lk / nn), i = (int64_t)((lk / n) % n), j = (int64_t)(lk % n);
    if (k >= k0 && k < k0 + BLK) {
      st->ov_last_a = (int64_t)snap_dec<T, FAST>(colT[(k - k0) * n + i]);
      st->ov_last_b = (int64_t)snap_dec<T, FAST>(rowS[(k - k0) * n + j]);
    }
  }
}

static __global__ void k_status_init(rbg_status* st) {
  st->events = 0;
  st->flags = 0;
  st->cycle_count = 0;
  st->cycle_first_key = ~0ull;
  st->cycle_last_key = 0;
  st->ov_first_key = ~0ull;
  st->ov_last_key = 0;
  st->cycle_first_a = st->cycle_first_b = st->cycle_last_a = st->cycle_last_b = 0;
  st->ov_last_a = st->ov_last_b = 0;
  st->diameter_max = 0;
  for (int i = 0; i < 8; i++) st->aux[i] = 0;
}

// Precondition of the fast kernels: every value >= -1, no off-diagonal +-1,
// every diagonal value in {-1, 0, 1}.  With such diagonals the products that
// involve a pivot cell (k, k) can neither overflow nor reveal a cycle, which
// is what lets the fast kernels skip the pivot row and column.  Also reports
// whether any off-diagonal entry lies below (RBG_F_HAS_LOWER) or above
// (RBG_F_HAS_UPPER) the diagonal: a matrix with entries on one side only
// can take the triangular schedule.
template <typename T>
__global__ void k_fw_precheck(const T* W, int64_t n, int64_t ld, rbg_status* st) {
  int64_t total = n * n;
  bool bad = false, lower = false, upper = false;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / n, j = e % n;
    T v = W[i * ld + j];
    if (v < (T)-1) bad = true;
    if (i != j && (v == (T)1 || v == (T)-1)) bad = true;
    if (i == j && v > (T)1) bad = true;
    if (v != 0 && i > j) lower = true;
    if (v != 0 && i < j) upper = true;
  }
  // one atomic per wave
  const uint32_t f = (__any(bad) ? (uint32_t)RBG_F_PRECOND_FAIL : 0u) | (__any(lower) ? (uint32_t)RBG_F_HAS_LOWER : 0u) |
                     (__any(upper) ? (uint32_t)RBG_F_HAS_UPPER : 0u);
  if ((threadIdx.x & 63) == 0 && f) atomicOr(&st->flags, f);
}

// The fast path's copy-in, one pass: W (npad x npad, zero padding) = the
// E-domain encoding of A (n x n, row stride lda), with k_fw_precheck's flags
// and the largest input value in aux[FW_AUX_INIT] (the triangular schedules
// count it in the trajectory maximum: k_fw_take_initial).  Workgroups take
// rows; each thread four consecutive columns per pass.
constexpr int FW_AUX_INIT = 7;
template <typename T>
__global__ __launch_bounds__(NT) void k_fw_load(const T* A, int64_t n, int64_t lda, T* W, int64_t npad,
                                                 rbg_status* st) {
  bool bad = false, lower = false, upper = false;
  int64_t vmax = 0;
  const bool vec = (lda & 3) == 0 && (n & 3) == 0 && ((uintptr_t)A & (4 * sizeof(T) - 1)) == 0;
  for (int64_t i = blockIdx.x; i < npad; i += gridDim.x) {
    for (int64_t j0 = 4 * (int64_t)threadIdx.x; j0 < npad; j0 += 4 * NT) {
      T v[4] = {0, 0, 0, 0};
      if (i < n && j0 < n) {
        if (vec) {
          if constexpr (sizeof(T) == 4) {
            const uint4 q = *(const uint4*)(A + i * lda + j0);
            v[0] = (T)q.x, v[1] = (T)q.y, v[2] = (T)q.z, v[3] = (T)q.w;
          } else {
            const ulonglong2 q0 = *(const ulonglong2*)(A + i * lda + j0);
            const ulonglong2 q1 = *(const ulonglong2*)(A + i * lda + j0 + 2);
            v[0] = (T)q0.x, v[1] = (T)q0.y, v[2] = (T)q1.x, v[3] = (T)q1.y;
          }
        } else {
#pragma unroll
          for (int c = 0; c < 4; c++)
            if (j0 + c < n) v[c] = A[i * lda + j0 + c];
        }
      }
      T o[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int64_t j = j0 + c;
        const T x = v[c];
        if (x < (T)-1) bad = true;
        if (i != j && (x == (T)1 || x == (T)-1)) bad = true;
        if (i == j && x > (T)1) bad = true;
        if (x != 0 && i > j) lower = true;
        if (x != 0 && i < j) upper = true;
        vmax = (int64_t)x > vmax ? (int64_t)x : vmax;
        o[c] = e_enc<T>(x);
      }
      if constexpr (sizeof(T) == 4) {
        *(uint4*)(W + i * npad + j0) = make_uint4((uint32_t)o[0], (uint32_t)o[1], (uint32_t)o[2], (uint32_t)o[3]);
      } else {
        *(ulonglong2*)(W + i * npad + j0) = make_ulonglong2((uint64_t)o[0], (uint64_t)o[1]);
        *(ulonglong2*)(W + i * npad + j0 + 2) = make_ulonglong2((uint64_t)o[2], (uint64_t)o[3]);
      }
    }
  }
  const uint32_t f = (__any(bad) ? (uint32_t)RBG_F_PRECOND_FAIL : 0u) | (__any(lower) ? (uint32_t)RBG_F_HAS_LOWER : 0u) |
                     (__any(upper) ? (uint32_t)RBG_F_HAS_UPPER : 0u);
  vmax = wave_max_i64(vmax);
  if ((threadIdx.x & 63) == 0) {
    if (f) atomicOr(&st->flags, f);
    if (vmax > 0) atomicMax((long long*)&st->aux[FW_AUX_INIT], (long long)vmax);
  }
}

// The fast path's decode in place (every cell of the n x n matrix): rows per
// workgroup, four consecutive cells per thread and pass (ld and W 16-byte
// aligned; k_e_convert otherwise).
template <typename T>
__global__ __launch_bounds__(NT) void k_e_decode4(T* W, int64_t n, int64_t ld) {
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x)
    for (int64_t j0 = 4 * (int64_t)threadIdx.x; j0 < n; j0 += 4 * NT) {
      T* p = W + i * ld + j0;
      if constexpr (sizeof(T) == 4) {
        uint4 q = *(const uint4*)p;
        q.x = (uint32_t)e_dec<T>((T)q.x), q.y = (uint32_t)e_dec<T>((T)q.y);
        q.z = (uint32_t)e_dec<T>((T)q.z), q.w = (uint32_t)e_dec<T>((T)q.w);
        *(uint4*)p = q;
      } else {
        ulonglong2 a = *(const ulonglong2*)p, b = *(const ulonglong2*)(p + 2);
        a.x = (uint64_t)e_dec<T>((T)a.x), a.y = (uint64_t)e_dec<T>((T)a.y);
        b.x = (uint64_t)e_dec<T>((T)b.x), b.y = (uint64_t)e_dec<T>((T)b.y);
        *(ulonglong2*)p = a;
        *(ulonglong2*)(p + 2) = b;
      }
    }
}

// A closure of a matrix k_fw_load wrote (mode 2): the side the caller chose
// must match the load's flags (else general mode is asked for, as a wrong
// side in k_e_convert); a triangular schedule takes the initial maximum.
// The load's side flags and the maximum are then cleared, leaving the status
// as the closure of a plain copy would.
static __global__ void k_fw_take_initial(rbg_status* st, int tri) {
  const uint32_t f = st->flags;
  if ((tri == 1 && (f & RBG_F_HAS_UPPER)) || (tri == 2 && (f & RBG_F_HAS_LOWER)))
    atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
  if (tri) record_diameter(st, st->aux[FW_AUX_INIT]);
  atomicAnd(&st->flags, ~(uint32_t)(RBG_F_HAS_LOWER | RBG_F_HAS_UPPER));
  st->aux[FW_AUX_INIT] = 0;
}

// E-domain conversion in place.  With st non-null (the triangular schedule)
// the largest initial value also enters the trajectory maximum: the naive
// loop writes every cell at k = 0, and on a triangular input with unit
// diagonals that first write leaves the initial value as it is, while the
// triangular schedule never visits the cells its pivot blocks cannot change.
// tri (1 lower, 2 upper) also checks the side: an off-diagonal entry on the
// other side asks for general mode, so every later kernel of the fast run
// exits and the caller reruns the full schedule.
template <typename T, bool ENC>
__global__ void k_e_convert(T* W, int64_t n, int64_t ld, rbg_status* st, int tri = 0) {
  int64_t total = n * n;
  int64_t vmax = 0;
  bool wrong = false;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / n, j = e % n;
    T v = W[i * ld + j];
    W[i * ld + j] = ENC ? e_enc<T>(v) : e_dec<T>(v);
    if ((int64_t)v > vmax) vmax = (int64_t)v;
    if (v != 0 && ((tri == 1 && j > i) || (tri == 2 && j < i))) wrong = true;
  }
  if (st) {
    vmax = wave_max_i64(vmax);
    if ((threadIdx.x & 63) == 0) record_diameter(st, vmax);
    if (__any(wrong) && (threadIdx.x & 63) == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
  }
}

// ---------------------------------------------------------------------------
// Panel kernel: phases 1 and 2 of pivot block K fused.  Every workgroup walks
// the diagonal tile (redundantly; workgroup 0 owns its results) together with
// one row- or column-panel tile, one k at a time, in the natural domain with
// the full arithmetic.  It writes the snapshots phase 3 consumes.
// ---------------------------------------------------------------------------
template <typename T, bool FAST>
__global__ __launch_bounds__(NT) void k_fw_panel(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                 int width, int is_uns, rbg_status* st) {
  __shared__ T Dg[BLK][BLK + 1];
  __shared__ T Pn[BLK][BLK + 1];
  const int tid = threadIdx.x;
  const int nb = (int)(n / BLK);
  const int64_t k0 = K * BLK;
  const int par = ws.slot(K);
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);

  // panel role: 0 = diagonal only, 1 = row panel (K, J), 2 = column panel (J, K)
  int role = 0, J = 0;
  if (nb > 1) {
    int t = blockIdx.x;
    role = t < nb - 1 ? 1 : 2;
    int idx = role == 1 ? t : t - (nb - 1);
    J = idx < K ? idx : idx + 1;
  }
  const bool owner = blockIdx.x == 0;

  if (owner && tid == 0 && K > 0) capture_block<T, FAST>(st, ws.col<T>(ws.slot(K - 1)), ws.row<T>(ws.slot(K - 1)), k0 - BLK, n);

  // load tiles (natural domain)
  const int64_t pr0 = role == 2 ? (int64_t)J * BLK : k0;  // panel row origin
  const int64_t pc0 = role == 1 ? (int64_t)J * BLK : k0;  // panel col origin
  for (int e = tid; e < BLK * BLK; e += NT) {
    int r = e / BLK, c = e % BLK;
    Dg[r][c] = snap_dec<T, FAST>(W[(k0 + r) * ld + k0 + c]);
    if (role) Pn[r][c] = snap_dec<T, FAST>(W[(pr0 + r) * ld + pc0 + c]);
  }
  __syncthreads();

  const int i = tid >> 3;         // 0..31
  const int j0 = (tid & 7) * 4;   // 0..28
  int64_t dmax = 0;
  uint32_t evacc = 0;

  for (int kk = 0; kk < BLK; kk++) {
    const int64_t k = k0 + kk;
    T aD = Dg[i][kk];
    T bD[4], aP = 0, bP[4];
#pragma unroll
    for (int q = 0; q < 4; q++) bD[q] = Dg[kk][j0 + q];
    if (role == 1) {
      aP = aD;
#pragma unroll
      for (int q = 0; q < 4; q++) bP[q] = Pn[kk][j0 + q];
    } else if (role == 2) {
      aP = Pn[i][kk];
#pragma unroll
      for (int q = 0; q < 4; q++) bP[q] = bD[q];
    }
    // snapshots of step k
    if (owner) {
      if ((tid & 7) == 0) colT[kk * n + k0 + i] = snap_enc<T, FAST>(aD);
      if (i == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) rowS[kk * n + k0 + j0 + q] = snap_enc<T, FAST>(bD[q]);
      }
    }
    if (role == 1 && i == 0) {
#pragma unroll
      for (int q = 0; q < 4; q++) rowS[kk * n + pc0 + j0 + q] = snap_enc<T, FAST>(bP[q]);
    }
    if (role == 2 && (tid & 7) == 0) colT[kk * n + pr0 + i] = snap_enc<T, FAST>(aP);
    __syncthreads();
    // diagonal tile
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int j = j0 + q;
      uint32_t ev = 0;
      T p = avos_product<T>(aD, bD[q], width, is_uns != 0, ev);
      if (owner) {
        if (ev) {
          record_overflow(st, naive_key(k, k0 + i, k0 + j, n), ev);
          evacc |= ev;
        }
        if (i == j && !(p == (T)-1 || p == 0 || p == 1)) record_cycle(st, naive_key(k, k0 + i, k0 + i, n));
      }
      T v = avos_sum<T>(Dg[i][j], p);
      Dg[i][j] = v;
      if (owner && (int64_t)v > dmax) dmax = (int64_t)v;
    }
    if (role) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = j0 + q;
        uint32_t ev = 0;
        T p = avos_product<T>(aP, bP[q], width, is_uns != 0, ev);
        if (ev) record_overflow(st, naive_key(k, pr0 + i, pc0 + j, n), ev);
        T v = avos_sum<T>(Pn[i][j], p);
        Pn[i][j] = v;
        if ((int64_t)v > dmax) dmax = (int64_t)v;
      }
    }
    __syncthreads();
  }
  (void)evacc;
  // write back
  for (int e = tid; e < BLK * BLK; e += NT) {
    int r = e / BLK, c = e % BLK;
    if (owner) W[(k0 + r) * ld + k0 + c] = snap_enc<T, FAST>(Dg[r][c]);
    if (role) W[(pr0 + r) * ld + pc0 + c] = snap_enc<T, FAST>(Pn[r][c]);
  }
  dmax = wave_max_i64(dmax);
  if ((tid & 63) == 0) record_diameter(st, dmax);
}

// 16-byte vector moves of four consecutive elements (addresses are multiples
// of four elements: padded orders are multiples of 128).
__device__ inline void ld4(const uint32_t* p, uint32_t (&o)[4]) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ inline void ld4(const uint64_t* p, uint64_t (&o)[4]) {
  const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(p);
  const ulonglong2 b = *reinterpret_cast<const ulonglong2*>(p + 2);
  o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ inline void st4(uint32_t* p, const uint32_t (&v)[4]) {
  *reinterpret_cast<uint4*>(p) = make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ inline void st4(uint64_t* p, const uint64_t (&v)[4]) {
  *reinterpret_cast<ulonglong2*>(p) = make_ulonglong2(v[0], v[1]);
  *reinterpret_cast<ulonglong2*>(p + 2) = make_ulonglong2(v[2], v[3]);
}
// Device-coherent stores (relaxed device-scope atomic stores: written through
// the XCD's L2), for cells another XCD reads later in the same launch.  Such
// a store is complete, for every XCD, once the wave's vmcnt has drained.
template <typename U>
__device__ __forceinline__ void st_dev(U* p, U v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Four cells with 8-byte device-scope stores (a 4-byte store per cell costs
// the persistent int32 closure a third of its time); the caller drains them
// with wait_vm() before publishing.
__device__ __forceinline__ void st4_dev(uint32_t* p, const uint32_t (&v)[4]) {
  st_dev((uint64_t*)p, (uint64_t)v[0] | ((uint64_t)v[1] << 32));
  st_dev((uint64_t*)p + 1, (uint64_t)v[2] | ((uint64_t)v[3] << 32));
}
__device__ __forceinline__ void st4_dev(uint64_t* p, const uint64_t (&v)[4]) {
#pragma unroll
  for (int x = 0; x < 4; x++) st_dev(p + x, v[x]);
}
// A raw buffer resource over [base, base + 2 GiB) for the stores below.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, 0x00020000);
}
// Four cells at byte offset `off` of a buffer resource with one or two
// 16-byte device-scope stores (cache policy sc1, the bit device-scope atomic
// stores carry): written through the XCD's L2; the caller drains them with
// wait_vm() before publishing.  (The compiler sees these stores, and keeps
// their data registers until the hardware has read them.)
__device__ __forceinline__ void st4_buf(__amdgpu_buffer_rsrc_t rs, int off, const uint32_t (&v)[4]) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  const v4u d = {v[0], v[1], v[2], v[3]};
  __builtin_amdgcn_raw_buffer_store_b128(d, rs, off, 0, 16);
}
__device__ __forceinline__ void st4_buf(__amdgpu_buffer_rsrc_t rs, int off, const uint64_t (&v)[4]) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  const v4u d0 = {(uint32_t)v[0], (uint32_t)(v[0] >> 32), (uint32_t)v[1], (uint32_t)(v[1] >> 32)};
  const v4u d1 = {(uint32_t)v[2], (uint32_t)(v[2] >> 32), (uint32_t)v[3], (uint32_t)(v[3] >> 32)};
  __builtin_amdgcn_raw_buffer_store_b128(d0, rs, off, 0, 16);
  __builtin_amdgcn_raw_buffer_store_b128(d1, rs, off + 16, 0, 16);
}
// Wait until this wave's memory operations are complete.
__device__ inline void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ---------------------------------------------------------------------------
// Panel kernel, fast mode, for one matrix (k_fw_panel_fast): phases 1 and 2
// of pivot block K in the E domain.  Workgroup P (PANEL_NT threads, P = 0 ..
// n/32 - 2) takes panel J = P < K ? P : P + 1, that is the row-panel tile
// (block K, block J) and the column-panel tile (block J, block K), and walks
// both through the 32 steps of block K together with the diagonal block
// (redundantly; workgroup 0 owns the diagonal block's results).  Thread
// (i, j) keeps cell (i, j) of each of the three tiles in registers:
//     diagonal   D[i][j]  = D[i][j]  (+) D[i][k]  (x) D[k][j]
//     row panel  R[i][j]  = R[i][j]  (+) D[i][k]  (x) R[k][j]
//     col panel  C[i][j]  = C[i][j]  (+) C[i][k]  (x) D[k][j]
// Right operands are used as (shift, low bits) pairs: y -> (MSB(y), y minus
// its top bit), so a product is one shift-or.  The threads that hold the
// pivot row of step k+1 compute its pairs, and the threads of row and column
// k+1 write the step-(k+1) snapshots, as they finish step k; one barrier
// separates the steps.
//
// Pivot exclusions without tests: while the block is walked its diagonal
// cells hold "absent" in LDS (their values are kept aside), so at step k the
// products of the pivot row and pivot column have an absent operand and leave
// those cells as they are -- what the natural-domain rules give for a
// product with a +-1 diagonal entry.  A diagonal cell could only receive a
// present product through a cycle, and cycles are found after the walk from
// the snapshots (step k, vertex v: W[v][k] and W[k][v] both present); a cycle
// sets RBG_F_NEED_GENERAL, since the caller reruns general mode for its
// exact detections.
//
// A product with an absent operand is some value with the top bit set; the
// min can store it into an absent cell.  Such a value must not be shifted
// again (the shift could clear its top bit), so left operands are
// canonicalised to all ones when read, and so is every snapshot phase 3
// reads.  Per product, MSB(x) + MSB(y) > width - 2 raises RBG_F_NEED_GENERAL
// before any value could overflow.  Each workgroup also leaves, per k, the
// largest MSB of its column-panel and row-panel snapshots in block J's guard
// words (plain stores; phase 3 combines the blocks of its own tile).
// ---------------------------------------------------------------------------
constexpr int PANEL_NT = BLK * BLK;

template <typename T>
__device__ inline void fw_panel_pair_body(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws, int width,
                                          rbg_status* st, int bid) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr U ALL = ~(U)0;
  __shared__ U Dg[BLK][BLK + 1];
  __shared__ U Pr[BLK][BLK + 1];
  __shared__ U Pc[BLK][BLK + 1];
  // snapshots of step kk: [kk][x]
  __shared__ U sDc[BLK][BLK + 1];  // diagonal column kk, row x
  __shared__ U sDr[BLK][BLK + 1];  // diagonal row kk, column x
  __shared__ U sPr[BLK][BLK + 1];  // row-panel row kk, column x
  __shared__ U sPc[BLK][BLK + 1];  // column-panel column kk, row x
  // (shift, low bits) of the pivot rows, by step parity
  __shared__ U MYd[2][BLK][2];
  __shared__ U MYr[2][BLK][2];
  __shared__ U sDiag[BLK];
  __shared__ T sMax[PANEL_NT / 64];
  __shared__ int sbad;
  const int tid = threadIdx.x;
  const int i = tid >> 5, j = tid & (BLK - 1);
  const int64_t k0 = K * BLK;
  const int J = bid < K ? bid : bid + 1;
  const int64_t j0 = (int64_t)J * BLK;
  const bool owner = bid == 0;
  const int par = ws.slot(K);
  const int lim = width - 2;
  auto canon = [](U x) -> U { return (T)x < 0 ? ALL : x; };
  auto pair = [](U b, U& m, U& y) {
    if ((T)b < 0) {
      m = 0;
      y = TOP;
    } else {
      m = (U)msb_u(b);
      y = b ^ ((U)1 << m);
    }
  };
  auto amsb = [](U a) -> int { return (T)a > 1 ? msb_u(a) : 0; };

  // the three cells of this thread, loaded before anything waits
  U d = canon((U)W[(k0 + i) * ld + k0 + j]);
  U r = canon((U)W[(k0 + i) * ld + j0 + j]);
  U c = canon((U)W[(j0 + i) * ld + k0 + j]);
  if (tid == 0) sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
  if (i == j) {
    sDiag[i] = d;
    d = ALL;
  }
  Dg[i][j] = d;
  Pr[i][j] = r;
  Pc[i][j] = c;
  // step 0: pivot row 0's pairs and the step-0 snapshots
  if (i == 0) {
    pair(d, MYd[0][j][0], MYd[0][j][1]);
    pair(r, MYr[0][j][0], MYr[0][j][1]);
    sDr[0][j] = d;
    sPr[0][j] = r;
  }
  if (j == 0) {
    sDc[0][i] = d;
    sPc[0][i] = c;
  }
  __syncthreads();
  if (sbad) return;
  if (owner && tid == 0 && K > 0) capture_block<T, true>(st, ws.col<T>(ws.slot(K - 1)), ws.row<T>(ws.slot(K - 1)), k0 - BLK, n);
  const int64_t seen = st->diameter_max;  // a lower bound of the value at exit

  T dd = 0, dr = 0, dc = 0;
  int g = 0;
#pragma unroll 2
  for (int kk = 0; kk < BLK; kk++) {
    const int q = kk & 1;
    const U ad = canon(Dg[i][kk]);
    const U ac = canon(Pc[i][kk]);
    const U md = MYd[q][j][0], yd = MYd[q][j][1];
    const U mr = MYr[q][j][0], yr = MYr[q][j][1];
    const int gd = amsb(ad) + (int)(md > mr ? md : mr);
    const int gc = amsb(ac) + (int)md;
    g = g > gd ? g : gd;
    g = g > gc ? g : gc;
    const U pd = (ad << md) | yd;
    const U pr = (ad << mr) | yr;
    const U pc = (ac << md) | yd;
    d = d < pd ? d : pd;
    r = r < pr ? r : pr;
    c = c < pc ? c : pc;
    dd = dd > (T)d ? dd : (T)d;
    dr = dr > (T)r ? dr : (T)r;
    dc = dc > (T)c ? dc : (T)c;
    Dg[i][j] = d;
    Pr[i][j] = r;
    Pc[i][j] = c;
    if (kk + 1 < BLK) {
      if (i == kk + 1) {
        const U cd = canon(d), cr = canon(r);
        pair(cd, MYd[q ^ 1][j][0], MYd[q ^ 1][j][1]);
        pair(cr, MYr[q ^ 1][j][0], MYr[q ^ 1][j][1]);
        sDr[kk + 1][j] = cd;
        sPr[kk + 1][j] = cr;
      }
      if (j == kk + 1) {
        sDc[kk + 1][i] = canon(d);
        sPc[kk + 1][i] = canon(c);
      }
    }
    __syncthreads();
  }
  if (g > lim) sbad = 1;
  // cycles: step kk, vertex x with W[x][kk] and W[kk][x] both present
  if (owner && i != j && (T)sDc[i][j] >= 0 && (T)sDr[i][j] >= 0) {
    record_cycle(st, naive_key(k0 + i, k0 + j, k0 + j, n));
    sbad = 1;
  }
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
  // tiles and snapshots out, coalesced (thread (i, j): snapshot step i, index j)
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);
  if (owner) {
    W[(k0 + i) * ld + k0 + j] = (T)(i == j ? sDiag[i] : d);
    colT[i * n + k0 + j] = (T)sDc[i][j];
    rowS[i * n + k0 + j] = (T)sDr[i][j];
  }
  W[(k0 + i) * ld + j0 + j] = (T)r;
  rowS[i * n + j0 + j] = (T)sPr[i][j];
  W[(j0 + i) * ld + k0 + j] = (T)c;
  colT[i * n + j0 + j] = (T)sPc[i][j];
  // guard words of block J: the largest MSB of column-panel snapshot kk (left
  // operands of rows in block J) and of row-panel snapshot kk (right operands
  // of columns in block J).  Block K's are zero: phase 3 skips its rows, and
  // its columns' products were checked here.
  int ga = amsb(sPc[i][j]);
  int gb = (T)sPr[i][j] >= 0 ? msb_u(sPr[i][j]) : 0;
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) {
    const int ta = __shfl_xor(ga, o), tb = __shfl_xor(gb, o);
    ga = ga > ta ? ga : ta;
    gb = gb > tb ? gb : tb;
  }
  if (j == 0) {
    int* gw = ws.gblk();
    gw[(J * 2 + 0) * BLK + i] = ga;
    gw[(J * 2 + 1) * BLK + i] = gb;
    if (owner) {
      gw[(K * 2 + 0) * BLK + i] = 0;
      gw[(K * 2 + 1) * BLK + i] = 0;
    }
  }
  // trajectory maximum: the workgroup's, then one atomic
  T m = dr > dc ? dr : dc;
  if (owner) {
    m = m > dd ? m : dd;
    if (i == j && (T)sDiag[i] > m) m = (T)sDiag[i];
  }
  const int64_t wm = wave_max_i64((int64_t)m);
  if ((tid & 63) == 0) sMax[tid >> 6] = (T)wm;
  __syncthreads();
  if (tid == 0) {
    T x = 0;
    for (int w = 0; w < PANEL_NT / 64; w++) x = x > sMax[w] ? x : sMax[w];
    if ((int64_t)x > seen) record_diameter(st, (int64_t)x);
  }
}

// ---------------------------------------------------------------------------
// Fast mode, phases 1 and 2 for the batched launcher (fw_panel_lines_body:
// throughput over thousands of components).  Each workgroup first walks
// the 32 x 32 diagonal block of pivot block K through its 32 steps in LDS
// (four cells per thread, one barrier per step); workgroup 0 owns the result
// and its events.  Then each of its 256 threads owns one panel line: a column
// j of the row panel (rows of the block, j outside it) or a row i of the
// column panel.  Given the diagonal block's snapshots a line evolves on its
// own:
//     W[r][j] = W[r][j] (+) colsnap_k[r] (x) W[k][j]   (r != k)
// with W[k][j] invariant at step k, so the thread walks its 32 cells through
// the 32 steps in registers -- 31 shift-or + min per step, the shift and low
// bits of the line's own operand computed once per step -- and writes the
// line's snapshots (W[k][j] at step k) for phase 3.  The MSB guard is checked
// per product on the diagonal block and per step, against the largest
// operand on the block's side, on the lines.
//
// Operands are canonicalised (absent = all ones) before use: a product with
// an absent operand is some value with the top bit set, which the min may
// store, and which must not be shifted again.
// ---------------------------------------------------------------------------
template <typename U>
struct LinesShared {
  U D[BLK][BLK + 1];
  U sC[BLK][BLK + 1];  // [kk][r] = W[r][kk] at step kk (diagonal block)
  U sR[BLK][BLK + 1];  // [kk][c] = W[kk][c] at step kk (diagonal block)
  U Mb[BLK][BLK];      // shift and low bits of sR
  U Yb[BLK][BLK];
  int maxA[BLK], maxB[BLK], sbad;
};

// The panel lines a step needs: row-panel lines (columns) [r0, r1) and
// column-panel lines (rows) [c0, c1) -- every line in the full schedule; in
// the triangular one the lines whose snapshots the step's tile rectangle
// reads (the others lie on the absent side and are not read).
struct LineRange {
  int64_t r0, r1, c0, c1;
};

// dsn (lower triangular components): block K's step snapshots, the diagonal
// block already closed in W with its events and guard (k_fw_diag_walk); the
// workgroup takes the snapshots instead of walking the block.
template <typename T>
__device__ __forceinline__ void fw_panel_lines_body(LinesShared<typename unsigned_of<T>::type>& L, T* W, int64_t n,
                                                    int64_t ld, int64_t K, FwWork ws, int width, rbg_status* st,
                                                    int bid, LineRange lr, const T* dsn = nullptr) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  auto& D = L.D;
  auto& sC = L.sC;
  auto& sR = L.sR;
  auto& Mb = L.Mb;
  auto& Yb = L.Yb;
  auto& maxA = L.maxA;
  auto& maxB = L.maxB;
  auto& sbad = L.sbad;
  const int tid = threadIdx.x;
  const int64_t k0 = K * BLK;
  const int par = ws.slot(K);
  const bool owner = bid == 0;
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);
  auto canon = [](U x) -> U { return (T)x < 0 ? ~(U)0 : x; };
  {
    // a workgroup with no needed line leaves at once (workgroup 0 owns the
    // diagonal block and always stays)
    const int64_t g0 = (int64_t)bid * NT, g1 = g0 + NT;
    const int64_t ra = g0 > lr.r0 ? g0 : lr.r0, rb = (g1 < n ? g1 : n) < lr.r1 ? (g1 < n ? g1 : n) : lr.r1;
    const int64_t ca = (g0 > n ? g0 - n : 0) > lr.c0 ? (g0 > n ? g0 - n : 0) : lr.c0;
    const int64_t cb = (g1 > n ? g1 - n : 0) < lr.c1 ? (g1 > n ? g1 - n : 0) : lr.c1;
    if (!owner && ra >= rb && ca >= cb) return;
  }
  if (tid == 0) sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
  __syncthreads();
  if (sbad) return;
  if (owner && tid == 0 && K > 0) capture_block<T, true>(st, ws.col<T>(ws.slot(K - 1)), ws.row<T>(ws.slot(K - 1)), k0 - BLK, n);
  const int64_t seen = st->diameter_max;  // a lower bound of the value at exit
  if (dsn) {
    for (int e = tid; e < BLK * BLK; e += NT) {
      sC[e / BLK][e % BLK] = (U)dsn[e];
      sR[e / BLK][e % BLK] = (U)dsn[BLK * BLK + e];
    }
  } else {
    for (int e = tid; e < BLK * BLK; e += NT) D[e / BLK][e % BLK] = canon((U)W[(k0 + e / BLK) * ld + k0 + e % BLK]);
  }
  // this thread's panel line, loaded while the diagonal block is walked
  const bool rowp = (int64_t)bid * NT + tid < n;
  const int64_t line = rowp ? (int64_t)bid * NT + tid : (int64_t)bid * NT + tid - n;
  const bool active = (int64_t)bid * NT + tid < 2 * n && !(line >= k0 && line < k0 + BLK) &&
                      (rowp ? line >= lr.r0 && line < lr.r1 : line >= lr.c0 && line < lr.c1);
  U c[BLK];
  U* Wu = (U*)W;
  if (active) {
    if (rowp) {
#pragma unroll
      for (int r = 0; r < BLK; r++) c[r] = Wu[(k0 + r) * ld + line];
    } else {
#pragma unroll
      for (int h = 0; h < BLK / 4; h++) {
        U q[4];
        ld4(Wu + line * ld + k0 + 4 * h, q);
#pragma unroll
        for (int x = 0; x < 4; x++) c[4 * h + x] = q[x];
      }
    }
  }
  __syncthreads();

  // ---- the diagonal block ----
  const int i = tid >> 3;
  const int j0 = (tid & 7) * 4;
  const int lim = width - 2;
  U dv[4];
#pragma unroll
  for (int x = 0; x < 4; x++) dv[x] = D[i][j0 + x];
  T dmax = 0;
  bool bad = false;
  for (int kk = 0; kk < (dsn ? 0 : BLK); kk++) {
    const U aD = canon(D[i][kk]);
    U bD[4];
#pragma unroll
    for (int x = 0; x < 4; x++) bD[x] = canon(D[kk][j0 + x]);
    if ((tid & 7) == 0) sC[kk][i] = aD;
    if (i == 0) {
#pragma unroll
      for (int x = 0; x < 4; x++) sR[kk][j0 + x] = bD[x];
    }
    const int maD = (T)aD > 1 ? msb_u(aD) : 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const int j = j0 + x;
      if (i != kk && j != kk) {
        U m, y;
        if ((T)bD[x] < 0) {
          m = 0;
          y = TOP;
        } else {
          m = (U)msb_u(bD[x]);
          y = bD[x] & (((U)1 << m) - 1);
        }
        if ((T)bD[x] > 1 && maD + (int)m > lim) bad = true;
        const U p = (aD << m) | y;
        if (owner && i == j && (T)p >= 0) record_cycle(st, naive_key(k0 + kk, k0 + i, k0 + i, n));
        dv[x] = dv[x] < p ? dv[x] : p;
        D[i][j] = dv[x];
      }
      if (owner && (T)dv[x] > dmax) dmax = (T)dv[x];
    }
    __syncthreads();
  }
  // shift / low bits of the block's rows, and the per-step operand maxima
  for (int e = tid; e < BLK * BLK; e += NT) {
    const int kk = e / BLK, x = e % BLK;
    const U b = sR[kk][x];
    if ((T)b < 0) {
      Mb[kk][x] = 0;
      Yb[kk][x] = TOP;
    } else {
      const U m = (U)msb_u(b);
      Mb[kk][x] = m;
      Yb[kk][x] = b & (((U)1 << m) - 1);
    }
  }
  if (bad) sbad = 1;
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
  if (tid < 2 * BLK) {
    const int kk = tid & (BLK - 1);
    int mx = 0;
    for (int x = 0; x < BLK; x++) {
      int m;
      if (tid < BLK) {
        const U a = sC[kk][x];
        m = (T)a > 1 ? msb_u(a) : 0;
      } else {
        m = (T)Yb[kk][x] < 0 ? 0 : (int)Mb[kk][x];
      }
      mx = m > mx ? m : mx;
    }
    if (tid < BLK) maxA[kk] = mx;
    else maxB[kk] = mx;
  }
  __syncthreads();
  if (owner) {
    for (int e = tid; e < BLK * BLK; e += NT) {
      const int r = e / BLK, cc = e % BLK;
      if (!dsn) W[(k0 + r) * ld + k0 + cc] = (T)D[r][cc];
      colT[r * n + k0 + cc] = (T)sC[r][cc];
      rowS[r * n + k0 + cc] = (T)sR[r][cc];
    }
  }

  // ---- the panel lines ----
  T d0 = dmax, d1 = 0;
  bool lbad = false;
  if (active) {
#pragma unroll
    for (int kk = 0; kk < BLK; kk++) {
      const U pk = canon(c[kk]);  // invariant at step kk
      if (rowp) rowS[kk * n + line] = (T)pk;
      else colT[kk * n + line] = (T)pk;
      d0 = (T)pk > d0 ? (T)pk : d0;
      if (rowp) {
        U m, y;
        if ((T)pk < 0) {
          m = 0;
          y = TOP;
        } else {
          m = (U)msb_u(pk);
          y = pk & (((U)1 << m) - 1);
        }
        if ((T)pk > 1 && maxA[kk] + (int)m > lim) lbad = true;
#pragma unroll
        for (int r = 0; r < BLK; r++) {
          if (r != kk) {
            const U p = (sC[kk][r] << m) | y;
            const U v = c[r] < p ? c[r] : p;
            c[r] = v;
            if (r & 1) d1 = (T)v > d1 ? (T)v : d1;
            else d0 = (T)v > d0 ? (T)v : d0;
          }
        }
      } else {
        if ((T)pk > 1 && (int)msb_u(pk) + maxB[kk] > lim) lbad = true;
#pragma unroll
        for (int x = 0; x < BLK; x++) {
          if (x != kk) {
            const U p = (pk << Mb[kk][x]) | Yb[kk][x];
            const U v = c[x] < p ? c[x] : p;
            c[x] = v;
            if (x & 1) d1 = (T)v > d1 ? (T)v : d1;
            else d0 = (T)v > d0 ? (T)v : d0;
          }
        }
      }
    }
    if (lbad) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    if (rowp) {
#pragma unroll
      for (int r = 0; r < BLK; r++) Wu[(k0 + r) * ld + line] = c[r];
    } else {
#pragma unroll
      for (int h = 0; h < BLK / 4; h++) {
        U q[4];
#pragma unroll
        for (int x = 0; x < 4; x++) q[x] = c[4 * h + x];
        st4(Wu + line * ld + k0 + 4 * h, q);
      }
    }
  }
  const int64_t dm = wave_max_i64((int64_t)(d0 > d1 ? d0 : d1));
  if ((tid & 63) == 0 && dm > seen) record_diameter(st, dm);
}

// ---------------------------------------------------------------------------
// Phase 3, fast mode (E domain).  One (16 RPT) x (16 CPT) tile per workgroup
// of 16 x 16 threads; a thread holds RPT rows x CPT columns in VGPRs for the
// whole pivot block.  Its columns are CPT/4 chunks of four, 64 columns apart,
// so that the eight lanes an LDS read serves together cover all 32 banks.
// Per k a thread reads RPT left operands and CPT (shift, low-bits) pairs from
// LDS; a relaxation is one shift-or and one unsigned min, the trajectory
// maximum one signed max3 per two steps.  Threads whose rows lie in the pivot
// block skip (the panel kernel owns those rows); cells of the pivot columns
// are relaxed again with the same snapshot products, which leaves them and
// the trajectory maximum unchanged.
// ---------------------------------------------------------------------------
// PANEL_GUARD: the per-k MSB guard comes from the per-block guard words the
// panel kernel left (one matrix): the largest left-operand MSB over the
// tile's row blocks plus the largest right-operand MSB over its column
// blocks.  Otherwise the tile computes the same bound from the snapshots it
// stages (the batched launcher's panels leave no guard words).
// SHIFT_BYTES: the shifts are staged as bytes (36 KB of LDS for a 128-tile
// instead of 48 KB, so four workgroups fit a CU when the registers allow it),
// at 12 more VALU instructions per 128 relaxations to take them apart.
// ahead (the look-ahead schedule, k_fw_step_fast): the next pivot block's
// rows and columns, [k0 + 32, k0 + 64), belong to the next panel, which runs
// in the same launch: the tile skips those rows as it skips the pivot rows,
// and does not store into those columns.
// The LDS arrays come from the kernel (TileShared), so that a kernel with
// two workgroup roles can overlay them with the other role's.
template <typename U, int TL, bool SHIFT_BYTES>
struct TileShared {
  using MsT = typename std::conditional<SHIFT_BYTES, uint8_t, U>::type;
  U As[BLK][TL];
  MsT Ms[BLK][TL];
  U Ys[BLK][TL];
  int sGa[BLK], sGb[BLK];
};

// DEV: device-coherent stores (the persistent closure, whose next reader of
// these cells may run on another XCD in the same launch).
template <typename T, int RPT, int CPT, bool PANEL_GUARD, bool SHIFT_BYTES, bool DEV = false>
__device__ __forceinline__ void fw_tile_fast_body(TileShared<typename unsigned_of<T>::type, 16 * CPT, SHIFT_BYTES>& L, T* W,
                                         int64_t n, int64_t ld, int64_t K, FwWork ws, int width, rbg_status* st,
                                         int I, int J, int ahead = 0) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr int TL = 16 * CPT;  // tile edge
  constexpr int NCH = CPT / 4;  // column chunks per thread
  static_assert(RPT == CPT && CPT % 4 == 0, "square thread blocks of whole chunks");
  auto& As = L.As;
  auto& Ms = L.Ms;
  auto& Ys = L.Ys;
  auto& sGa = L.sGa;
  auto& sGb = L.sGb;
  const int tid = threadIdx.x;
  const int64_t k0 = K * BLK;
  const int par = ws.slot(K);
  const U* colT = (const U*)ws.col<T>(par);
  const U* rowS = (const U*)ws.row<T>(par);

  const int tr = tid >> 4, tc = tid & 15;
  const int64_t r0 = (int64_t)I * TL + tr * RPT;
  const int64_t cb = (int64_t)J * TL + tc * 4;  // column of chunk 0; chunk h adds 64 h
  const bool skip = r0 >= k0 && r0 < k0 + (ahead ? 2 : 1) * BLK;

  // every global load first: the tile, this thread's share of the
  // snapshots, the status flags and (one matrix) the guard words
  U w[RPT][CPT];
  U* Wu = (U*)W;
  if (!skip) {
#pragma unroll
    for (int r = 0; r < RPT; r++)
#pragma unroll
      for (int h = 0; h < NCH; h++) {
        U q[4];
        ld4(Wu + (r0 + r) * ld + cb + 64 * h, q);
#pragma unroll
        for (int x = 0; x < 4; x++) w[r][4 * h + x] = q[x];
      }
  }
  constexpr int GL = TL / 4;            // threads staging one k (a lane group)
  constexpr int SPT = BLK * GL / NT;    // staged chunks per thread
  U sa[SPT][4], sb[SPT][4];
#pragma unroll
  for (int q = 0; q < SPT; q++) {
    const int e = tid + q * NT, kk = e / GL, x = (e % GL) * 4;
    if (ahead == 2) {
      // the persistent closure: left operands of rows above block K and
      // right operands of columns right of it are absent; block K's own
      // come from the diagonal block's snapshots
      const int64_t ri = (int64_t)I * TL + x, cj = (int64_t)J * TL + x;
      const U* dsn = (const U*)ws.dsnap + K * 2 * BLK * BLK;
      if (ri < k0) {
#pragma unroll
        for (int c = 0; c < 4; c++) sa[q][c] = ~(U)0;
      } else if (ri < k0 + BLK) {
        ld4(dsn + kk * BLK + (ri - k0), sa[q]);
      } else {
        ld4(colT + kk * n + ri, sa[q]);
      }
      if (cj >= k0 + BLK) {
#pragma unroll
        for (int c = 0; c < 4; c++) sb[q][c] = ~(U)0;
      } else if (cj >= k0) {
        ld4(dsn + BLK * BLK + kk * BLK + (cj - k0), sb[q]);
      } else {
        ld4(rowS + kk * n + cj, sb[q]);
      }
    } else {
      ld4(colT + kk * n + (int64_t)I * TL + x, sa[q]);
      ld4(rowS + kk * n + (int64_t)J * TL + x, sb[q]);
    }
  }
  const uint32_t flags = st->flags;
  const int64_t seen = st->diameter_max;  // a lower bound of the value at exit
  int pga = 0, pgb = 0;
  if (PANEL_GUARD && tid < BLK) {
    constexpr int NBT = TL / BLK;  // 32-blocks per tile edge
    const int* gw = ws.gblk();
#pragma unroll
    for (int q = 0; q < NBT; q++) {
      const int a = gw[((I * NBT + q) * 2 + 0) * BLK + tid];
      const int b = gw[((J * NBT + q) * 2 + 1) * BLK + tid];
      pga = a > pga ? a : pga;
      pgb = b > pgb ? b : pgb;
    }
  }
  // (shift, low bits) pairs into LDS; without guard words, per k the largest
  // MSB among this tile's left operands and among its right operands.  One
  // lane group (within one wave) stages each k: its first lane clears the two
  // words and the group's LDS max operations follow it in the wave's order.
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
    if (!PANEL_GUARD) {
      if (e % GL == 0) {
        sGa[kk] = 0;
        sGb[kk] = 0;
      }
      if (ga) atomicMax(&sGa[kk], ga);
      if (gb) atomicMax(&sGb[kk], gb);
    }
    st4(&As[kk][x], sa[q]);
    if (SHIFT_BYTES)
      *(uint32_t*)&Ms[kk][x] = (uint32_t)m[0] | ((uint32_t)m[1] << 8) | ((uint32_t)m[2] << 16) | ((uint32_t)m[3] << 24);
    else
      st4((U*)&Ms[kk][x], m);
    st4(&Ys[kk][x], y);
  }
  // exit when an earlier kernel already asked for general mode (no barrier
  // follows, so threads may leave on their own reads of the flags); a failed
  // MSB guard only asks for it (the values this launch then computes are
  // discarded with the rest of the fast run)
  if (PANEL_GUARD && tid < BLK && pga + pgb > width - 2) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
  __syncthreads();
  if (!PANEL_GUARD && tid < BLK && sGa[tid] + sGb[tid] > width - 2)
    atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
  if (flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) return;
  if (skip) return;

  // diagonal cells of this thread (tiles with I == J): four at most, rows
  // dr..dr+3 against chunk dh's columns; mode = 2 dh + dr / 4, or -1
  int dmode = -1;
  if (I == J) {
    const int d = tc * 4 - tr * RPT;
#pragma unroll
    for (int h = 0; h < NCH; h++) {
      const int dr = d + 64 * h;
      if (dr == 0 || (RPT == 8 && dr == 4)) dmode = 2 * h + dr / 4;
    }
  }
  T d0 = 0, d1 = 0;
  // a shift is used modulo the operand width, which is what the shift
  // instructions do with their amount: a chunk's four shift bytes come in one
  // LDS word and byte q is that word moved right by 8 q, no masking needed
  constexpr uint32_t SHM = 8 * sizeof(U) - 1;
  for (int kk = 0; kk < BLK; kk += 2) {
    U a0[RPT], a1[RPT], y0[CPT], y1[CPT];
    uint32_t m0[CPT], m1[CPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      a0[r] = As[kk][tr * RPT + r];
      a1[r] = As[kk + 1][tr * RPT + r];
    }
    if (SHIFT_BYTES) {
#pragma unroll
      for (int h = 0; h < NCH; h++) {
        const uint32_t p0 = *(const uint32_t*)&Ms[kk][64 * h + tc * 4];
        const uint32_t p1 = *(const uint32_t*)&Ms[kk + 1][64 * h + tc * 4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          m0[4 * h + q] = p0 >> (8 * q);
          m1[4 * h + q] = p1 >> (8 * q);
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPT; c++) {
        const int x = 64 * (c >> 2) + tc * 4 + (c & 3);
        m0[c] = (uint32_t)Ms[kk][x];
        m1[c] = (uint32_t)Ms[kk + 1][x];
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const int x = 64 * (c >> 2) + tc * 4 + (c & 3);
      y0[c] = Ys[kk][x];
      y1[c] = Ys[kk + 1][x];
    }
    if (dmode >= 0) {
#pragma unroll
      for (int mode = 0; mode < 2 * NCH; mode++) {
        if (mode == dmode) {
          const int dr = (mode & 1) * 4, dc = (mode >> 1) * 4;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            if (dr + q < RPT) {
              const U p0 = (a0[dr + q] << (m0[dc + q] & SHM)) | y0[dc + q];
              const U p1 = (a1[dr + q] << (m1[dc + q] & SHM)) | y1[dc + q];
              if ((T)p0 >= 0) record_cycle(st, naive_key(k0 + kk, r0 + dr + q, r0 + dr + q, n));
              if ((T)p1 >= 0) record_cycle(st, naive_key(k0 + kk + 1, r0 + dr + q, r0 + dr + q, n));
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; r++) {
#pragma unroll
      for (int c = 0; c < CPT; c++) {
        U p0 = (a0[r] << (m0[c] & SHM)) | y0[c];
        U p1 = (a1[r] << (m1[c] & SHM)) | y1[c];
        U v1 = w[r][c] < p0 ? w[r][c] : p0;
        U v2 = v1 < p1 ? v1 : p1;
        w[r][c] = v2;
        T t1 = (T)v1, t2 = (T)v2;
        T tm = t1 > t2 ? t1 : t2;
        if (c & 1) d1 = tm > d1 ? tm : d1;
        else d0 = tm > d0 ? tm : d0;
      }
    }
  }
  // device-coherent stores: 16-byte buffer stores through a resource based
  // at the tile's first cell (offsets within a tile fit 32 bits at any order)
  const __amdgpu_buffer_rsrc_t rs = buf_rsrc(Wu + (int64_t)I * TL * ld + (int64_t)J * TL);
#pragma unroll
  for (int h = 0; h < NCH; h++) {
    const int64_t c = cb + 64 * h;
    // ahead 1: the next panel's columns; ahead 2: every column right of the
    // pivot block, which a lower triangular step leaves unchanged and later
    // panels of the persistent closure may be writing
    if (ahead && c >= k0 + BLK && (ahead == 2 || c < k0 + 2 * BLK)) continue;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      U q[4];
#pragma unroll
      for (int x = 0; x < 4; x++) q[x] = w[r][4 * h + x];
      if constexpr (DEV)
        st4_buf(rs, (int)(((r0 + r - (int64_t)I * TL) * ld + c - (int64_t)J * TL) * sizeof(U)), q);
      else
        st4(Wu + (r0 + r) * ld + c, q);
    }
  }
  int64_t dm = (int64_t)(d0 > d1 ? d0 : d1);
  dm = wave_max_i64(dm);
  if ((tid & 63) == 0 && dm > seen) record_diameter(st, dm);
}

template <typename T>
__global__ __launch_bounds__(PANEL_NT) void k_fw_panel_fast(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                            int width, rbg_status* st) {
  fw_panel_pair_body<T>(W, n, ld, K, ws, width, st, blockIdx.x);
}

// Phase-3 tile shape per element type: 8 x 8 int32 cells per thread (128 x
// 128 tiles), 4 x 4 int64 cells (64 x 64 tiles: an 8 x 8 int64 block needs more
// registers than leave room for a second wave per SIMD).  The int64 tile
// stages its shifts as bytes and fits four workgroups per CU (126 VGPRs, 34 KB
// of LDS); the int32 tile keeps word shifts and three per CU: at four it
// measured slower (182 us against 175 us per N = 8192 step), the byte
// unpacking and a 128-register schedule costing more than the partial last
// round of 768-workgroup rounds that four per CU removes.
template <typename T>
struct FastTile {
  static constexpr int RC = sizeof(T) == 4 ? 8 : 4;
  static constexpr int EDGE = 16 * RC;
  static constexpr bool SHIFT_BYTES = sizeof(T) == 8;
  static constexpr int PER_CU = SHIFT_BYTES ? 4 : 3;  // workgroups per CU the registers are sized for
};

// The tiles one phase-3 launch covers: the rectangle of tile rows
// [i0, i0 + grid / jn) and tile columns [j0, j0 + jn), workgroup b taking
// tile (i0 + b / jn, j0 + b % jn).  The full schedule launches the whole
// matrix (i0 = j0 = 0, jn = n / edge); the triangular schedule only the tiles
// that pivot block K can change (TriRect).
struct TileRect {
  int i0, j0, jn;
};

// Small int32 matrices (at most SMALL_TILE_MAX_N): 64 x 64 tiles, four times
// as many workgroups per pivot step, since a 128-tile grid leaves most CUs
// idle there (N = 512: 16 workgroups for 256 CUs).  The triangular schedule
// also takes them for the pivot steps whose 128-tile rectangle is small.
constexpr int64_t SMALL_TILE_MAX_N = 2048;
template <typename T>
__global__ __launch_bounds__(NT, 4) void k_fw_tile_fast_small(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                             int width, rbg_status* st, TileRect r) {
  __shared__ TileShared<typename unsigned_of<T>::type, 64, false> L;
  const int b = blockIdx.x;
  fw_tile_fast_body<T, 4, 4, true, false>(L, W, n, ld, K, ws, width, st, r.i0 + b / r.jn, r.j0 + b % r.jn);
}

template <typename T, int PER_CU = FastTile<T>::PER_CU>
__global__ __launch_bounds__(NT, PER_CU) void k_fw_tile_fast(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                            int width, rbg_status* st, TileRect r) {
  __shared__ TileShared<typename unsigned_of<T>::type, FastTile<T>::EDGE, FastTile<T>::SHIFT_BYTES> L;
  const int b = blockIdx.x;
  fw_tile_fast_body<T, FastTile<T>::RC, FastTile<T>::RC, true, FastTile<T>::SHIFT_BYTES>(
      L, W, n, ld, K, ws, width, st, r.i0 + b / r.jn, r.j0 + b % r.jn);
}

// The diagonal blocks of W into dst ([K][32][32], FwWork::dcopy): one
// workgroup per block.  In a lower triangular closure no step changes
// diagonal block K before pivot block K's panel, so the copy taken before the
// first step is what that panel must start from.
template <typename T>
__global__ __launch_bounds__(NT) void k_fw_diag_copy(const T* W, int64_t ld, T* dst) {
  const int64_t k0 = (int64_t)blockIdx.x * BLK;
  for (int e = threadIdx.x; e < BLK * BLK; e += NT)
    dst[k0 * BLK + e] = W[(k0 + e / BLK) * ld + k0 + e % BLK];
}

// Panel of the look-ahead schedule (fw_panel_quad_body), for lower
// triangular matrices: the pair panel's walk (fw_panel_pair_body) with 256
// threads, thread (i4, j) holding rows i4, i4 + 8, i4 + 16, i4 + 24 of column
// j of the diagonal tile and of the workgroup's one live panel tile -- the
// row-panel tile (K, J) when J < K, else the column-panel tile (J, K); the
// other one lies above the diagonal, stays absent, and only its snapshots
// (absent) are written.  256 threads, so that it can share a launch with
// 256-thread tiles.
// Before the walk, a row-panel workgroup applies pivot block K - 1 to its
// tile from that block's snapshots, as phase 3 would have:
//     R[i][j] = R[i][j] (+) colsnap_k[k0 + i] (x) rowsnap_k[j0 + j],
//     k = k0 - 32 .. k0 - 1,
// all operands fixed, so one pass over the 32 steps with no barrier.  These
// are the only cells of block K's row and column panels that block K - 1
// changes (its column panel would need W[k][j] with k < j).
// The MSB guard is checked per step from the snapshots after the walk (the
// largest left-operand MSB plus the largest shift among step kk's
// snapshots bounds every product of step kk); a failure discards the run
// like any other.  Guard words are not written: the tiles of this schedule
// check the guard on the operands they stage.
template <typename U>
struct QuadStrip {  // the strip's operands, before the walk
  U sL[BLK][BLK + 1];        // colsnap_k[k0 + x] for step kk
  U sY[BLK][BLK + 1];        // low bits of rowsnap_k[j0 + x]
  uint8_t sM[BLK][BLK + 4];  // and its shift
};
template <typename U>
struct QuadSnaps {     // the walk's snapshots
  U sDc[BLK][BLK + 1];  // diagonal column kk at step kk, row x
  U sDr[BLK][BLK + 1];  // diagonal row kk, column x
  U sPr[BLK][BLK + 1];  // row-panel row kk, column x
  U sPc[BLK][BLK + 1];  // column-panel column kk, row x
};
template <typename U>
struct QuadShared {
  U Dg[BLK][BLK + 1];
  U Pc[BLK][BLK + 1];
  union {  // one after the other: the strip's operands, then the snapshots
    QuadStrip<U> s;
    QuadSnaps<U> w;
  } x;
  U MYd[2][BLK][2];  // (shift, low bits) of the pivot rows, by step parity
  U MYr[2][BLK][2];
  U sDiag[BLK];
  int64_t sMax[NT / 64];
  int sbad, sguard;  // sguard: the strip's MSB guard failed
};

// DEV: device-coherent stores (the persistent closure, whose readers of these
// cells may run on other XCDs in the same launch).
template <typename T, bool DEV = false>
__device__ __forceinline__ void fw_panel_quad_body(QuadShared<typename unsigned_of<T>::type>& L, T* W, int64_t n,
                                                   int64_t ld, int64_t K, FwWork ws, int width, rbg_status* st,
                                                   int bid) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr U ALL = ~(U)0;
  const int tid = threadIdx.x;
  const int j = tid & (BLK - 1), i4 = tid >> 5;
  const int64_t k0 = K * BLK;
  const int J = bid < K ? bid : bid + 1;
  const int64_t j0 = (int64_t)J * BLK;
  const bool owner = bid == 0;
  const bool rowp = J < K;  // the live tile: row panel (K, J), else column panel (J, K)
  const int par = ws.slot(K);
  const int lim = width - 2;
  auto canon = [](U x) -> U { return (T)x < 0 ? ALL : x; };
  auto pair = [](U b, U& m, U& y) {
    if ((T)b < 0) {
      m = 0;
      y = TOP;
    } else {
      m = (U)msb_u(b);
      y = b ^ ((U)1 << m);
    }
  };
  auto amsb = [](U a) -> int { return (T)a > 1 ? msb_u(a) : 0; };

  // the cells of this thread, loaded before anything waits: p is the live
  // panel tile's cell (row panel: W[k0 + i][j0 + j], column panel:
  // W[j0 + i][k0 + j])
  U d[4], p[4];
  const U* dsrc = (const U*)ws.dcopy + K * BLK * BLK;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = i4 + 8 * q;
    d[q] = canon(dsrc[i * BLK + j]);
    p[q] = canon((U)(rowp ? W[(k0 + i) * ld + j0 + j] : W[(j0 + i) * ld + k0 + j]));
  }
  if (tid == 0) {
    L.sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
    L.sguard = 0;
  }
  if (rowp) {
    // the strip's operands: all loads first, then the LDS stores
    const U* pcol = (const U*)ws.col<T>(ws.slot(K - 1));
    const U* prow = (const U*)ws.row<T>(ws.slot(K - 1));
    constexpr int PER = BLK * BLK / NT;
    U la[PER], lb[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = tid + q * NT, kk = e / BLK, x = e % BLK;
      la[q] = pcol[kk * n + k0 + x];
      lb[q] = prow[kk * n + j0 + x];
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = tid + q * NT, kk = e / BLK, x = e % BLK;
      L.x.s.sL[kk][x] = la[q];
      U m, y;
      pair(lb[q], m, y);
      L.x.s.sM[kk][x] = m;
      L.x.s.sY[kk][x] = y;
    }
  }
  __syncthreads();
  if (L.sbad) return;
  T ds = 0;
  if (rowp) {
    // per step, the largest left-operand MSB plus the largest shift
    if (tid < BLK) {
      int ga = 0, gb = 0;
      for (int x = 0; x < BLK; x++) {
        const int a = amsb(L.x.s.sL[tid][x]);
        const int m = (T)L.x.s.sY[tid][x] < 0 ? 0 : (int)L.x.s.sM[tid][x];
        ga = a > ga ? a : ga;
        gb = m > gb ? m : gb;
      }
      if (ga + gb > lim) L.sguard = 1;
    }
    for (int kk = 0; kk < BLK; kk++) {
      const U m = L.x.s.sM[kk][j], y = L.x.s.sY[kk][j];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const U v = (L.x.s.sL[kk][i4 + 8 * q] << m) | y;
        p[q] = p[q] < v ? p[q] : v;
        ds = ds > (T)p[q] ? ds : (T)p[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) p[q] = canon(p[q]);
  }
  __syncthreads();  // the strip's operands are read; their LDS takes the snapshots
  // the walk's starting state
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = i4 + 8 * q;
    if (i == j) {
      L.sDiag[i] = d[q];
      d[q] = ALL;
    }
    L.Dg[i][j] = d[q];
    if (!rowp) L.Pc[i][j] = p[q];
    if (j == 0) {
      L.x.w.sDc[0][i] = d[q];
      L.x.w.sPc[0][i] = rowp ? ALL : p[q];
    }
  }
  if (i4 == 0) {  // pivot row 0 (q = 0): its pairs and the step-0 row snapshots
    pair(d[0], L.MYd[0][j][0], L.MYd[0][j][1]);
    L.x.w.sDr[0][j] = d[0];
    if (rowp) pair(p[0], L.MYr[0][j][0], L.MYr[0][j][1]);
    L.x.w.sPr[0][j] = rowp ? p[0] : ALL;
  }
  __syncthreads();
  if (L.sguard) {
    if (tid == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
  const int64_t seen = st->diameter_max;  // a lower bound of the value at exit

  T dd = 0, dp = 0;
  for (int kk = 0; kk < BLK; kk++) {
    const int qq = kk & 1;
    const U md = L.MYd[qq][j][0], yd = L.MYd[qq][j][1];
    if (rowp) {
      //   diagonal  D[i][j] = D[i][j] (+) D[i][k] (x) D[k][j]
      //   row panel R[i][j] = R[i][j] (+) D[i][k] (x) R[k][j]
      const U mr = L.MYr[qq][j][0], yr = L.MYr[qq][j][1];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i4 + 8 * q;
        const U ad = canon(L.Dg[i][kk]);
        const U pd = (ad << md) | yd;
        const U pr = (ad << mr) | yr;
        d[q] = d[q] < pd ? d[q] : pd;
        p[q] = p[q] < pr ? p[q] : pr;
        dd = dd > (T)d[q] ? dd : (T)d[q];
        dp = dp > (T)p[q] ? dp : (T)p[q];
        L.Dg[i][j] = d[q];
      }
    } else {
      //   column panel C[i][j] = C[i][j] (+) C[i][k] (x) D[k][j]
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i4 + 8 * q;
        const U ad = canon(L.Dg[i][kk]);
        const U ac = canon(L.Pc[i][kk]);
        const U pd = (ad << md) | yd;
        const U pc = (ac << md) | yd;
        d[q] = d[q] < pd ? d[q] : pd;
        p[q] = p[q] < pc ? p[q] : pc;
        dd = dd > (T)d[q] ? dd : (T)d[q];
        dp = dp > (T)p[q] ? dp : (T)p[q];
        L.Dg[i][j] = d[q];
        L.Pc[i][j] = p[q];
      }
    }
    if (kk + 1 < BLK) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i4 + 8 * q;
        if (i == kk + 1) {
          const U cd = canon(d[q]);
          pair(cd, L.MYd[qq ^ 1][j][0], L.MYd[qq ^ 1][j][1]);
          L.x.w.sDr[kk + 1][j] = cd;
          if (rowp) {
            const U cr = canon(p[q]);
            pair(cr, L.MYr[qq ^ 1][j][0], L.MYr[qq ^ 1][j][1]);
            L.x.w.sPr[kk + 1][j] = cr;
          } else {
            L.x.w.sPr[kk + 1][j] = ALL;
          }
        }
        if (j == kk + 1) {
          L.x.w.sDc[kk + 1][i] = canon(d[q]);
          L.x.w.sPc[kk + 1][i] = rowp ? ALL : canon(p[q]);
        }
      }
    }
    __syncthreads();
  }
  // the MSB guard of each step, from its snapshots: left operands are the
  // column snapshots (diagonal, column panel), right operands the row
  // snapshots (diagonal, row panel); thread (i4, j) takes index j of steps
  // i4 + 8 q, the 32 lanes of a step combine by shuffles
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int kk = i4 + 8 * q;
    const int a1 = amsb(L.x.w.sDc[kk][j]), a2 = amsb(L.x.w.sPc[kk][j]);
    const U b1 = L.x.w.sDr[kk][j], b2 = L.x.w.sPr[kk][j];
    const int m1 = (T)b1 < 0 ? 0 : msb_u(b1), m2 = (T)b2 < 0 ? 0 : msb_u(b2);
    int ga = a1 > a2 ? a1 : a2, gb = m1 > m2 ? m1 : m2;
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
      const int ta = __shfl_xor(ga, o), tb = __shfl_xor(gb, o);
      ga = ga > ta ? ga : ta;
      gb = gb > tb ? gb : tb;
    }
    if (j == 0 && ga + gb > lim) L.sbad = 1;
  }
  // cycles: step x, vertex y with W[y][x] and W[x][y] both present
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = i4 + 8 * q;
    if (owner && i != j && (T)L.x.w.sDc[i][j] >= 0 && (T)L.x.w.sDr[i][j] >= 0) {
      record_cycle(st, naive_key(k0 + i, k0 + j, k0 + j, n));
      L.sbad = 1;
    }
  }
  __syncthreads();
  if (L.sbad) {
    if (tid == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
  // tiles and snapshots out, coalesced (thread (i, j): snapshot step i, index j)
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);
  auto put = [](T* at, T v) {
    if constexpr (DEV)
      st_dev(at, v);
    else
      *at = v;
  };
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = i4 + 8 * q;
    if (owner) {
      put(&W[(k0 + i) * ld + k0 + j], (T)(i == j ? L.sDiag[i] : d[q]));
      put(&colT[i * n + k0 + j], (T)L.x.w.sDc[i][j]);
      put(&rowS[i * n + k0 + j], (T)L.x.w.sDr[i][j]);
    }
    if (rowp)
      put(&W[(k0 + i) * ld + j0 + j], (T)p[q]);
    else
      put(&W[(j0 + i) * ld + k0 + j], (T)p[q]);
    put(&rowS[i * n + j0 + j], (T)L.x.w.sPr[i][j]);
    put(&colT[i * n + j0 + j], (T)L.x.w.sPc[i][j]);
  }
  // (no capture of the previous block's detections: a detection in this
  // schedule asks for general mode, whose run reports them)
  // trajectory maximum: the workgroup's, then one atomic
  T m = dp > ds ? dp : ds;
  if (owner) {
    m = m > dd ? m : dd;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = i4 + 8 * q;
      if (i == j && (T)L.sDiag[i] > m) m = (T)L.sDiag[i];
    }
  }
  const int64_t wm = wave_max_i64((int64_t)m);
  if ((tid & 63) == 0) L.sMax[tid >> 6] = wm;
  __syncthreads();
  if (tid == 0) {
    int64_t x = 0;
    for (int w = 0; w < NT / 64; w++) x = x > L.sMax[w] ? x : L.sMax[w];
    if (x > seen) record_diameter(st, x);
  }
}

// One pivot step of the look-ahead schedule (lower triangular matrices, fast
// mode): npanel workgroups run the panel of pivot block K + 1
// (fw_panel_quad_body, which first applies block K to block K + 1's rows),
// the others the phase-3 tiles of block K in rectangle r with those rows
// and columns left to the panel (ahead).  Neither role reads what the other
// writes: in a lower triangular matrix block K changes no cell of block
// K + 1's columns, and the tiles skip block K + 1's rows.  So a pivot step is
// one launch and the panel runs beside the tiles instead of before them.
// npanel = 0: the last step, tiles only.  The tiles check the MSB guard on
// the operands they stage; the two roles share one LDS allocation.
template <typename T, int RC, bool SHIFT_BYTES, int PER_CU>
__global__ __launch_bounds__(NT, PER_CU) void k_fw_step_fast(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                            int width, rbg_status* st, TileRect r, int npanel) {
  using U = typename unsigned_of<T>::type;
  __shared__ union {
    TileShared<U, 16 * RC, SHIFT_BYTES> t;
    QuadShared<U> p;
  } L;
  // int32: the tiles take the first workgroups and the panel the last.  A
  // launch of the 128-tiles holds one to two rounds; the panel then starts
  // as the first round drains instead of holding a third of its slots
  // (12.54 against 12.87 ms per N = 8192 closure).  int64: the panel first.
  // Its launches hold several rounds of 64-tiles, and a panel dispatched
  // behind them would end the launch late (7.50 against 6.38 ms per
  // N = 4096 closure).
  constexpr bool PANEL_LAST = sizeof(T) == 4;
  const int b = blockIdx.x;
  const int ntile = (int)gridDim.x - npanel;
  const int pb = PANEL_LAST ? b - ntile : b;  // this workgroup's panel, if in [0, npanel)
  if (pb >= 0 && pb < npanel) {
    // the panel is the launch's critical path: its waves take issue
    // priority over the tile waves that share their SIMDs
    __builtin_amdgcn_s_setprio(3);
    fw_panel_quad_body<T>(L.p, W, n, ld, K + 1, ws, width, st, pb);
    return;
  }
  const int t = PANEL_LAST ? b : b - npanel;
  fw_tile_fast_body<T, RC, RC, false, SHIFT_BYTES>(L.t, W, n, ld, K, ws, width, st, r.i0 + t / r.jn,
                                                   r.j0 + t % r.jn, npanel > 0);
}

// Triangular schedule (the blocked form of the reference's
// assume_upper_triangular loop, _shortest_path.pyx:365-377).  When every
// off-diagonal entry lies below the diagonal (tri = 1), a product
// W[i][k] (x) W[k][j] can be present only for i >= k >= j, and closing keeps
// the matrix lower triangular; so pivot block [k0, k0 + 32) changes only
// cells with i > k0 and j < k0 + 31, outside its own rows and columns.  The
// tiles of edge `edge` that hold such cells form the rectangle of tile rows
// k0 / edge .. nt - 1 and tile columns 0 .. (k0 + 31) / edge; for tri = 2
// (upper) it is the mirror image.  Every other tile would only take products
// with an absent operand, which leave its cells, and the trajectory maximum,
// as they are.  Returns the rectangle and its tile count.
__host__ __device__ inline TileRect tri_rect(int tri, int64_t n, int edge, int64_t K, int* count) {
  const int nt = (int)(n / edge);
  const int64_t k0 = K * BLK;
  const int lo = (int)(k0 / edge), hi = (int)((k0 + BLK - 1) / edge);
  TileRect r{0, 0, nt};
  if (tri == 1) {
    r = {lo, 0, hi + 1};
    *count = (nt - lo) * (hi + 1);
  } else if (tri == 2) {
    r = {0, lo, nt - lo};
    *count = (hi + 1) * (nt - lo);
  } else {
    *count = nt * nt;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Persistent closure of a lower triangular matrix (fast mode): one launch for
// every pivot step.  In a lower triangular matrix a product W[i][k] (x)
// W[k][j] can be present only for i >= k >= j, so:
//   - no step before K's own changes diagonal block K or block K's columns:
//     k_fw_diag_walk closes every diagonal block before the launch, leaving
//     each step's snapshots in dsnap; block K's column panel needs only
//     those and the input;
//   - in block K's column panel, row i at step kk reads its own cell
//     (i, kk), which no earlier step of the block changed: every row is an
//     independent loop, and the column snapshots are the input's values;
//   - in block K's row panel, row kk is final at step kk (later steps would
//     need entries above the diagonal): every column is an independent
//     forward substitution, and the row snapshots are its final values;
//   - a tile's snapshot segments above block K's rows and right of its
//     columns are absent.
// Work items, one per workgroup, in ticket order:
//   P(0), then for K = 0 .. nb - 1: P(K + 1) (while K + 1 < nb), T(K)
// P(K) = the column-panel items C(K, g) (64 rows each, from block K + 1
// down) and the row-panel items R(K, J), J < K (block (K, J), which first
// takes pivot block K - 1's products, as the tiles of step K - 1 would have),
// in the order C(K, 0) (block K + 1's rows, which R(K + 1, .) read), R(K, 0 ..
// K - 1), C(K, 1 ..).
// T(K) = the phase-3 tiles of block K's triangular rectangle (tri_rect), the
// tile row holding block K + 2 first.  An item waits only for the items that
// last wrote what it reads (flow_cond), all of smaller tickets, so the
// smallest unfinished item can always run: the waits form no cycle, whatever
// workgroups are resident.  No cell has two writers at once: the tiles of
// step K skip block K's and block K + 1's rows and store no column right of
// block K (a lower triangular step leaves those cells unchanged).  A tile may
// read a cell right of block K that a panel is writing; either value is on
// the cell's trajectory, the step's products there are absent, and the value
// is not stored.
// Cells and snapshots are written with device-coherent stores (through the
// XCD's L2); a finished item drains its stores, then publishes its counters.
// A waiting item polls the counters with device-scope loads, under a time
// limit (a wait that runs out sets RBG_F_STALLED | RBG_F_NEED_GENERAL and
// every workgroup leaves), then invalidates its XCD's L2 before reading.
// ---------------------------------------------------------------------------
constexpr int FLOW_COL_ROWS = 64;  // rows of one column-panel item: 4 threads per row

struct FlowCtl {
  // int32 words, zero before the launch: [2] stall, then pdone[nb],
  // tdone[nb], tstep[nt * nt], pflag[nb][nb]:
  //   pdone[K]: P(K) items finished; tdone[K]: T(K) tiles finished;
  //   tstep[I nt + J]: 1 + the last step tile (I, J) finished;
  //   pflag[K][idx]: 1 once item idx of P(K) finished
  int* c;
  int nb, nt, edge;
  int64_t total;  // items of the closure
  // the items in ticket order, packed (flow_pack), or null: then each
  // workgroup finds its item by flow_item's search
  const uint32_t* items = nullptr;
  __host__ __device__ int* stall() const { return c + 2; }
  __host__ __device__ int64_t pdone(int K) const { return 4 + K; }
  __host__ __device__ int64_t tdone(int K) const { return 4 + nb + K; }
  __host__ __device__ int64_t tstep(int I, int J) const { return 4 + 2 * (int64_t)nb + (int64_t)I * nt + J; }
  __host__ __device__ int64_t pflag(int K, int idx) const {
    return 4 + 2 * (int64_t)nb + (int64_t)nt * nt + (int64_t)K * nb + idx;
  }
  __host__ __device__ static int64_t words(int nb, int nt) { return 4 + 2 * (int64_t)nb + (int64_t)nt * nt + (int64_t)nb * nb; }
  __host__ __device__ int sh() const { return edge == 128 ? 2 : 1; }  // log2 of 32-blocks per tile edge
  // step K's rectangle: tile rows lo(K) .. nt - 1, tile columns 0 .. lo(K)
  __host__ __device__ int lo(int K) const { return K >> sh(); }
  __host__ __device__ int64_t tiles(int K) const { return (int64_t)(nt - lo(K)) * (lo(K) + 1); }
  // the tile row of step K taken first: block K + 2's, whose row panel waits for it
  __host__ __device__ int prow(int K) const { return K + 2 < nb ? lo(K + 2) : lo(K); }
  // column-panel items of P(K): the rows below block K, two blocks at a time
  __host__ __device__ int ncol(int K) const { return (nb - K) >> 1; }
  __host__ __device__ int np(int K) const { return K + ncol(K); }
  // P(K)'s item numbers: C(K, 0), R(K, 0 .. K - 1), C(K, 1 ..)
  __host__ __device__ int ridx(int K, int J) const { return J + (ncol(K) > 0 ? 1 : 0); }
  __host__ __device__ int cidx(int K, int g) const { return g == 0 ? 0 : K + g; }
  // item idx of P(K): a row panel (sets J) or a column panel (sets g)
  __host__ __device__ bool is_row(int K, int idx, int& J, int& g) const {
    const int c = ncol(K) > 0 ? 1 : 0;
    if (c && idx == 0) { g = 0; return false; }
    if (idx < K + c) { J = idx - c; return true; }
    g = idx - K;
    return false;
  }
  // sum over m = 1 .. M of ceil(m / 2), M >= -1
  __host__ __device__ static int64_t s2(int64_t M) {
    if (M <= 0) return 0;
    const int64_t Q = M >> 1, R = M & 1;
    return Q * (Q + 1) + R * (Q + 1);
  }
  // items of P(0) .. P(K): ncol(K') = ceil((nb - 1 - K') / 2)
  __host__ __device__ int64_t psum(int K) const { return (int64_t)K * (K + 1) / 2 + s2(nb - 1) - s2(nb - 2 - K); }
  // tiles of T(0) .. T(K - 1): with 2^sh pivot blocks per tile row and
  // K = Q 2^sh + r, steps q 2^sh .. q 2^sh + 2^sh - 1 share lo = q, so the
  // sum is 2^sh F(Q) + r (nt - Q)(Q + 1), F(Q) = sum over q < Q of (nt - q)(q + 1)
  __host__ __device__ int64_t tsum(int K) const {
    const int64_t Q = K >> sh(), r = K & ((1 << sh()) - 1), N = nt;
    const int64_t F = N * Q * (Q - 1) / 2 + N * Q - (Q - 1) * Q * (2 * Q - 1) / 6 - Q * (Q - 1) / 2;
    return (F << sh()) + r * (N - Q) * (Q + 1);
  }
  // first ticket of group g = P(g + 1), T(g), 0 <= g < nb
  __host__ __device__ int64_t before(int g) const { return psum(g) + tsum(g); }
  __host__ __device__ int64_t count() const { return psum(nb - 1) + tsum(nb); }
};

// Ticket x (0 <= x < total) -> its item: role 0, item a of P(K) (FlowCtl::
// is_row tells which panel); role 1, tile (a, b) of step K.
struct FlowItem {
  int role, K, a, b;
};
__host__ __device__ __forceinline__ FlowItem flow_item(const FlowCtl& f, int64_t x) {
  if (x < f.np(0)) return FlowItem{0, 0, (int)x, 0};
  int lo = 0, hi = f.nb - 1;  // the group of ticket x
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (f.before(mid) <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int g = lo;
  uint32_t off = (uint32_t)(x - f.before(g));
  if (g + 1 < f.nb) {
    if (off < (uint32_t)f.np(g + 1)) return FlowItem{0, g + 1, (int)off, 0};
    off -= f.np(g + 1);
  }
  const int l = f.lo(g), pr = f.prow(g);
  const uint32_t jn = (uint32_t)l + 1;
  if (off < jn) return FlowItem{1, g, pr, (int)off};
  off -= jn;
  const int a = l + (int)(off / jn);
  return FlowItem{1, g, a >= pr ? a + 1 : a, (int)(off % jn)};
}

// The waits of an item: condition i (0 <= i < FLOW_MAXW) is "counter word
// `word` >= target"; returns false past the last one.
//   P(K), K >= 4: all of P(K - 3) and T(K - 4), the last readers of the
//     snapshot buffer P(K) writes.  A column-panel item waits for nothing
//     else.  Row panel (K, J) also waits for what its first part reads --
//     block J's row snapshots of step K - 1 (item (K - 1, J); the diagonal
//     block's, J = K - 1, come from dsnap) and block K's column snapshots
//     of step K - 1 (column item 0 of P(K - 1), which also wrote block
//     (K, K - 1)) -- and for the tile of step K - 2 that holds block (K, J),
//     the last tile to write it (the tiles of step K - 1 skip block K's rows).
//   T(K, I, J): the row-panel items of P(K) for tile column J's blocks left
//     of block K, the column-panel items for tile row I's blocks below it,
//     and tile (I, J)'s step K - 1 when that step covered it.
// An item packed into 32 bits (role 1, K 10, a 11, b 10): a table of them
// spares every workgroup flow_item's search (several us of dependent scalar
// arithmetic per item).  Orders up to 32 x FLOW_PACK_MAX_NB.
constexpr int FLOW_PACK_MAX_NB = 1024;
__host__ __device__ inline uint32_t flow_pack(const FlowItem& q) {
  return (uint32_t)q.role << 31 | (uint32_t)q.K << 21 | (uint32_t)q.a << 10 | (uint32_t)q.b;
}
__host__ __device__ inline FlowItem flow_unpack(uint32_t v) {
  return FlowItem{(int)(v >> 31), (int)((v >> 21) & 1023), (int)((v >> 10) & 2047), (int)(v & 1023)};
}

constexpr int FLOW_MAXW = 10;
__host__ __device__ inline bool flow_cond(const FlowCtl& f, const FlowItem& q, int i, int64_t& word, int& target) {
  const int bpt = 1 << f.sh(), K = q.K;
  if (q.role == 0) {
    if (K >= 4) {
      if (i == 0) { word = f.pdone(K - 3); target = f.np(K - 3); return true; }
      if (i == 1) { word = f.tdone(K - 4); target = (int)f.tiles(K - 4); return true; }
      i -= 2;
    }
    int J = 0, g = 0;
    if (!f.is_row(K, q.a, J, g)) return false;  // a column-panel item
    if (J < K - 1) {
      if (i == 0) { word = f.pflag(K - 1, f.ridx(K - 1, J)); target = 1; return true; }
      i--;
    }
    if (i == 0) { word = f.pflag(K - 1, f.cidx(K - 1, 0)); target = 1; return true; }
    if (i == 1 && K >= 2 && (J >> f.sh()) <= f.lo(K - 2)) {
      word = f.tstep(f.lo(K), J >> f.sh());
      target = K - 1;
      return true;
    }
    return false;
  }
  const int I = q.a, J = q.b;
  // row panels for tile column J's blocks s < K
  for (int s = J * bpt; s < J * bpt + bpt && s < K; s++) {
    if (i == 0) { word = f.pflag(K, f.ridx(K, s)); target = 1; return true; }
    i--;
  }
  // column panels for tile row I's blocks s > K: C(K, (s - K - 1) / 2)
  int last = -1;
  for (int s = I * bpt; s < I * bpt + bpt; s++) {
    if (s <= K) continue;
    const int g = (s - K - 1) >> 1;
    if (g == last) continue;
    last = g;
    if (i == 0) { word = f.pflag(K, f.cidx(K, g)); target = 1; return true; }
    i--;
  }
  if (i == 0 && K >= 1 && J <= f.lo(K - 1)) { word = f.tstep(I, J); target = K; return true; }
  return false;
}

__device__ inline int flow_load(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Wave 0 of a workgroup: lane i waits until condition i of the item holds
// (counters only grow), all lanes polling together.  False when a wait ran
// out (then the closure is marked stalled) or another workgroup's did.
__device__ __forceinline__ bool flow_wait(const FlowCtl& f, const FlowItem& q, int64_t limit, rbg_status* st) {
  const int lane = threadIdx.x & 63;
  int64_t word = 0;
  int target = 0;
  const bool mine = lane < FLOW_MAXW && flow_cond(f, q, lane, word, target);
  // lane FLOW_MAXW reads the stall word in the same round trip
  const bool watch = lane == FLOW_MAXW;
  if (watch) word = f.stall() - f.c;
  bool timing = false;
  int64_t t0 = 0;
  for (;;) {
    const int v = (mine || watch) ? flow_load(f.c + word) : 0;
    if (__any(watch && v != 0)) return false;
    if (__all(!mine || v >= target)) return true;
    const int64_t now = wall_clock64();
    if (!timing) {
      timing = true;
      t0 = now;
    } else if (now - t0 > limit) {
      if (lane == 0) {
        __hip_atomic_store(f.stall(), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicOr(&st->flags, (uint32_t)(RBG_F_STALLED | RBG_F_NEED_GENERAL));
      }
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Every diagonal block of a lower triangular matrix closed before the
// persistent launch (block K: workgroup K, one cell per thread): the pair
// panel's walk of the diagonal block alone.  Leaves the closed block in W and
// each step's snapshots in dsnap[K]; checks the MSB guard of every product,
// finds cycles from the snapshots, and records the block's trajectory
// maximum.
template <typename T>
__device__ __forceinline__ void fw_diag_walk_body(T* W, int64_t n, int64_t ld, T* dsnap, int width, rbg_status* st,
                                                  int64_t K) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr U ALL = ~(U)0;
  __shared__ U Dg[BLK][BLK + 1];
  __shared__ U sDc[BLK][BLK + 1];  // column kk at step kk, row x
  __shared__ U sDr[BLK][BLK + 1];  // row kk at step kk, column x
  __shared__ U MY[2][BLK][2];      // (shift, low bits) of the pivot row, by step parity
  __shared__ U sDiag[BLK];
  __shared__ int64_t sMax[PANEL_NT / 64];
  __shared__ int sbad;
  const int tid = threadIdx.x, i = tid >> 5, j = tid & (BLK - 1);
  const int64_t k0 = K * BLK;
  const int lim = width - 2;
  auto canon = [](U x) -> U { return (T)x < 0 ? ALL : x; };
  auto pair = [](U b, U& m, U& y) {
    if ((T)b < 0) {
      m = 0;
      y = TOP;
    } else {
      m = (U)msb_u(b);
      y = b ^ ((U)1 << m);
    }
  };
  U d = canon((U)W[(k0 + i) * ld + k0 + j]);
  if (tid == 0) sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
  if (i == j) {
    sDiag[i] = d;
    d = ALL;
  }
  Dg[i][j] = d;
  if (i == 0) {
    pair(d, MY[0][j][0], MY[0][j][1]);
    sDr[0][j] = d;
  }
  if (j == 0) sDc[0][i] = d;
  __syncthreads();
  if (sbad) return;
  const int64_t seen = st->diameter_max;
  T dd = 0;
  int g = 0;
  for (int kk = 0; kk < BLK; kk++) {
    const int q = kk & 1;
    const U ad = canon(Dg[i][kk]);
    const U md = MY[q][j][0], yd = MY[q][j][1];
    const int gd = ((T)ad > 1 ? msb_u(ad) : 0) + (int)md;
    g = g > gd ? g : gd;
    const U pd = (ad << md) | yd;
    d = d < pd ? d : pd;
    dd = dd > (T)d ? dd : (T)d;
    Dg[i][j] = d;
    if (kk + 1 < BLK) {
      if (i == kk + 1) {
        const U cd = canon(d);
        pair(cd, MY[q ^ 1][j][0], MY[q ^ 1][j][1]);
        sDr[kk + 1][j] = cd;
      }
      if (j == kk + 1) sDc[kk + 1][i] = canon(d);
    }
    __syncthreads();
  }
  if (g > lim) sbad = 1;
  // cycles: step x, vertex y with W[y][x] and W[x][y] both present
  if (i != j && (T)sDc[i][j] >= 0 && (T)sDr[i][j] >= 0) {
    record_cycle(st, naive_key(k0 + i, k0 + j, k0 + j, n));
    sbad = 1;
  }
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
  W[(k0 + i) * ld + k0 + j] = (T)(i == j ? sDiag[i] : d);
  dsnap[K * 2 * BLK * BLK + i * BLK + j] = (T)sDc[i][j];
  dsnap[K * 2 * BLK * BLK + BLK * BLK + i * BLK + j] = (T)sDr[i][j];
  T m = dd;
  if (i == j && (T)sDiag[i] > m) m = (T)sDiag[i];
  const int64_t wm = wave_max_i64((int64_t)m);
  if ((tid & 63) == 0) sMax[tid >> 6] = wm;
  __syncthreads();
  if (tid == 0) {
    int64_t x = 0;
    for (int w = 0; w < PANEL_NT / 64; w++) x = x > sMax[w] ? x : sMax[w];
    if (x > seen) record_diameter(st, x);
  }
}

// Column panel of pivot block K in the persistent closure (lower triangular):
// rows r = 32 (K + 1) + 64 g + t / 4.  Row r at step kk of the block reads
// its own cell (r, kk), which no earlier step of the block changed, and
// block K's diagonal row snapshot kk (dsnap[K]):
//     C[r][x] = C[r][x] (+) C[r][kk] (x) D_kk[kk][x],   x < kk
// (products with x >= kk have an absent right operand).  Every cell's
// sequence needs only the row as the closure found it, so the four threads
// of a row take every fourth cell (x = t % 4 + 4 m).  The step's column
// snapshot of row r is C[r][kk] as the closure found it.  Writes the cells
// and the column snapshots with device-coherent stores.
template <typename U>
struct FlowColShared {
  U C0[FLOW_COL_ROWS][BLK + 1];  // the rows as the closure found them
  U M[BLK][BLK];                 // (shift, low bits) of the diagonal row snapshots [kk][x]
  U Y[BLK][BLK];
  int G[BLK];                    // per kk, the largest shift among them
  int sbad;
};
template <typename T, bool DEV = true>
__device__ __forceinline__ void fw_flow_col_body(FlowColShared<typename unsigned_of<T>::type>& L, T* W, int64_t n,
                                                 int64_t ld, int K, int g, FwWork ws, int width, rbg_status* st) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr U ALL = ~(U)0;
  auto canon = [](U x) -> U { return (T)x < 0 ? ALL : x; };
  auto put = [](T* at, T v) {
    if constexpr (DEV)
      st_dev(at, v);
    else
      *at = v;
  };
  const int tid = threadIdx.x, h = tid & 3, row = tid >> 2;
  const int64_t k0 = (int64_t)K * BLK;
  const int64_t r = k0 + BLK + (int64_t)g * FLOW_COL_ROWS + row;
  const bool live = r < n;
  U* Wu = (U*)W;
  {
    // thread (row, h) loads cells 8 h .. 8 h + 7 of its row
    U v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = ALL;
    if (live) {
      ld4(Wu + r * ld + k0 + 8 * h, *reinterpret_cast<U(*)[4]>(&v[0]));
      ld4(Wu + r * ld + k0 + 8 * h + 4, *reinterpret_cast<U(*)[4]>(&v[4]));
    }
#pragma unroll
    for (int q = 0; q < 8; q++) L.C0[row][8 * h + q] = canon(v[q]);
  }
  const U* drow = (const U*)ws.dsnap + (int64_t)K * 2 * BLK * BLK + BLK * BLK;
  for (int e = tid; e < BLK * BLK; e += NT) {
    const U b = drow[e];
    const bool abs = (T)b < 0;
    const int m = abs ? 0 : msb_u(b);
    L.M[e / BLK][e % BLK] = (U)m;
    L.Y[e / BLK][e % BLK] = abs ? TOP : b ^ ((U)1 << m);
  }
  if (tid == 0) L.sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 1 : 0;
  __syncthreads();
  if (tid < BLK) {
    int gm = 0;
    for (int x = 0; x < tid; x++) gm = (int)L.M[tid][x] > gm ? (int)L.M[tid][x] : gm;
    L.G[tid] = gm;
  }
  __syncthreads();
  if (L.sbad) return;
  const int64_t seen = st->diameter_max;
  T dm = 0;
  int ga = 0;
  if (live) {
    T* colT = ws.col<T>(ws.slot(K));
    U c[BLK / 4];  // cells h + 4 m
#pragma unroll
    for (int m = 0; m < BLK / 4; m++) {
      c[m] = L.C0[row][h + 4 * m];
      put(&colT[(h + 4 * m) * n + r], (T)c[m]);
    }
    for (int kk = 1; kk < BLK; kk++) {
      const U a = L.C0[row][kk];  // as the closure found it: no earlier step of the block changed it
      const int am = (T)a > 1 ? msb_u(a) : 0;
      ga = am + L.G[kk] > ga ? am + L.G[kk] : ga;
#pragma unroll
      for (int m = 0; m < BLK / 4; m++) {
        const int x = h + 4 * m;
        if (x < kk) {
          const U p = (a << L.M[kk][x]) | L.Y[kk][x];
          c[m] = c[m] < p ? c[m] : p;
          dm = dm > (T)c[m] ? dm : (T)c[m];
        }
      }
    }
#pragma unroll
    for (int m = 0; m < BLK / 4; m++) put((T*)&Wu[r * ld + k0 + h + 4 * m], (T)c[m]);
  }
  if (ga > width - 2) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
  const int64_t wm = wave_max_i64((int64_t)dm);
  if ((tid & 63) == 0 && wm > seen) record_diameter(st, wm);
}

// Row panel (K, J), J < K, of the persistent closure: block K's rows in block
// J's columns.  First pivot block K - 1's products (as its tiles would have
// taken them; they skip block K's rows), from step K - 1's snapshots:
//     R[i][j] = R[i][j] (+) colsnap_kk[k0 + i] (x) rowsnap_kk[j0 + j]
// then block K's steps: row kk of the panel is final at step kk (a later step
// would need an entry above the diagonal), so each column is a forward
// substitution,
//     R[i][j] = R[i][j] (+) D_kk[i][kk] (x) R[kk][j],   i > kk,
// D_kk[i][kk] being block K's diagonal column snapshots (dsnap[K]).  The
// step's row snapshot is R[kk][j].  Thread t takes column t >> 3, rows
// (t & 7) + 8 q: the eight threads of a column share a wave, and row kk's
// value goes to them by a shuffle, with no barrier between steps.
template <typename U>
struct FlowRowShared {
  U R[BLK][BLK + 1];        // the panel, for the transposition
  U sL[BLK][BLK + 1];       // step K - 1's column snapshots of block K's rows [kk][i]
  U sY[BLK][BLK + 1];       // step K - 1's row snapshots of block J's columns: low bits [kk][j]
  uint8_t sM[BLK][BLK + 4]; // and shifts
  U Dc[BLK][BLK + 1];       // block K's diagonal column snapshots [kk][i]
  int gL[BLK];              // per kk, the largest MSB among sL[kk][*]
  int gD[BLK];              // and among Dc[kk][i], i > kk
  int64_t sMax[NT / 64];
  int sbad;
};
template <typename T, bool DEV = true>
__device__ __forceinline__ void fw_flow_row_body(FlowRowShared<typename unsigned_of<T>::type>& L, T* W, int64_t n,
                                                 int64_t ld, int K, int J, FwWork ws, int width, rbg_status* st) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr U ALL = ~(U)0;
  auto canon = [](U x) -> U { return (T)x < 0 ? ALL : x; };
  auto amsb = [](U a) -> int { return (T)a > 1 ? msb_u(a) : 0; };
  auto put = [](T* at, T v) {
    if constexpr (DEV)
      st_dev(at, v);
    else
      *at = v;
  };
  const int tid = threadIdx.x;
  const int64_t k0 = (int64_t)K * BLK, j0 = (int64_t)J * BLK;
  constexpr int PER = BLK * BLK / NT;  // 4
  // every load first: the panel (coalesced), step K - 1's snapshots, block K's diagonal snapshots
  const U* pcol = (const U*)ws.col<T>(ws.slot(K - 1)) + k0;
  const U* prow;
  int64_t prs;  // the row snapshots' stride: the diagonal block's come from dsnap
  if (J == K - 1) {
    prow = (const U*)ws.dsnap + (int64_t)(K - 1) * 2 * BLK * BLK + BLK * BLK;
    prs = BLK;
  } else {
    prow = (const U*)ws.row<T>(ws.slot(K - 1)) + j0;
    prs = n;
  }
  const U* dcol = (const U*)ws.dsnap + (int64_t)K * 2 * BLK * BLK;
  U w[PER], la[PER], lb[PER], ld_[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + q * NT, y = e / BLK, x = e % BLK;
    w[q] = ((const U*)W)[(k0 + y) * ld + j0 + x];
    la[q] = pcol[y * n + x];
    lb[q] = prow[y * prs + x];
    ld_[q] = dcol[y * BLK + x];
  }
  if (tid == 0) L.sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 1 : 0;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = tid + q * NT, y = e / BLK, x = e % BLK;
    L.R[y][x] = canon(w[q]);
    L.sL[y][x] = canon(la[q]);
    const U b = lb[q];
    const bool abs = (T)b < 0;
    const int m = abs ? 0 : msb_u(b);
    L.sM[y][x] = (uint8_t)m;
    L.sY[y][x] = abs ? TOP : b ^ ((U)1 << m);
    L.Dc[y][x] = canon(ld_[q]);
  }
  __syncthreads();
  if (L.sbad) return;
  // the MSB guard per step: every column meets every row, so the largest left
  // operand MSB of step kk plus a column's shift bounds that column's products
  if (tid < 2 * BLK) {
    const int kk = tid & (BLK - 1);
    int g = 0;
    for (int x = 0; x < BLK; x++) {
      const int a = tid < BLK ? amsb(L.sL[kk][x]) : (x > kk ? amsb(L.Dc[kk][x]) : 0);
      g = a > g ? a : g;
    }
    (tid < BLK ? L.gL : L.gD)[kk] = g;
  }
  __syncthreads();
  const int64_t seen = st->diameter_max;
  const int lim = width - 2;
  const int j = tid >> 3, i4 = tid & 7;
  U r[4];
#pragma unroll
  for (int q = 0; q < 4; q++) r[q] = L.R[i4 + 8 * q][j];
  T dm = 0;
  int ga = 0;
  // pivot block K - 1
#pragma unroll 4
  for (int kk = 0; kk < BLK; kk++) {
    const U m = L.sM[kk][j], y = L.sY[kk][j];
    const int mb = (T)y < 0 ? 0 : (int)m;
    const int gk = L.gL[kk] + mb;
    ga = gk > ga ? gk : ga;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const U a = L.sL[kk][i4 + 8 * q];
      const U v = (a << m) | y;
      r[q] = r[q] < v ? r[q] : v;
      dm = dm > (T)r[q] ? dm : (T)r[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) r[q] = canon(r[q]);
  // pivot block K: forward substitution down each column
  T* rowS = ws.row<T>(ws.slot(K));
  const int base = (threadIdx.x & 63) & ~7;
#pragma unroll
  for (int kk = 0; kk < BLK; kk++) {
    const U b = canon((U)__shfl((T)r[kk >> 3], base + (kk & 7), 64));
    if (i4 == (kk & 7)) put(&rowS[kk * n + j0 + j], (T)b);
    const bool abs = (T)b < 0;
    const U m = abs ? 0 : (U)msb_u(b);
    const U y = abs ? TOP : b ^ ((U)1 << m);
    const int gk = L.gD[kk] + (int)m;
    ga = gk > ga ? gk : ga;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (i4 + 8 * q > kk) {
        const U a = L.Dc[kk][i4 + 8 * q];
        const U v = (a << m) | y;
        r[q] = r[q] < v ? r[q] : v;
        dm = dm > (T)r[q] ? dm : (T)r[q];
      }
    }
  }
  if (ga > lim) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
#pragma unroll
  for (int q = 0; q < 4; q++) put(&W[(k0 + i4 + 8 * q) * ld + j0 + j], (T)r[q]);
  const int64_t wm = wave_max_i64((int64_t)dm);
  if ((threadIdx.x & 63) == 0 && wm > seen) record_diameter(st, wm);
}

template <typename T>
__global__ __launch_bounds__(PANEL_NT) void k_fw_diag_walk(T* W, int64_t n, int64_t ld, T* dsnap, int width,
                                                           rbg_status* st) {
  fw_diag_walk_body<T>(W, n, ld, dsnap, width, st, blockIdx.x);
}

static __global__ void k_flow_zero(int* c, int64_t words) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    c[i] = 0;
}

// The packed item table of a closure (FlowCtl::items), one thread per ticket.
static __global__ void k_flow_table(FlowCtl f, uint32_t* items) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < f.total; x += (int64_t)gridDim.x * blockDim.x)
    items[x] = flow_pack(flow_item(f, x));
}

// The shared memory of one persistent work item: a tile's, a row panel's or a
// column panel's.
template <typename U, int RC, bool SHIFT_BYTES>
struct FlowShared {
  union {
    TileShared<U, 16 * RC, SHIFT_BYTES> t;
    FlowRowShared<U> r;
    FlowColShared<U> c;
  };
  int role;
};

// Work item q of the persistent closure of one matrix: waits for the items
// that last wrote what it reads, runs, then publishes its counters.
// limit: the longest wait, in wall-clock ticks.
template <typename T, int RC, bool SHIFT_BYTES>
__device__ __forceinline__ void fw_flow_item(FlowShared<typename unsigned_of<T>::type, RC, SHIFT_BYTES>& L,
                                             const FlowItem& q, T* W, int64_t n, int64_t ld, const FwWork& ws,
                                             int width, rbg_status* st, const FlowCtl& f, int64_t limit) {
  if (threadIdx.x < 64) {
    bool ok = flow_wait(f, q, limit, st);
    if (threadIdx.x == 0) {
      // the items waited for published their cells: drop stale lines
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      L.role = ok ? q.role : -1;
    }
  }
  __syncthreads();
  if (L.role < 0) return;
  const int K = q.K, a = q.a, b = q.b;
  if (q.role == 0) {
    // the panels are the closure's critical path: their waves take issue
    // priority over the tile waves that share their SIMDs
    __builtin_amdgcn_s_setprio(3);
    int J = 0, g = 0;
    if (f.is_row(K, a, J, g))
      fw_flow_row_body<T>(L.r, W, n, ld, K, J, ws, width, st);
    else
      fw_flow_col_body<T>(L.c, W, n, ld, K, g, ws, width, st);
  } else {
    fw_tile_fast_body<T, RC, RC, false, SHIFT_BYTES, true>(L.t, W, n, ld, K, ws, width, st, a, b, 2);
  }
  // publication: the item's cells and snapshots went out with device-coherent
  // stores; once every wave's stores are complete, the counters may say so
  wait_vm();
  __syncthreads();
  if (threadIdx.x == 0) {
    int* c = f.c;
    if (q.role == 0) {
      __hip_atomic_store(c + f.pflag(K, a), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(c + f.pdone(K), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(c + f.tstep(a, b), K + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(c + f.tdone(K), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// One work item per workgroup: the grid holds every item of the closure in
// ticket order, and workgroup x takes ticket x.  The hardware starts each
// XCD's workgroups in index order, so a workgroup that has not started
// leaves every later one of its XCD unstarted too: the smallest unfinished
// ticket is always running, or starts as soon as an earlier one ends.
template <typename T, int RC, bool SHIFT_BYTES, int PER_CU>
__global__ __launch_bounds__(NT, PER_CU) void k_fw_flow(T* W, int64_t n, int64_t ld, FwWork ws, int width,
                                                       rbg_status* st, FlowCtl f, int64_t limit) {
  __shared__ FlowShared<typename unsigned_of<T>::type, RC, SHIFT_BYTES> L;
  const FlowItem q = f.items ? flow_unpack(f.items[blockIdx.x]) : flow_item(f, blockIdx.x);
  fw_flow_item<T, RC, SHIFT_BYTES>(L, q, W, n, ld, ws, width, st, f, limit);
}

// ---------------------------------------------------------------------------
// Phase 3, general mode (natural domain, full arithmetic and event keys).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT) void k_fw_tile_general(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                        int width, int is_uns, rbg_status* st) {
  __shared__ T As[BLK][TILE];
  __shared__ T Bs[BLK][TILE];
  const int tid = threadIdx.x;
  const int nt = (int)(n / TILE);
  const int I = blockIdx.x / nt, J = blockIdx.x % nt;
  const int64_t k0 = K * BLK;
  const int par = ws.slot(K);
  const T* colT = ws.col<T>(par);
  const T* rowS = ws.row<T>(par);
  for (int e = tid; e < BLK * TILE; e += NT) {
    int kk = e / TILE, x = e % TILE;
    As[kk][x] = colT[kk * n + (int64_t)I * TILE + x];
    Bs[kk][x] = rowS[kk * n + (int64_t)J * TILE + x];
  }
  __syncthreads();
  const int tr = tid >> 4, tc = tid & 15;
  const int64_t r0 = (int64_t)I * TILE + tr * RPT;
  const int64_t c0 = (int64_t)J * TILE + tc * CPT;
  if ((r0 >= k0 && r0 < k0 + BLK) || (c0 >= k0 && c0 < k0 + BLK)) return;
  T w[RPT][CPT];
#pragma unroll
  for (int r = 0; r < RPT; r++)
#pragma unroll
    for (int c = 0; c < CPT; c++) w[r][c] = W[(r0 + r) * ld + c0 + c];
  int64_t dmax = 0;
  for (int kk = 0; kk < BLK; kk++) {
    const int64_t k = k0 + kk;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      T a = As[kk][tr * RPT + r];
#pragma unroll
      for (int c = 0; c < CPT; c++) {
        uint32_t ev = 0;
        T p = avos_product<T>(a, Bs[kk][tc * CPT + c], width, is_uns != 0, ev);
        if (ev) record_overflow(st, naive_key(k, r0 + r, c0 + c, n), ev);
        if (r0 + r == c0 + c && !(p == (T)-1 || p == 0 || p == 1))
          record_cycle(st, naive_key(k, r0 + r, r0 + r, n));
        T v = avos_sum<T>(w[r][c], p);
        w[r][c] = v;
        if ((int64_t)v > dmax) dmax = (int64_t)v;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RPT; r++)
#pragma unroll
    for (int c = 0; c < CPT; c++) W[(r0 + r) * ld + c0 + c] = w[r][c];
  dmax = wave_max_i64(dmax);
  if ((tid & 63) == 0) record_diameter(st, dmax);
}

template <typename T, bool FAST>
__global__ void k_fw_capture_last(rbg_status* st, FwWork ws, int64_t K, int64_t n) {
  const int par = ws.slot(K);
  capture_block<T, FAST>(st, ws.col<T>(par), ws.row<T>(par), K * BLK, n);
}

// Workspace layouts: FwWork's snapshot buffers (two, or four for the
// persistent closure) and guard words; for the lower triangular schedules
// the diagonal-block copy (FwWork::dcopy, n x 32); for the persistent
// closure the diagonal blocks' step snapshots, its counters (FlowCtl; tile
// rows counted for the smallest tile edge, 64) and its packed item table.
// rbg_fw_workspace_bytes is the first, rbg_fw_tri_workspace_bytes the
// largest.
static inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }
static inline int64_t dcopy_offset(int64_t n, int elem, int ring) { return align256(FwWork::bytes(n, elem, ring)); }
static inline int64_t la_ws_bytes(int64_t n, int elem) { return dcopy_offset(n, elem, 2) + n * BLK * elem; }
static inline int64_t dsnap_offset(int64_t n, int elem) { return align256(dcopy_offset(n, elem, 4) + n * BLK * elem); }
static inline int64_t flow_ctl_offset(int64_t n, int elem) { return align256(dsnap_offset(n, elem) + 2 * n * BLK * elem); }
static inline int64_t flow_items_offset(int64_t n, int elem) {
  return align256(flow_ctl_offset(n, elem) + FlowCtl::words((int)(n / BLK), (int)(n / 64)) * 4);
}

// The persistent closure's schedule for order n with tiles of `edge` (128 or
// 64), counters in `wsp` (no item table).
static inline FlowCtl flow_ctl(int64_t n, int elem, void* wsp, int edge) {
  FlowCtl f{wsp ? (int*)((char*)wsp + flow_ctl_offset(n, elem)) : nullptr, (int)(n / BLK), (int)(n / edge), edge, 0};
  f.total = f.count();
  return f;
}

// The single-matrix choice: int32 takes 128-tiles (64-tiles up to
// SMALL_TILE_MAX_N), int64 64-tiles; the item table follows the counters.
static inline int flow_edge(int64_t n, int elem) { return (elem == 4 && n > SMALL_TILE_MAX_N) ? 128 : 64; }
static inline FlowCtl flow_ctl(int64_t n, int elem, void* wsp) {
  FlowCtl f = flow_ctl(n, elem, wsp, flow_edge(n, elem));
  if (wsp && f.nb <= FLOW_PACK_MAX_NB) f.items = (const uint32_t*)((char*)wsp + flow_items_offset(n, elem));
  return f;
}
static inline int64_t flow_ws_bytes(int64_t n, int elem) {
  return flow_items_offset(n, elem) + 4 * flow_ctl(n, elem, nullptr).total;
}

// Relaxations a persistent closure of order n executes, counted in cells x 32
// steps: the tiles' cells outside the skipped rows; a row panel's strip (one
// block) and forward substitution (496 of a block's cells per 32 steps), a
// column panel's 496 per 32 rows (the diagonal blocks run before the launch).
static inline double flow_relaxations(const FlowCtl& f, int64_t n) {
  double cells = 0;
  for (int K = 0; K < f.nb; K++) {
    const int jn = f.lo(K) + 1;
    cells += (double)f.tiles(K) * f.edge * f.edge - (double)(K + 1 < f.nb ? 2 : 1) * BLK * jn * f.edge;
    cells += (double)K * (BLK * BLK + 496) + (double)(n - (int64_t)(K + 1) * BLK) * 496 / BLK;
  }
  return cells * BLK;
}

// The longest wait of one persistent work item: two seconds of the device's
// wall clock (a step's tiles take well under a millisecond at N = 8192), or
// the test override (rbg_fw_flow_wait_limit).
extern std::atomic<int64_t> g_flow_limit_override;
static inline int64_t flow_limit_ticks() {
  const int64_t o = g_flow_limit_override.load();
  if (o > 0) return o;
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz < 1)
    khz = 100000;
  return (int64_t)khz * 1000 * 2;
}

}  // namespace rbg
