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
#include "rbg_common.hpp"

namespace rbg {

constexpr int BLK = 32;    // pivot block (k values per step)
constexpr int TILE = 128;  // phase-3 tile edge
constexpr int NT = 256;    // threads per workgroup
constexpr int RPT = 8;     // phase-3 rows per thread
constexpr int CPT = 8;     // phase-3 cols per thread

struct FwWork {
  // two snapshot buffers (block parity), each BLK x n for columns and rows,
  // then 2 x 2 x BLK guard words
  void* base;
  int64_t n;
  int elem;
  template <typename T>
  __host__ __device__ T* col(int parity) const {
    return (T*)base + (size_t)parity * 2 * BLK * n;
  }
  template <typename T>
  __host__ __device__ T* row(int parity) const {
    return (T*)base + (size_t)parity * 2 * BLK * n + (size_t)BLK * n;
  }
  __host__ __device__ int* guard(int parity) const {  // [0..BLK) col msb, [BLK..2BLK) row msb
    return (int*)((char*)base + (size_t)4 * BLK * n * elem) + parity * 2 * BLK;
  }
  static int64_t bytes(int64_t n, int elem) { return (int64_t)4 * BLK * n * elem + 4 * BLK * 4 + 256; }
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

static __global__ void k_guard_init(int* g) {
  if (threadIdx.x < 4 * BLK) g[threadIdx.x] = 0;
}

// Precondition of the fast kernels: every value >= -1, no off-diagonal +-1,
// every diagonal value in {-1, 0, 1}.  With such diagonals the products that
// involve a pivot cell (k, k) can neither overflow nor reveal a cycle, which
// is what lets the fast kernels skip the pivot row and column.
template <typename T>
__global__ void k_fw_precheck(const T* W, int64_t n, int64_t ld, rbg_status* st) {
  int64_t total = n * n;
  bool bad = false;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / n, j = e % n;
    T v = W[i * ld + j];
    if (v < (T)-1) bad = true;
    if (i != j && (v == (T)1 || v == (T)-1)) bad = true;
    if (i == j && v > (T)1) bad = true;
  }
  if (bad) atomicOr(&st->flags, (uint32_t)RBG_F_PRECOND_FAIL);
}

template <typename T, bool ENC>
__global__ void k_e_convert(T* W, int64_t n, int64_t ld) {
  int64_t total = n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / n, j = e % n;
    T v = W[i * ld + j];
    W[i * ld + j] = ENC ? e_enc<T>(v) : e_dec<T>(v);
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
  __shared__ int sguard[BLK];
  const int tid = threadIdx.x;
  const int nb = (int)(n / BLK);
  const int64_t k0 = K * BLK;
  const int par = (int)(K & 1);
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);
  int* guard = ws.guard(par);

  // panel role: 0 = diagonal only, 1 = row panel (K, J), 2 = column panel (J, K)
  int role = 0, J = 0;
  if (nb > 1) {
    int t = blockIdx.x;
    role = t < nb - 1 ? 1 : 2;
    int idx = role == 1 ? t : t - (nb - 1);
    J = idx < K ? idx : idx + 1;
  }
  const bool owner = blockIdx.x == 0;

  if (owner && tid == 0) {
    if (K > 0) capture_block<T, FAST>(st, ws.col<T>(par ^ 1), ws.row<T>(par ^ 1), k0 - BLK, n);
    int* gnext = ws.guard(par ^ 1);
    for (int q = 0; q < 2 * BLK; q++) gnext[q] = 0;
  }
  if (tid < BLK) sguard[tid] = 0;

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
      int mm = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        rowS[kk * n + pc0 + j0 + q] = snap_enc<T, FAST>(bP[q]);
        int m = bP[q] > 1 ? msb<T>(bP[q]) : 0;
        mm = m > mm ? m : mm;
      }
      atomicMax(&sguard[kk], mm);
    }
    if (role == 2 && (tid & 7) == 0) {
      colT[kk * n + pr0 + i] = snap_enc<T, FAST>(aP);
      atomicMax(&sguard[kk], aP > 1 ? msb<T>(aP) : 0);
    }
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
  if (tid < BLK) {
    int g = sguard[tid];
    if (g) atomicMax(&guard[(role == 1 ? BLK : 0) + tid], g);
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

// ---------------------------------------------------------------------------
// Panel kernel, fast mode, for one matrix (k_fw_panel_fast): same schedule as
// k_fw_panel, E-domain arithmetic.  2 (n/32 - 1) workgroups each walk the
// diagonal block together with one 32 x 32 tile of the row or column panel,
// which spreads a block's short dependent chain over the whole chip.  The
// panels also leave, per k, the largest MSB of the column and row snapshots
// in the guard words that phase 3 reads.
// At step k the pivot row and column do not change (they are skipped), so the
// step reads them while it updates every other cell and needs one barrier.
// Operands off the pivot row/column are never +-1 (precheck), so the product
// is (x << MSB(y)) | low(y) there.  A per-product MSB check raises
// RBG_F_NEED_GENERAL before any value could overflow.
// A product with an absent operand is some value with the top bit set; the
// min can store it into an absent cell.  Such a value must not be shifted
// again (the shift could clear its top bit), so every operand read here is
// first canonicalised to all ones, and so is every snapshot phase 3 reads.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void fw_panel_tiles_body(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws, int width,
                                          rbg_status* st, int bid) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  __shared__ U Dg[BLK][BLK + 1];
  __shared__ U Pn[BLK][BLK + 1];
  // snapshots of step kk, staged in LDS and written out after the loop so the
  // per-step barrier never waits on global stores
  __shared__ U sDc[BLK][BLK + 1];  // [kk][i] = diag W[i][kk]
  __shared__ U sDr[BLK][BLK + 1];  // [kk][j] = diag W[kk][j]
  __shared__ U sP[BLK][BLK + 1];   // row panel: [kk][j] = W[kk][j]; column panel: [kk][i] = W[i][kk]
  __shared__ int sbad;
  const int tid = threadIdx.x;
  const int nb = (int)(n / BLK);
  const int64_t k0 = K * BLK;
  const int par = (int)(K & 1);
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);
  int* guard = ws.guard(par);

  int role = 0, J = 0;
  if (nb > 1) {
    int t = bid;
    role = t < nb - 1 ? 1 : 2;
    int idx = role == 1 ? t : t - (nb - 1);
    J = idx < K ? idx : idx + 1;
  }
  const bool owner = bid == 0;
  // a workgroup-uniform exit once the matrix is known to need general mode
  if (tid == 0) sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
  __syncthreads();
  if (sbad) return;
  if (owner && tid == 0) {
    if (K > 0) capture_block<T, true>(st, ws.col<T>(par ^ 1), ws.row<T>(par ^ 1), k0 - BLK, n);
    int* gnext = ws.guard(par ^ 1);
    for (int q = 0; q < 2 * BLK; q++) gnext[q] = 0;
  }

  const int64_t pr0 = role == 2 ? (int64_t)J * BLK : k0;
  const int64_t pc0 = role == 1 ? (int64_t)J * BLK : k0;
  auto canon = [](U x) -> U { return (T)x < 0 ? ~(U)0 : x; };
  for (int e = tid; e < BLK * BLK; e += NT) {
    int r = e / BLK, c = e % BLK;
    Dg[r][c] = canon((U)W[(k0 + r) * ld + k0 + c]);
    if (role) Pn[r][c] = canon((U)W[(pr0 + r) * ld + pc0 + c]);
  }
  __syncthreads();

  const int i = tid >> 3;
  const int j0 = (tid & 7) * 4;
  const int lim = width - 2;
  U dv[4], pv[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    dv[q] = Dg[i][j0 + q];
    pv[q] = role ? Pn[i][j0 + q] : (U)0;
  }
  T dmax = 0;
  bool bad = false;
  for (int kk = 0; kk < BLK; kk++) {
    const int64_t k = k0 + kk;
    const U aD = canon(Dg[i][kk]);
    U bD[4], bP[4], aP = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) bD[q] = canon(Dg[kk][j0 + q]);
    if (role == 1) {
      aP = aD;
#pragma unroll
      for (int q = 0; q < 4; q++) bP[q] = canon(Pn[kk][j0 + q]);
    } else if (role == 2) {
      aP = canon(Pn[i][kk]);
#pragma unroll
      for (int q = 0; q < 4; q++) bP[q] = bD[q];
    }
    if ((tid & 7) == 0) {
      sDc[kk][i] = aD;
      if (role == 2) sP[kk][i] = aP;
    }
    if (i == 0) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        sDr[kk][j0 + q] = bD[q];
        if (role == 1) sP[kk][j0 + q] = bP[q];
      }
    }
    // diagonal tile: row kk and column kk are invariant at step kk
    const int maD = (T)aD > 1 ? msb_u(aD) : 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int j = j0 + q;
      if (i != kk && j != kk) {
        U m, y;
        if ((T)bD[q] < 0) {
          m = 0;
          y = TOP;
        } else {
          m = (U)msb_u(bD[q]);
          y = bD[q] & (((U)1 << m) - 1);
        }
        if ((T)bD[q] > 1 && maD + (int)m > lim) bad = true;
        U p = (aD << m) | y;
        if (owner && i == j && (T)p >= 0) record_cycle(st, naive_key(k, k0 + i, k0 + i, n));
        dv[q] = dv[q] < p ? dv[q] : p;
        Dg[i][j] = dv[q];
      }
      if (owner && (T)dv[q] > dmax) dmax = (T)dv[q];
    }
    if (role) {
      const int maP = (T)aP > 1 ? msb_u(aP) : 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = j0 + q;
        const bool pivot = role == 1 ? (i == kk) : (j == kk);
        if (!pivot) {
          U m, y;
          if ((T)bP[q] < 0) {
            m = 0;
            y = TOP;
          } else {
            m = (U)msb_u(bP[q]);
            y = bP[q] & (((U)1 << m) - 1);
          }
          if ((T)bP[q] > 1 && maP + (int)m > lim) bad = true;
          U p = (aP << m) | y;
          pv[q] = pv[q] < p ? pv[q] : p;
          Pn[i][j] = pv[q];
        }
        if ((T)pv[q] > dmax) dmax = (T)pv[q];
      }
    }
    __syncthreads();
  }
  if (bad) sbad = 1;
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
  // tiles and snapshots out, coalesced
  for (int e = tid; e < BLK * BLK; e += NT) {
    int r = e / BLK, c = e % BLK;
    if (owner) {
      W[(k0 + r) * ld + k0 + c] = (T)Dg[r][c];
      colT[r * n + k0 + c] = (T)sDc[r][c];
      rowS[r * n + k0 + c] = (T)sDr[r][c];
    }
    if (role == 1) {
      W[(pr0 + r) * ld + pc0 + c] = (T)Pn[r][c];
      rowS[r * n + pc0 + c] = (T)sP[r][c];
    } else if (role == 2) {
      W[(pr0 + r) * ld + pc0 + c] = (T)Pn[r][c];
      colT[r * n + pr0 + c] = (T)sP[r][c];
    }
  }
  // per-k MSB guard for phase 3 from this panel's snapshots
  if (role && tid < BLK) {
    int mm = 0;
    for (int c = 0; c < BLK; c++) {
      U v = sP[tid][c];
      int m = (T)v > 1 ? msb_u(v) : 0;
      mm = m > mm ? m : mm;
    }
    if (mm) atomicMax(&guard[(role == 1 ? BLK : 0) + tid], mm);
  }
  int64_t dm = wave_max_i64((int64_t)dmax);
  if ((tid & 63) == 0) record_diameter(st, dm);
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
template <typename T>
__device__ inline void fw_panel_lines_body(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws, int width,
                                          rbg_status* st, int bid) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  __shared__ U D[BLK][BLK + 1];
  __shared__ U sC[BLK][BLK + 1];  // [kk][r] = W[r][kk] at step kk (diagonal block)
  __shared__ U sR[BLK][BLK + 1];  // [kk][c] = W[kk][c] at step kk (diagonal block)
  __shared__ U Mb[BLK][BLK];      // shift and low bits of sR
  __shared__ U Yb[BLK][BLK];
  __shared__ int maxA[BLK], maxB[BLK], sbad;
  const int tid = threadIdx.x;
  const int64_t k0 = K * BLK;
  const int par = (int)(K & 1);
  const bool owner = bid == 0;
  T* colT = ws.col<T>(par);
  T* rowS = ws.row<T>(par);
  auto canon = [](U x) -> U { return (T)x < 0 ? ~(U)0 : x; };
  if (tid == 0) sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
  __syncthreads();
  if (sbad) return;
  if (owner && tid == 0 && K > 0) capture_block<T, true>(st, ws.col<T>(par ^ 1), ws.row<T>(par ^ 1), k0 - BLK, n);
  for (int e = tid; e < BLK * BLK; e += NT) D[e / BLK][e % BLK] = canon((U)W[(k0 + e / BLK) * ld + k0 + e % BLK]);
  // this thread's panel line, loaded while the diagonal block is walked
  const bool rowp = (int64_t)bid * NT + tid < n;
  const int64_t line = rowp ? (int64_t)bid * NT + tid : (int64_t)bid * NT + tid - n;
  const bool active = (int64_t)bid * NT + tid < 2 * n && !(line >= k0 && line < k0 + BLK);
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
  for (int kk = 0; kk < BLK; kk++) {
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
      W[(k0 + r) * ld + k0 + cc] = (T)D[r][cc];
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
  if ((tid & 63) == 0) record_diameter(st, dm);
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
// GLOBAL_GUARD: the per-k MSB guard comes from the guard words the panel
// kernel left (one matrix); otherwise the tile computes it from the
// snapshots it stages (the batched launcher's panels leave none).
template <typename T, int RPT, int CPT, bool GLOBAL_GUARD>
__device__ inline void fw_tile_fast_body(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws, int width,
                                         rbg_status* st, int bid) {
  using U = typename unsigned_of<T>::type;
  constexpr U TOP = (U)1 << (sizeof(U) * 8 - 1);
  constexpr int TL = 16 * CPT;  // tile edge
  constexpr int NCH = CPT / 4;  // column chunks per thread
  static_assert(RPT == CPT && CPT % 4 == 0, "square thread blocks of whole chunks");
  __shared__ U As[BLK][TL];
  __shared__ U Ms[BLK][TL];
  __shared__ U Ys[BLK][TL];
  __shared__ int sbad, sGa[BLK], sGb[BLK];
  const int tid = threadIdx.x;
  const int nt = (int)(n / TL);
  const int I = bid / nt, J = bid % nt;
  const int64_t k0 = K * BLK;
  const int par = (int)(K & 1);
  const U* colT = (const U*)ws.col<T>(par);
  const U* rowS = (const U*)ws.row<T>(par);

  const int tr = tid >> 4, tc = tid & 15;
  const int64_t r0 = (int64_t)I * TL + tr * RPT;
  const int64_t cb = (int64_t)J * TL + tc * 4;  // column of chunk 0; chunk h adds 64 h
  const bool skip = r0 >= k0 && r0 < k0 + BLK;

  // issue the tile loads first
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

  // a workgroup-uniform exit when an earlier kernel already asked for general mode
  if (tid == 0) sbad = (st->flags & (RBG_F_NEED_GENERAL | RBG_F_PRECOND_FAIL)) ? 2 : 0;
  if (tid < BLK) sGa[tid] = sGb[tid] = 0;
  __syncthreads();
  if (GLOBAL_GUARD && tid < BLK) {
    const int* guard = ws.guard(par);
    if (guard[tid] + guard[BLK + tid] > width - 2) atomicOr(&sbad, 1);
  }
  // stage the snapshots; per k, the largest MSB among this tile's left
  // operands and among its right operands (the MSB guard of this tile)
  for (int e = tid; e < BLK * TL / 4; e += NT) {
    const int kk = e / (TL / 4), x = (e % (TL / 4)) * 4;
    U a[4], b[4], m[4], y[4];
    ld4(colT + kk * n + (int64_t)I * TL + x, a);
    ld4(rowS + kk * n + (int64_t)J * TL + x, b);
    int ga = 0, gb = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if ((T)b[q] < 0) {  // absent
        m[q] = 0;
        y[q] = TOP;
      } else {
        m[q] = (U)msb_u(b[q]);
        y[q] = b[q] & (((U)1 << m[q]) - 1);
      }
      const int mq = (int)m[q], aq = (T)a[q] > 1 ? msb_u(a[q]) : 0;
      gb = mq > gb ? mq : gb;
      ga = aq > ga ? aq : ga;
    }
    if (!GLOBAL_GUARD) {
      if (ga) atomicMax(&sGa[kk], ga);
      if (gb) atomicMax(&sGb[kk], gb);
    }
    st4(&As[kk][x], a);
    st4(&Ms[kk][x], m);
    st4(&Ys[kk][x], y);
  }
  __syncthreads();
  if (!GLOBAL_GUARD) {
    if (tid < BLK && sGa[tid] + sGb[tid] > width - 2) atomicOr(&sbad, 1);
    __syncthreads();
  }
  if (sbad) {
    if (tid == 0 && (sbad & 1)) atomicOr(&st->flags, (uint32_t)RBG_F_NEED_GENERAL);
    return;
  }
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
  for (int kk = 0; kk < BLK; kk += 2) {
    U a0[RPT], a1[RPT], m0[CPT], m1[CPT], y0[CPT], y1[CPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      a0[r] = As[kk][tr * RPT + r];
      a1[r] = As[kk + 1][tr * RPT + r];
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const int x = 64 * (c >> 2) + tc * 4 + (c & 3);
      m0[c] = Ms[kk][x];
      y0[c] = Ys[kk][x];
      m1[c] = Ms[kk + 1][x];
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
              const U p0 = (a0[dr + q] << m0[dc + q]) | y0[dc + q];
              const U p1 = (a1[dr + q] << m1[dc + q]) | y1[dc + q];
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
        U p0 = (a0[r] << m0[c]) | y0[c];
        U p1 = (a1[r] << m1[c]) | y1[c];
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
#pragma unroll
  for (int r = 0; r < RPT; r++)
#pragma unroll
    for (int h = 0; h < NCH; h++) {
      U q[4];
#pragma unroll
      for (int x = 0; x < 4; x++) q[x] = w[r][4 * h + x];
      st4(Wu + (r0 + r) * ld + cb + 64 * h, q);
    }
  int64_t dm = (int64_t)(d0 > d1 ? d0 : d1);
  dm = wave_max_i64(dm);
  if ((tid & 63) == 0) record_diameter(st, dm);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_fw_panel_fast(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                      int width, rbg_status* st) {
  fw_panel_tiles_body<T>(W, n, ld, K, ws, width, st, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_fw_tile_fast(T* W, int64_t n, int64_t ld, int64_t K, FwWork ws,
                                                     int width, rbg_status* st) {
  fw_tile_fast_body<T, RPT, CPT, true>(W, n, ld, K, ws, width, st, blockIdx.x);
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
  const int par = (int)(K & 1);
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
  const int par = (int)(K & 1);
  capture_block<T, FAST>(st, ws.col<T>(par), ws.row<T>(par), K * BLK, n);
}

}  // namespace rbg
