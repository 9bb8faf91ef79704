// Shared device helpers: the status block, event recording and the
// "E domain" used by the fast Floyd-Warshall kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <utility>
#include <vector>

#include "../../include/rbg_hip.h"
#include "avos.hpp"

#define RBG_CHECK_LAUNCH()                         \
  do {                                             \
    hipError_t e_ = hipGetLastError();             \
    if (e_ != hipSuccess) return (int)e_;          \
  } while (0)

namespace rbg {

// ---- optional per-launch timing ---------------------------------------------
// While enabled, selected launches are bracketed by hipEvents on their stream
// (bench and profiling only).  Kinds: 0 FW panel, 1 FW tile (128 x 128 for
// int32), 2 DAG level, 3 SpGEMM rows, 4 batched FW, 5 FW tile (64 x 64 int32
// tiles of the small-matrix and triangular schedules).  Each launch may
// carry a count of the work it does (relaxations for the tiles).
constexpr int PROF_KINDS = 8;
struct Prof {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, int>> pairs[PROF_KINDS];
  std::vector<double> work[PROF_KINDS];
  int used = 0;
  std::mutex mu;
  hipEvent_t take() {
    if (used == (int)pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
};
inline Prof g_prof;

inline int prof_start(hipStream_t s) {
  if (!g_prof.on) return -1;
  std::lock_guard<std::mutex> g(g_prof.mu);
  int idx = g_prof.used;
  hipEvent_t e = g_prof.take();
  if (!e) return -1;
  (void)hipEventRecord(e, s);
  return idx;
}
inline void prof_stop(int kind, int start_idx, hipStream_t s, double work = 0.0) {
  if (start_idx < 0) return;
  std::lock_guard<std::mutex> g(g_prof.mu);
  int idx = g_prof.used;
  hipEvent_t e = g_prof.take();
  if (!e) return;
  (void)hipEventRecord(e, s);
  g_prof.pairs[kind].push_back({start_idx, idx});
  g_prof.work[kind].push_back(work);
}

// ---- event recording -------------------------------------------------------
__device__ inline void record_cycle(rbg_status* st, uint64_t key) {
  atomicMin((unsigned long long*)&st->cycle_first_key, (unsigned long long)key);
  atomicMax((unsigned long long*)&st->cycle_last_key, (unsigned long long)key);
  atomicAdd((unsigned long long*)&st->cycle_count, 1ull);
}

__device__ inline void record_overflow(rbg_status* st, uint64_t key, uint32_t ev) {
  atomicOr(&st->events, ev);
  atomicMin((unsigned long long*)&st->ov_first_key, (unsigned long long)key);
  atomicMax((unsigned long long*)&st->ov_last_key, (unsigned long long)key);
}

__device__ inline void record_diameter(rbg_status* st, int64_t v) {
  if (v > 0) atomicMax((long long*)&st->diameter_max, (long long)v);
}

// Wave-level max reduction (wave64) followed by one atomic per wave.
__device__ inline int64_t wave_max_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    int64_t o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// ---- E domain --------------------------------------------------------------
// Fast Floyd-Warshall keeps the matrix in an encoding where the AVOS sum is a
// single unsigned min and the trajectory maximum a single signed max:
//   value v >= 1  ->  v
//   0 (no edge)   ->  all ones  (loses every unsigned min; negative for max)
//   RED_ONE (-1)  ->  0         (wins every unsigned min: RED_ONE absorbs)
// A left operand x in this encoding already behaves as the product needs: an
// absent x (all ones) shifted left keeps its top bit, so (x << m) | ylow stays
// "absent" for every shift m < bits.
template <typename T>
__host__ __device__ inline T e_enc(T v) {
  return v >= 1 ? v : (v == 0 ? (T)-1 : (T)0);
}
template <typename T>
__host__ __device__ inline T e_dec(T e) {
  return e == 0 ? (T)-1 : (e < 0 ? (T)0 : e);
}

}  // namespace rbg
