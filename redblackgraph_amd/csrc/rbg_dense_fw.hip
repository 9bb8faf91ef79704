// Dense AVOS transitive closure of one matrix: the host-side schedule of the
// snapshot-blocked Floyd-Warshall kernels (fw_kernels.hpp) and the C ABI.
#include "fw_kernels.hpp"

namespace rbg {

std::atomic<int64_t> g_flow_limit_override{0};

// 128-tile rectangles below this many workgroups (a third of a round of
// three per CU) run as 64-tiles instead, in the triangular schedule: four
// times the workgroups for the steps that touch few tiles.
constexpr int TRI_SMALL_BELOW = 256;

// Lower triangular closures with a workspace of rbg_fw_tri_workspace_bytes
// take the persistent closure (k_fw_flow): one launch for all pivot blocks
// (one box: N = 8192 int32 10.87 ms, N = 4096 int64 4.34 ms per closure);
// with a smaller workspace, or tri = 3, the look-ahead schedule
// (k_fw_step_fast): one launch per pivot block, the next block's panel
// inside it (12.56 and 6.36 ms).
constexpr bool LOOKAHEAD = true;
constexpr bool FLOW = true;

// tri: 0 = full schedule; 1 / 2 = the input's off-diagonal entries all lie
// below / above the diagonal (fast mode only: general mode always runs the
// full schedule); 3 = below, with one launch per pivot step even when the
// workspace holds the persistent closure's.
template <typename T, bool PROF = true>
static int fw_dense(T* W, int64_t n, int64_t ld, int width, int is_uns, int mode, int tri, void* wsp,
                    int64_t ws_bytes, rbg_status* st, hipStream_t s) {
  if (!W || !wsp || !st || n <= 0 || n % TILE || ld < n || tri < 0 || tri > 3 || mode < 0 || mode > 2)
    return RBG_E_ARG;
  if (ws_bytes < FwWork::bytes(n, (int)sizeof(T))) return RBG_E_WORKSPACE;
  // mode 2: fast, W written by rbg_fw_load_* (already in the E domain)
  const bool loaded = mode == 2;
  const bool fast = mode == 0 || loaded;
  if (fast && (width != (int)sizeof(T) * 8 || is_uns)) return RBG_E_ARG;
  const bool per_step = tri == 3;
  if (per_step) tri = 1;
  if (!fast) tri = 0;
  const int elem = (int)sizeof(T);
  // Look-ahead (lower triangular): pivot block K's tiles and block K + 1's
  // panel share a launch; only block 0's panel runs on its own.  With room
  // for it, the persistent closure instead.
  const bool la = fast && tri == 1 && LOOKAHEAD && ws_bytes >= la_ws_bytes(n, elem);
  const bool flow = la && FLOW && !per_step && ws_bytes >= flow_ws_bytes(n, elem);
  FwWork ws{wsp, n, elem, flow ? 4 : 2};
  const int nb = (int)(n / BLK);
  const int nt = (int)(n / TILE);
  const int conv_grid = (int)((n * n + NT * 8 - 1) / (NT * 8)) < 4096 ? (int)((n * n + NT * 8 - 1) / (NT * 8)) : 4096;
  if (loaded) {
    hipLaunchKernelGGL(k_fw_take_initial, dim3(1), dim3(1), 0, s, st, tri);
    RBG_CHECK_LAUNCH();
  } else if (fast) {
    hipLaunchKernelGGL((k_e_convert<T, true>), dim3(conv_grid), dim3(NT), 0, s, W, n, ld, tri ? st : nullptr, tri);
    RBG_CHECK_LAUNCH();
  }
  if (flow) {
    // every diagonal block closed up front (none changes before its step)
    ws.dsnap = (char*)wsp + dsnap_offset(n, elem);
    hipLaunchKernelGGL((k_fw_diag_walk<T>), dim3(nb), dim3(PANEL_NT), 0, s, W, n, ld, (T*)ws.dsnap, width, st);
    RBG_CHECK_LAUNCH();
  } else if (la) {
    ws.dcopy = (char*)wsp + dcopy_offset(n, elem, ws.ring);
    hipLaunchKernelGGL((k_fw_diag_copy<T>), dim3(nb), dim3(NT), 0, s, W, ld, (T*)ws.dcopy);
    RBG_CHECK_LAUNCH();
  }
  const int panel_grid = nb > 1 ? 2 * (nb - 1) : 1;
  auto panel = [&](int K, hipStream_t q) -> int {
    int p0 = PROF ? prof_start(q) : -1;
    if (fast) {
      hipLaunchKernelGGL((k_fw_panel_fast<T>), dim3(nb - 1), dim3(PANEL_NT), 0, q, W, n, ld, (int64_t)K, ws, width,
                         st);
    } else {
      hipLaunchKernelGGL((k_fw_panel<T, false>), dim3(panel_grid), dim3(NT), 0, q, W, n, ld, (int64_t)K, ws,
                         width, is_uns, st);
    }
    RBG_CHECK_LAUNCH();
    if (PROF) prof_stop(0, p0, q);
    return 0;
  };
  // the tiles of pivot block K, and with the look-ahead schedule (la) the
  // panel of block K + 1 in the same launch
  auto tiles = [&](int K, hipStream_t q, bool la) -> int {
    int p1 = PROF ? prof_start(q) : -1;
    int edge = TILE, count = nt * nt, kind = 1;
    TileRect r{0, 0, nt};
    const bool ahead = la && K + 1 < nb;
    if (fast) {
      edge = FastTile<T>::EDGE;
      r = tri_rect(tri, n, edge, K, &count);
      if constexpr (sizeof(T) == 4) {
        if (n <= SMALL_TILE_MAX_N || (tri && count < TRI_SMALL_BELOW)) {
          edge = 64;
          r = tri_rect(tri, n, edge, K, &count);
          kind = 5;
        }
      }
      if (la) {
        // the next block's panel: one workgroup per row/column panel tile
        // pair, as the pair panel
        const int npanel = ahead ? nb - 1 : 0;
        if constexpr (sizeof(T) == 4) {
          if (kind == 5)
            hipLaunchKernelGGL((k_fw_step_fast<T, 4, false, 4>), dim3(npanel + count), dim3(NT), 0, q, W, n, ld,
                               (int64_t)K, ws, width, st, r, npanel);
          else
            hipLaunchKernelGGL((k_fw_step_fast<T, 8, false, 3>), dim3(npanel + count), dim3(NT), 0, q, W, n, ld,
                               (int64_t)K, ws, width, st, r, npanel);
        } else {
          // int64: the 64 x 64 tile at three workgroups per CU (the int64
          // panel's LDS leaves no room for a fourth)
          hipLaunchKernelGGL((k_fw_step_fast<T, 4, true, 3>), dim3(npanel + count), dim3(NT), 0, q, W, n, ld,
                             (int64_t)K, ws, width, st, r, npanel);
        }
      } else if (kind == 5) {
        hipLaunchKernelGGL((k_fw_tile_fast_small<T>), dim3(count), dim3(NT), 0, q, W, n, ld, (int64_t)K, ws, width,
                           st, r);
      } else {
        hipLaunchKernelGGL((k_fw_tile_fast<T>), dim3(count), dim3(NT), 0, q, W, n, ld, (int64_t)K, ws, width, st,
                           r);
      }
    } else {
      hipLaunchKernelGGL((k_fw_tile_general<T>), dim3(nt * nt), dim3(NT), 0, q, W, n, ld, (int64_t)K, ws,
                         width, is_uns, st);
    }
    RBG_CHECK_LAUNCH();
    // relaxations the launch executes, 32 steps per cell: the tiles' cells
    // outside the skipped rows (the pivot block's, and with a panel in the
    // launch the next block's, are in every rectangle); a fused panel adds
    // the diagonal block, one live panel tile per workgroup and the strip
    // tiles left of the next pivot block
    if (PROF) {
      double cells = (double)count * edge * edge - (double)(ahead ? 2 : 1) * BLK * r.jn * edge;
      if (ahead) cells += (double)BLK * BLK * (1 + (nb - 1) + (K + 1));
      prof_stop(kind, p1, q, cells * BLK);
    }
    return 0;
  };
  if (flow) {
    const FlowCtl f = flow_ctl(n, elem, wsp);
    const int edge = f.edge;
    if (f.total > 0x7fffffffLL) return RBG_E_ARG;  // one workgroup per item
    const int64_t words = FlowCtl::words(nb, f.nt);
    hipLaunchKernelGGL(k_flow_zero, dim3((int)((words + NT - 1) / NT < 1024 ? (words + NT - 1) / NT : 1024)),
                       dim3(NT), 0, s, f.c, words);
    RBG_CHECK_LAUNCH();
    if (f.items) {
      const int tg = (int)((f.total + NT - 1) / NT < 1024 ? (f.total + NT - 1) / NT : 1024);
      hipLaunchKernelGGL(k_flow_table, dim3(tg), dim3(NT), 0, s, f, (uint32_t*)f.items);
      RBG_CHECK_LAUNCH();
    }
    const int64_t limit = flow_limit_ticks();
    const dim3 grid((unsigned)f.total);
    int p = PROF ? prof_start(s) : -1;
    if constexpr (sizeof(T) == 8)
      hipLaunchKernelGGL((k_fw_flow<T, 4, true, 3>), grid, dim3(NT), 0, s, W, n, ld, ws, width, st, f, limit);
    else if (edge == 128)
      hipLaunchKernelGGL((k_fw_flow<T, 8, false, 3>), grid, dim3(NT), 0, s, W, n, ld, ws, width, st, f, limit);
    else
      hipLaunchKernelGGL((k_fw_flow<T, 4, false, 4>), grid, dim3(NT), 0, s, W, n, ld, ws, width, st, f, limit);
    RBG_CHECK_LAUNCH();
    if (PROF) prof_stop(6, p, s, flow_relaxations(f, n));
  } else {
    for (int K = 0; K < nb; K++) {
      int rc = (!la || K == 0) ? panel(K, s) : 0;
      if (!rc) rc = tiles(K, s, la);
      if (rc) return rc;
    }
  }
  if (fast) {
    hipLaunchKernelGGL((k_fw_capture_last<T, true>), dim3(1), dim3(1), 0, s, st, ws, (int64_t)(nb - 1), n);
    RBG_CHECK_LAUNCH();
    if (ld % 4 == 0 && ((uintptr_t)W & 15) == 0)
      hipLaunchKernelGGL((k_e_decode4<T>), dim3((unsigned)(n < 4096 ? n : 4096)), dim3(NT), 0, s, W, n, ld);
    else
      hipLaunchKernelGGL((k_e_convert<T, false>), dim3(conv_grid), dim3(NT), 0, s, W, n, ld, nullptr);
    RBG_CHECK_LAUNCH();
  } else {
    hipLaunchKernelGGL((k_fw_capture_last<T, false>), dim3(1), dim3(1), 0, s, st, ws, (int64_t)(nb - 1), n);
    RBG_CHECK_LAUNCH();
  }
  return 0;
}

template <typename T>
static int fw_load(const T* A, int64_t n, int64_t lda, T* W, int64_t npad, rbg_status* st, hipStream_t s) {
  if (!A || !W || !st || n <= 0 || lda < n || npad < n || npad % 4 || ((uintptr_t)W & 15)) return RBG_E_ARG;
  const int grid = (int)(npad < 4096 ? npad : 4096);
  hipLaunchKernelGGL((k_fw_load<T>), dim3(grid), dim3(NT), 0, s, A, n, lda, W, npad, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int fw_precheck(const T* W, int64_t n, int64_t ld, rbg_status* st, hipStream_t s) {
  if (!W || !st || n <= 0 || ld < n) return RBG_E_ARG;
  int grid = (int)((n * n + NT * 8 - 1) / (NT * 8));
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL((k_fw_precheck<T>), dim3(grid), dim3(NT), 0, s, W, n, ld, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

// ---- allocation-free closure ----------------------------------------------
// The snapshot workspace of rbg_fw_dense_noalloc_* lives in this file-scope
// device array.  Host code never passes `scratch` itself to a kernel (on the
// host the name is only a stand-in for the symbol): each device's address of
// its own copy is looked up once, with that device current, and kept.
constexpr int64_t NOALLOC_SCRATCH_BYTES = 8912896;  // 8.5 MiB
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// This device's address of `scratch`, or an error code.
static int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return RBG_E_ARG;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

template <typename T>
static int fw_dense_noalloc(T* W, int64_t n, int64_t ld, int width, int is_uns, int mode, rbg_status* st,
                            hipStream_t s) {
  if (!W || !st || n <= 0 || n % TILE || ld < n) return RBG_E_ARG;
  const int64_t need = FwWork::bytes(n, (int)sizeof(T));
  if (need > NOALLOC_SCRATCH_BYTES) return RBG_E_WORKSPACE;
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  // every snapshot and guard word is written by the panel kernel of a step
  // before that step (or the next) reads it, so nothing needs clearing
  return fw_dense<T, false>(W, n, ld, width, is_uns, mode, 0, ws, NOALLOC_SCRATCH_BYTES, st, s);
}

}  // namespace rbg

using namespace rbg;

extern "C" {

int rbg_status_init(rbg_status* st, rbg_stream_t stream) {
  if (!st) return RBG_E_ARG;
  hipLaunchKernelGGL(k_status_init, dim3(1), dim3(1), 0, (hipStream_t)stream, st);
  RBG_CHECK_LAUNCH();
  return 0;
}

int64_t rbg_fw_workspace_bytes(int64_t n, int32_t elem_bytes) { return FwWork::bytes(n, elem_bytes); }

int64_t rbg_fw_flow_wait_limit(int64_t ticks) { return g_flow_limit_override.exchange(ticks > 0 ? ticks : 0); }

int64_t rbg_fw_tri_workspace_bytes(int64_t n, int32_t elem_bytes) { return flow_ws_bytes(n, elem_bytes); }

int64_t rbg_fw_flow_schedule(int64_t n, int32_t elem_bytes, int32_t edge, int32_t* items, int64_t cap) {
  if (n <= 0 || n % TILE || (elem_bytes != 4 && elem_bytes != 8) || (edge && edge != 64 && edge != 128))
    return RBG_E_ARG;
  const FlowCtl f = flow_ctl(n, elem_bytes, nullptr, edge ? edge : flow_edge(n, elem_bytes));
  if (items && cap >= 4 * f.total) {
    for (int64_t x = 0; x < f.total; x++) {
      const FlowItem q = flow_item(f, x);
      items[4 * x] = q.role;
      items[4 * x + 1] = q.K;
      items[4 * x + 2] = q.a;
      items[4 * x + 3] = q.b;
    }
  }
  return f.total;
}

int64_t rbg_fw_flow_waits(int64_t n, int32_t elem_bytes, int32_t edge, int64_t* waits, int64_t cap) {
  if (n <= 0 || n % TILE || (elem_bytes != 4 && elem_bytes != 8) || (edge && edge != 64 && edge != 128))
    return RBG_E_ARG;
  const FlowCtl f = flow_ctl(n, elem_bytes, nullptr, edge ? edge : flow_edge(n, elem_bytes));
  if (waits && cap >= 2 * FLOW_MAXW * f.total) {
    for (int64_t x = 0; x < f.total; x++) {
      const FlowItem q = flow_item(f, x);
      for (int i = 0; i < FLOW_MAXW; i++) {
        int64_t word = -1;
        int target = 0;
        if (!flow_cond(f, q, i, word, target)) word = -1;
        waits[2 * (FLOW_MAXW * x + i)] = word;
        waits[2 * (FLOW_MAXW * x + i) + 1] = target;
      }
    }
  }
  return f.total;
}

int rbg_prof_enable(int32_t on) {
  std::lock_guard<std::mutex> g(g_prof.mu);
  if (on == 1) {  // start afresh
    g_prof.used = 0;
    for (int k = 0; k < PROF_KINDS; k++) {
      g_prof.pairs[k].clear();
      g_prof.work[k].clear();
    }
  }
  g_prof.on = on != 0;  // 2 resumes and 0 pauses, keeping what was recorded
  bool empty = true;
  for (int k = 0; k < PROF_KINDS; k++) empty = empty && g_prof.pairs[k].empty();
  if (empty) g_prof.used = 0;
  return 0;
}

int rbg_prof_read_work(int32_t kind, double* total_ms, int64_t* launches, double* work) {
  if (kind < 0 || kind >= PROF_KINDS) return RBG_E_ARG;
  std::lock_guard<std::mutex> g(g_prof.mu);
  double tot = 0.0, wk = 0.0;
  for (auto& pr : g_prof.pairs[kind]) {
    hipError_t e = hipEventSynchronize(g_prof.pool[pr.second]);
    if (e != hipSuccess) return (int)e;
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, g_prof.pool[pr.first], g_prof.pool[pr.second]);
    if (e != hipSuccess) return (int)e;
    tot += ms;
  }
  for (double w : g_prof.work[kind]) wk += w;
  if (total_ms) *total_ms = tot;
  if (launches) *launches = (int64_t)g_prof.pairs[kind].size();
  if (work) *work = wk;
  g_prof.pairs[kind].clear();
  g_prof.work[kind].clear();
  bool empty = true;
  for (int k = 0; k < PROF_KINDS; k++) empty = empty && g_prof.pairs[k].empty();
  if (empty) g_prof.used = 0;
  return 0;
}

int rbg_prof_read(int32_t kind, double* total_ms, int64_t* launches) {
  return rbg_prof_read_work(kind, total_ms, launches, nullptr);
}

int rbg_fw_prof_enable(int32_t on) { return rbg_prof_enable(on); }

int rbg_fw_prof_read(double* panel_ms, int64_t* panel_launches, double* tile_ms, int64_t* tile_launches) {
  int rc = rbg_prof_read(0, panel_ms, panel_launches);
  if (rc) return rc;
  return rbg_prof_read(1, tile_ms, tile_launches);
}

int rbg_fw_load_i32(const int32_t* A, int64_t n, int64_t lda, int32_t* W, int64_t npad, rbg_status* st,
                    rbg_stream_t stream) {
  return fw_load<int32_t>(A, n, lda, W, npad, st, (hipStream_t)stream);
}
int rbg_fw_load_i64(const int64_t* A, int64_t n, int64_t lda, int64_t* W, int64_t npad, rbg_status* st,
                    rbg_stream_t stream) {
  return fw_load<int64_t>(A, n, lda, W, npad, st, (hipStream_t)stream);
}
int rbg_fw_precheck_i32(const int32_t* W, int64_t n, int64_t ld, rbg_status* st, rbg_stream_t stream) {
  return fw_precheck<int32_t>(W, n, ld, st, (hipStream_t)stream);
}
int rbg_fw_precheck_i64(const int64_t* W, int64_t n, int64_t ld, rbg_status* st, rbg_stream_t stream) {
  return fw_precheck<int64_t>(W, n, ld, st, (hipStream_t)stream);
}
int rbg_fw_dense_i32(int32_t* W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                     void* workspace, int64_t workspace_bytes, rbg_status* st, rbg_stream_t stream) {
  if (width != 8 && width != 16 && width != 32) return RBG_E_ARG;
  return fw_dense<int32_t>(W, n, ld, width, is_unsigned, mode, 0, workspace, workspace_bytes, st,
                           (hipStream_t)stream);
}
int rbg_fw_dense_i64(int64_t* W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                     void* workspace, int64_t workspace_bytes, rbg_status* st, rbg_stream_t stream) {
  if (width != 32 && width != 64) return RBG_E_ARG;
  return fw_dense<int64_t>(W, n, ld, width, is_unsigned, mode, 0, workspace, workspace_bytes, st,
                           (hipStream_t)stream);
}
int rbg_fw_dense_tri_i32(int32_t* W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                         int32_t tri, void* workspace, int64_t workspace_bytes, rbg_status* st,
                         rbg_stream_t stream) {
  if (width != 8 && width != 16 && width != 32) return RBG_E_ARG;
  return fw_dense<int32_t>(W, n, ld, width, is_unsigned, mode, tri, workspace, workspace_bytes, st,
                           (hipStream_t)stream);
}
int rbg_fw_dense_tri_i64(int64_t* W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                         int32_t tri, void* workspace, int64_t workspace_bytes, rbg_status* st,
                         rbg_stream_t stream) {
  if (width != 32 && width != 64) return RBG_E_ARG;
  return fw_dense<int64_t>(W, n, ld, width, is_unsigned, mode, tri, workspace, workspace_bytes, st,
                           (hipStream_t)stream);
}

int64_t rbg_fw_noalloc_max_bytes(void) { return NOALLOC_SCRATCH_BYTES; }
int rbg_fw_dense_noalloc_i32(int32_t* W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                             rbg_status* st, rbg_stream_t stream) {
  if (width != 8 && width != 16 && width != 32) return RBG_E_ARG;
  return fw_dense_noalloc<int32_t>(W, n, ld, width, is_unsigned, mode, st, (hipStream_t)stream);
}
int rbg_fw_dense_noalloc_i64(int64_t* W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                             rbg_status* st, rbg_stream_t stream) {
  if (width != 32 && width != 64) return RBG_E_ARG;
  return fw_dense_noalloc<int64_t>(W, n, ld, width, is_unsigned, mode, st, (hipStream_t)stream);
}

}  // extern "C"
