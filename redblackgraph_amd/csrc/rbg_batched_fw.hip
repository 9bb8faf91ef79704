// Batched dense closure: many independent matrices ("components") closed by one
// schedule of launches, recordable as a single hipGraph.
//
// Reference: the per-graph closure is warshall.c.in:50-107 (through
// core/redblack.py:49-50); the reference closes components one at a time
// (sparse/csgraph/transitive_closure.py:85-195 component_wise_closure).  Here
// every component is padded to a multiple of 128 and stored contiguously in a
// pool; components are ordered by padded order, largest first, so that at
// pivot-block step K the components still running (npad / 32 > K) form a
// prefix of every work list.  Step K is one panel launch (groups of 256
// panel lines) and one tile launch per tile shape, each over the work of all
// running components, so a launch carries the panels (or tiles) of thousands
// of components at once.
// Arithmetic, snapshots, event keys and the MSB guard are the single-matrix
// fast kernels' (fw_kernels.hpp), per component.  A component whose input is
// outside the fast domain, or that needs general arithmetic or saw a cycle,
// is flagged in its status; the caller reruns it through the single-matrix
// path.
#include <algorithm>
#include <numeric>

#include "fw_kernels.hpp"

namespace rbg {

struct BComp {
  int32_t* W;
  void* snap;
  rbg_status* st;
  int32_t npad;
  int32_t tile;
  int32_t tri;  // 0 full schedule; 1 / 2: entries declared below / above the diagonal only
};

struct Item {
  int32_t comp;
  int32_t bid;
};

constexpr int CHUNK = 128;  // precheck / domain conversion granule

// workgroups of 256 panel lines per component (2 npad lines: its row-panel
// columns and column-panel rows)
static inline int lines_groups(int npad) { return (2 * npad + NT - 1) / NT; }

static __global__ void k_bfw_init(const BComp* comps, int32_t count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  rbg_status* st = comps[c].st;
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

// Fast-domain precheck of one 128 x 128 chunk (rules of k_fw_precheck).
static __global__ __launch_bounds__(NT) void k_bfw_precheck(const BComp* comps, const Item* items) {
  __shared__ int sbad;
  const Item it = items[blockIdx.x];
  const BComp c = comps[it.comp];
  const int nc = c.npad / CHUNK;
  const int64_t i0 = (int64_t)(it.bid / nc) * CHUNK, j0 = (int64_t)(it.bid % nc) * CHUNK;
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  bool bad = false;
  for (int e = threadIdx.x; e < CHUNK * CHUNK; e += NT) {
    const int64_t i = i0 + e / CHUNK, j = j0 + e % CHUNK;
    const int32_t v = c.W[i * c.npad + j];
    if (v < -1) bad = true;
    if (i != j && (v == 1 || v == -1)) bad = true;
    if (i == j && v > 1) bad = true;
    // a component declared triangular must be: its schedule skips the rest
    if (v != 0 && ((c.tri == 1 && i < j) || (c.tri == 2 && i > j))) bad = true;
  }
  if (bad) sbad = 1;
  __syncthreads();
  if (threadIdx.x == 0 && sbad) atomicOr(&c.st->flags, (uint32_t)RBG_F_PRECOND_FAIL);
}

// Domain conversion of one 128 x 128 chunk.  Entering the E domain, a
// triangular component's initial values also enter its trajectory maximum
// (the triangular schedule never visits the tiles its pivot blocks cannot
// change; see k_e_convert).
template <bool ENC>
static __global__ __launch_bounds__(NT) void k_bfw_convert(const BComp* comps, const Item* items) {
  const Item it = items[blockIdx.x];
  const BComp c = comps[it.comp];
  const int nc = c.npad / CHUNK;
  const int64_t i0 = (int64_t)(it.bid / nc) * CHUNK, j0 = (int64_t)(it.bid % nc) * CHUNK;
  int64_t vmax = 0;
  for (int e = threadIdx.x; e < CHUNK * CHUNK; e += NT) {
    int32_t* p = c.W + (i0 + e / CHUNK) * c.npad + j0 + e % CHUNK;
    const int32_t v = *p;
    *p = ENC ? e_enc<int32_t>(v) : e_dec<int32_t>(v);
    vmax = v > vmax ? v : vmax;
  }
  if (ENC && c.tri) {
    vmax = wave_max_i64(vmax);
    if ((threadIdx.x & 63) == 0) record_diameter(c.st, vmax);
  }
}

// phases 1 and 2: item = (component, group of 256 panel lines)
static __global__ __launch_bounds__(NT) void k_bfw_panel(const BComp* comps, const Item* items, int64_t K) {
  const Item it = items[blockIdx.x];
  const BComp c = comps[it.comp];
  FwWork ws{c.snap, c.npad, 4};
  __shared__ LinesShared<uint32_t> L;
  const int64_t n = c.npad;
  LineRange lr{0, n, 0, n};
  if (c.tri) {  // the lines the step's tile rectangle reads snapshots at
    int cnt = 0;
    const TileRect r = tri_rect(c.tri, n, c.tile, K, &cnt);
    const int64_t e = c.tile;
    if (c.tri == 1)
      lr = LineRange{0, (int64_t)(r.j0 + r.jn) * e, (int64_t)r.i0 * e, n};
    else
      lr = LineRange{(int64_t)r.j0 * e, n, 0, (int64_t)(r.i0 + cnt / r.jn) * e};
  }
  fw_panel_lines_body<int32_t>(L, c.W, n, n, K, ws, 32, c.st, it.bid, lr);
}

template <int RC>
static __global__ __launch_bounds__(NT, 3) void k_bfw_tile(const BComp* comps, const Item* items, int64_t K) {
  const Item it = items[blockIdx.x];
  const BComp c = comps[it.comp];
  FwWork ws{c.snap, c.npad, 4};
  const int nt = c.npad / (16 * RC);
  __shared__ TileShared<uint32_t, 16 * RC, false> L;
  fw_tile_fast_body<int32_t, RC, RC, false, false>(L, c.W, c.npad, c.npad, K, ws, 32, c.st, it.bid / nt, it.bid % nt);
}

// Operands of the last block's detections (the panel of the next block would
// have copied them; there is none).
static __global__ void k_bfw_capture(const BComp* comps, int32_t count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  const BComp b = comps[c];
  const int64_t nb = b.npad / BLK;
  const int par = (int)((nb - 1) & 1);
  FwWork ws{b.snap, b.npad, 4};
  capture_block<int32_t, true>(b.st, ws.col<int32_t>(par), ws.row<int32_t>(par), (nb - 1) * BLK, b.npad);
}

}  // namespace rbg

using namespace rbg;

struct rbg_batch {
  int32_t count = 0;
  int32_t maxnb = 0;
  BComp* d_comps = nullptr;
  // chunks | panels | tiles 128 | tiles 64 (full-schedule components) |
  // triangular components' tiles 128, step by step | the same for tiles 64
  Item* d_items = nullptr;
  int64_t n_chunk = 0, n_panel = 0, n_t128 = 0, n_t64 = 0, n_tri128 = 0, n_tri64 = 0;
  std::vector<int64_t> panel_cnt, t128_cnt, t64_cnt;  // per step: items of running components
  std::vector<int64_t> tri128_off, tri64_off;         // per step: [off[K], off[K + 1]) of the step lists
  std::vector<double> work;                           // per step: relaxations its tile launches execute
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int batch_launch(rbg_batch* b, hipStream_t s) {
  const int cgrid = (b->count + 255) / 256;
  hipLaunchKernelGGL(k_bfw_init, dim3(cgrid), dim3(256), 0, s, b->d_comps, b->count);
  RBG_CHECK_LAUNCH();
  const Item* chunks = b->d_items;
  const Item* panels = chunks + b->n_chunk;
  const Item* t128 = panels + b->n_panel;
  const Item* t64 = t128 + b->n_t128;
  const Item* tri128 = t64 + b->n_t64;
  const Item* tri64 = tri128 + b->n_tri128;
  hipLaunchKernelGGL(k_bfw_precheck, dim3(b->n_chunk), dim3(NT), 0, s, b->d_comps, chunks);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_bfw_convert<true>), dim3(b->n_chunk), dim3(NT), 0, s, b->d_comps, chunks);
  RBG_CHECK_LAUNCH();
  for (int K = 0; K < b->maxnb; K++) {
    hipLaunchKernelGGL(k_bfw_panel, dim3(b->panel_cnt[K]), dim3(NT), 0, s, b->d_comps, panels, (int64_t)K);
    RBG_CHECK_LAUNCH();
    int pi = prof_start(s);
    if (b->t128_cnt[K]) {
      hipLaunchKernelGGL((k_bfw_tile<8>), dim3(b->t128_cnt[K]), dim3(NT), 0, s, b->d_comps, t128, (int64_t)K);
      RBG_CHECK_LAUNCH();
    }
    if (b->t64_cnt[K]) {
      hipLaunchKernelGGL((k_bfw_tile<4>), dim3(b->t64_cnt[K]), dim3(NT), 0, s, b->d_comps, t64, (int64_t)K);
      RBG_CHECK_LAUNCH();
    }
    const int64_t c128 = b->tri128_off[K + 1] - b->tri128_off[K], c64 = b->tri64_off[K + 1] - b->tri64_off[K];
    if (c128) {
      hipLaunchKernelGGL((k_bfw_tile<8>), dim3(c128), dim3(NT), 0, s, b->d_comps, tri128 + b->tri128_off[K],
                         (int64_t)K);
      RBG_CHECK_LAUNCH();
    }
    if (c64) {
      hipLaunchKernelGGL((k_bfw_tile<4>), dim3(c64), dim3(NT), 0, s, b->d_comps, tri64 + b->tri64_off[K], (int64_t)K);
      RBG_CHECK_LAUNCH();
    }
    prof_stop(4, pi, s, b->work[K]);
  }
  hipLaunchKernelGGL(k_bfw_capture, dim3(cgrid), dim3(256), 0, s, b->d_comps, b->count);
  RBG_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_bfw_convert<false>), dim3(b->n_chunk), dim3(NT), 0, s, b->d_comps, chunks);
  RBG_CHECK_LAUNCH();
  return 0;
}

extern "C" {

int64_t rbg_batch_snap_bytes(int32_t npad) { return FwWork::bytes(npad, 4); }

int rbg_batch_create(const rbg_batch_comp* comps, int32_t count, rbg_batch** out) {
  if (!comps || count <= 0 || !out) return RBG_E_ARG;
  *out = nullptr;
  std::vector<BComp> c(count);
  for (int i = 0; i < count; i++) {
    const rbg_batch_comp& q = comps[i];
    if (!q.W || !q.snap || !q.st || q.npad < CHUNK || q.npad % CHUNK) return RBG_E_ARG;
    if (q.tile != 128 && q.tile != 64) return RBG_E_ARG;
    if (q.tri < 0 || q.tri > 2) return RBG_E_ARG;
    c[i] = BComp{q.W, q.snap, q.st, q.npad, q.tile, q.tri};
  }
  // largest first: the components running at step K are a prefix
  std::stable_sort(c.begin(), c.end(), [](const BComp& a, const BComp& b) { return a.npad > b.npad; });
  auto* b = new rbg_batch();
  b->count = count;
  b->maxnb = c[0].npad / BLK;
  std::vector<Item> chunks, panels, t128, t64, tri128, tri64;
  for (int i = 0; i < count; i++) {
    const int nc = c[i].npad / CHUNK, nt = c[i].npad / c[i].tile;
    for (int q = 0; q < nc * nc; q++) chunks.push_back({i, q});
    for (int q = 0; q < lines_groups(c[i].npad); q++) panels.push_back({i, q});
    if (c[i].tri) continue;
    auto& tl = c[i].tile == 128 ? t128 : t64;
    for (int q = 0; q < nt * nt; q++) tl.push_back({i, q});
  }
  b->n_chunk = (int64_t)chunks.size();
  b->n_panel = (int64_t)panels.size();
  b->n_t128 = (int64_t)t128.size();
  b->n_t64 = (int64_t)t64.size();
  b->panel_cnt.assign(b->maxnb, 0);
  b->t128_cnt.assign(b->maxnb, 0);
  b->t64_cnt.assign(b->maxnb, 0);
  b->tri128_off.assign(b->maxnb + 1, 0);
  b->tri64_off.assign(b->maxnb + 1, 0);
  b->work.assign(b->maxnb, 0.0);
  for (int i = 0; i < count; i++) {
    const int nb = c[i].npad / BLK, nt = c[i].npad / c[i].tile;
    for (int K = 0; K < nb; K++) {
      b->panel_cnt[K] += lines_groups(c[i].npad);
      if (!c[i].tri) {
        (c[i].tile == 128 ? b->t128_cnt : b->t64_cnt)[K] += (int64_t)nt * nt;
        b->work[K] += ((double)c[i].npad - BLK) * c[i].npad * BLK;  // every cell outside the pivot rows
      }
    }
  }
  // triangular components: per step, the tiles its pivot block can change
  // (tri_rect), of every component still running, in component order
  for (int K = 0; K < b->maxnb; K++) {
    for (int i = 0; i < count && c[i].npad / BLK > K; i++) {
      if (!c[i].tri) continue;
      const int nt = c[i].npad / c[i].tile;
      int cnt = 0;
      const TileRect r = tri_rect(c[i].tri, c[i].npad, c[i].tile, K, &cnt);
      auto& tl = c[i].tile == 128 ? tri128 : tri64;
      for (int q = 0; q < cnt; q++) tl.push_back({i, (r.i0 + q / r.jn) * nt + r.j0 + q % r.jn});
      // the rectangle's cells outside the pivot rows (the pivot tile row is in it)
      b->work[K] += ((double)cnt * c[i].tile * c[i].tile - (double)BLK * r.jn * c[i].tile) * BLK;
    }
    b->tri128_off[K + 1] = (int64_t)tri128.size();
    b->tri64_off[K + 1] = (int64_t)tri64.size();
  }
  b->n_tri128 = (int64_t)tri128.size();
  b->n_tri64 = (int64_t)tri64.size();
  std::vector<Item> all;
  all.reserve(chunks.size() + panels.size() + t128.size() + t64.size() + tri128.size() + tri64.size());
  for (auto* v : {&chunks, &panels, &t128, &t64, &tri128, &tri64}) all.insert(all.end(), v->begin(), v->end());
  hipError_t e = hipMalloc((void**)&b->d_comps, sizeof(BComp) * count);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_items, sizeof(Item) * all.size());
  if (e == hipSuccess) e = hipMemcpy(b->d_comps, c.data(), sizeof(BComp) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_items, all.data(), sizeof(Item) * all.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(b->d_comps);
    (void)hipFree(b->d_items);
    delete b;
    return (int)e;
  }
  *out = b;
  return 0;
}

int rbg_batch_info(const rbg_batch* b, int64_t* info) {
  if (!b || !info) return RBG_E_ARG;
  int64_t launches = 5;
  for (int K = 0; K < b->maxnb; K++)
    launches += 1 + (b->t128_cnt[K] > 0) + (b->t64_cnt[K] > 0) + (b->tri128_off[K + 1] > b->tri128_off[K]) +
                (b->tri64_off[K + 1] > b->tri64_off[K]);
  info[0] = b->count;
  info[1] = b->maxnb;
  info[2] = launches;
  info[3] = b->n_panel;
  info[4] = b->n_t128 + b->n_tri128;
  info[5] = b->n_t64 + b->n_tri64;
  info[6] = b->n_chunk;
  info[7] = b->exec != nullptr;
  return 0;
}

int rbg_batch_run(rbg_batch* b, rbg_stream_t stream) {
  if (!b) return RBG_E_ARG;
  return batch_launch(b, (hipStream_t)stream);
}

int rbg_batch_capture(rbg_batch* b, rbg_stream_t stream) {
  if (!b || !stream) return RBG_E_ARG;  // the legacy null stream cannot be captured
  hipStream_t s = (hipStream_t)stream;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return (int)e;
  int rc = batch_launch(b, s);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return (int)e;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return (int)e;
  }
  b->graph = g;
  return 0;
}

int rbg_batch_replay(rbg_batch* b, rbg_stream_t stream) {
  if (!b || !b->exec) return RBG_E_ARG;
  hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (int)e;
}

int rbg_batch_destroy(rbg_batch* b) {
  if (!b) return 0;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->d_comps);
  (void)hipFree(b->d_items);
  delete b;
  return 0;
}

}  // extern "C"
