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
// Lower triangular components whose workspace holds the persistent closure's
// (rbg_batch_tri_snap_bytes) instead take one persistent launch for all of
// them: the single-matrix persistent closure's work items (k_fw_flow), every
// component's in one grid, each waiting only on its own component's counters.
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
  int32_t tri;   // 0 full schedule; 1 / 2: entries declared below / above the diagonal only
  int32_t flow;  // 1: closed by the persistent launch (snapshot ring of four)
  // triangular, per-step launches: every diagonal block closed before the
  // steps (k_bfw_prewalk), its step snapshots here for the panels
  void* dsnap;
};

// A component of the persistent launch: its matrix, snapshot workspace (with
// the diagonal blocks' step snapshots at dsnap), status and padded order; its
// counters and schedule follow from its class.
struct BFlow {
  int32_t* W;
  void* snap;
  void* dsnap;
  rbg_status* st;
  int64_t n;
};

// The persistent launch's components of one padded order and tile edge, in
// launch order (largest order first).  Tickets run class by class and, within
// a class, by the components' own tickets with the components in lockstep:
// ticket base + t cnt + m is own ticket t of component first + m.  Every wait
// of an item names an item of its own component with a smaller own ticket,
// so a smaller ticket of the launch, as in the single-matrix launch.  The
// class's counters: `words` per component from word `ctl` of the launch's
// counter array; its packed items (flow_pack, in own-ticket order) from entry
// `items` of the launch's item table.
struct BClass {
  int64_t base;
  int64_t ctl;
  int64_t items;
  int32_t first;
  int32_t cnt;
  int32_t n;
  int32_t edge;
  int32_t words;
  int32_t pad;
};
constexpr int BFLOW_MAX_CLASSES = 64;
// Passed by value: the decode reads kernel arguments only.
struct BClasses {
  int32_t count = 0;
  int32_t dual = 0;  // some class has 64-tiles
  int64_t total = 0;
  int64_t ctl_words = 0;
  int64_t item_words = 0;
  BClass c[BFLOW_MAX_CLASSES];
};

// Classes of components with these (padded order, tile edge), in launch
// order; false when they need more than BFLOW_MAX_CLASSES classes, more than
// 2^31 - 1 tickets (one workgroup per ticket) or an order past the packed
// items' range.
static bool bclass_plan(const std::vector<std::pair<int, int>>& ne, BClasses& k) {
  k = BClasses{};
  for (size_t j = 0; j < ne.size();) {
    size_t e = j;
    while (e < ne.size() && ne[e] == ne[j]) e++;
    if (k.count == BFLOW_MAX_CLASSES || ne[j].first / BLK > FLOW_PACK_MAX_NB) return false;
    const FlowCtl f = flow_ctl(ne[j].first, 4, nullptr, ne[j].second);
    BClass& c = k.c[k.count++];
    c = BClass{k.total, k.ctl_words, k.item_words, (int32_t)j, (int32_t)(e - j), ne[j].first, ne[j].second,
               (int32_t)FlowCtl::words(f.nb, f.nt), 0};
    k.total += f.total * (int64_t)(e - j);
    k.ctl_words += (int64_t)c.words * (e - j);
    k.item_words += f.total;
    k.dual |= c.edge == 64;
    j = e;
  }
  return k.total <= 0x7fffffffLL;
}

// Ticket x -> class j, component, own ticket.
__host__ __device__ inline void bclass_locate(const BClasses& k, int64_t x, int& j, int& comp, int64_t& local) {
  int lo = 0, hi = k.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (k.c[mid].base <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  j = lo;
  const uint32_t d = (uint32_t)(x - k.c[lo].base), cnt = (uint32_t)k.c[lo].cnt;
  comp = k.c[lo].first + (int)(d % cnt);
  local = (int64_t)(d / cnt);
}

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

// Domain conversion of one 128 x 128 chunk, four cells per thread and pass.
// Entering the E domain it also checks the fast domain (rules of
// k_fw_precheck, and a component declared triangular must be: its schedule
// skips the rest), and a triangular component's initial values enter its
// trajectory maximum (the triangular schedule never visits the tiles its
// pivot blocks cannot change; see k_e_convert).
template <bool ENC>
static __global__ __launch_bounds__(NT) void k_bfw_convert(const BComp* comps, const Item* items) {
  const Item it = items[blockIdx.x];
  const BComp c = comps[it.comp];
  const int nc = c.npad / CHUNK, ci = it.bid / nc, cj = it.bid % nc;
  const int64_t i0 = (int64_t)ci * CHUNK, j0 = (int64_t)cj * CHUNK;
  // a chunk wholly on a triangular component's absent side is never read or
  // written by its closure: checked (it must be all zero), never converted
  const bool off = (c.tri == 1 && cj > ci) || (c.tri == 2 && cj < ci);
  if (!ENC && off) return;
  int64_t vmax = 0;
  bool bad = false;
  for (int e = 4 * threadIdx.x; e < CHUNK * CHUNK; e += 4 * NT) {
    const int64_t i = i0 + e / CHUNK, jb = j0 + e % CHUNK;
    uint4* p = (uint4*)(c.W + i * c.npad + jb);
    const uint4 q = *p;
    int32_t v[4] = {(int32_t)q.x, (int32_t)q.y, (int32_t)q.z, (int32_t)q.w};
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const int64_t j = jb + x;
      if (ENC) {
        if (v[x] < -1) bad = true;
        if (i != j && (v[x] == 1 || v[x] == -1)) bad = true;
        if (i == j && v[x] > 1) bad = true;
        if (v[x] != 0 && ((c.tri == 1 && i < j) || (c.tri == 2 && i > j))) bad = true;
        vmax = v[x] > vmax ? v[x] : vmax;
      }
      v[x] = ENC ? e_enc<int32_t>(v[x]) : e_dec<int32_t>(v[x]);
    }
    if (!off) *p = make_uint4((uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3]);
  }
  if (ENC) {
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&c.st->flags, (uint32_t)RBG_F_PRECOND_FAIL);
    if (c.tri) {
      vmax = wave_max_i64(vmax);
      if ((threadIdx.x & 63) == 0) record_diameter(c.st, vmax);
    }
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
  fw_panel_lines_body<int32_t>(L, c.W, n, n, K, ws, 32, c.st, it.bid, lr,
                               c.dsnap ? (const int32_t*)c.dsnap + K * 2 * BLK * BLK : nullptr);
}

// The per-step triangular components' diagonal blocks, closed before the
// steps (no step changes block K before step K in such a matrix): item =
// (component, block).
static __global__ __launch_bounds__(PANEL_NT) void k_bfw_prewalk(const BComp* comps, const Item* items) {
  const Item it = items[blockIdx.x];
  const BComp c = comps[it.comp];
  fw_diag_walk_body<int32_t>(c.W, c.npad, c.npad, (int32_t*)c.dsnap, 32, c.st, it.bid);
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
  FwWork ws{b.snap, b.npad, 4, b.flow ? 4 : 2};
  const int par = ws.slot(nb - 1);
  capture_block<int32_t, true>(b.st, ws.col<int32_t>(par), ws.row<int32_t>(par), (nb - 1) * BLK, b.npad);
}

// The persistent launch's pre-passes: item = (component, diagonal block),
// every diagonal block closed with its step snapshots (k_fw_diag_walk); and
// the counters cleared.
static __global__ __launch_bounds__(PANEL_NT) void k_bfw_diag_walk(const BFlow* comps, const Item* items) {
  const Item it = items[blockIdx.x];
  const BFlow c = comps[it.comp];
  fw_diag_walk_body<int32_t>(c.W, c.n, c.n, (int32_t*)c.dsnap, 32, c.st, it.bid);
}

static __global__ void k_bfw_flow_zero(int* ctl, int64_t words) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    ctl[i] = 0;
}

// Every work item of every persistent component, one per workgroup, in the
// ticket order of BClass; workgroup x takes ticket x.  The class, component
// and counters follow from the kernel arguments alone and the item from one
// table load; the component's pointers load meanwhile.
// DUAL: some class has 64-tiles (else every class has 128-tiles).
template <bool DUAL>
static __global__ __launch_bounds__(NT, 3) void k_bfw_flow(const BFlow* comps, int* ctl, const uint32_t* items,
                                                          BClasses k, int64_t limit) {
  __shared__ union {
    FlowShared<uint32_t, 8, false> e128;
    FlowShared<uint32_t, 4, false> e64;
  } L;
  int j = 0, comp = 0;
  int64_t local = 0;
  bclass_locate(k, blockIdx.x, j, comp, local);
  const int n = k.c[j].n, edge = k.c[j].edge;
  const FlowCtl f{ctl + k.c[j].ctl + (int64_t)(comp - k.c[j].first) * k.c[j].words, n / BLK, n / edge, edge, 0};
  const FlowItem q = flow_unpack(items[k.c[j].items + local]);
  const BFlow c = comps[comp];
  FwWork ws{c.snap, n, 4, 4};
  ws.dsnap = c.dsnap;
  if (!DUAL || edge == 128)
    fw_flow_item<int32_t, 8, false>(L.e128, q, c.W, n, n, ws, 32, c.st, f, limit);
  else
    fw_flow_item<int32_t, 4, false>(L.e64, q, c.W, n, n, ws, 32, c.st, f, limit);
}

}  // namespace rbg

using namespace rbg;

struct rbg_batch {
  int32_t count = 0;
  int32_t maxnb = 0;
  int32_t stepnb = 0;  // pivot steps of the per-step launches (components outside the persistent launch)
  BComp* d_comps = nullptr;
  // the persistent launch: its components, classes, counters, diagonal-walk items
  int32_t n_flow = 0;
  int64_t n_walk = 0;
  double flow_work = 0;  // relaxations it executes
  BClasses classes;
  BFlow* d_flow = nullptr;
  Item* d_walk = nullptr;
  int* d_ctl = nullptr;
  uint32_t* d_fitems = nullptr;
  // chunks | panels | tiles 128 | tiles 64 (full-schedule components) |
  // triangular components' tiles 128, step by step | the same for tiles 64
  Item* d_items = nullptr;
  int64_t n_chunk = 0, n_panel = 0, n_t128 = 0, n_t64 = 0, n_tri128 = 0, n_tri64 = 0;
  Item* d_prewalk = nullptr;  // (component, block) of the per-step components with dsnap
  int64_t n_prewalk = 0;
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
  hipLaunchKernelGGL((k_bfw_convert<true>), dim3(b->n_chunk), dim3(NT), 0, s, b->d_comps, chunks);
  RBG_CHECK_LAUNCH();
  if (b->n_flow) {
    hipLaunchKernelGGL(k_bfw_diag_walk, dim3(b->n_walk), dim3(PANEL_NT), 0, s, b->d_flow, b->d_walk);
    RBG_CHECK_LAUNCH();
    const int64_t words = b->classes.ctl_words;
    const int zg = (int)((words + NT - 1) / NT < 2048 ? (words + NT - 1) / NT : 2048);
    hipLaunchKernelGGL(k_bfw_flow_zero, dim3(zg), dim3(NT), 0, s, b->d_ctl, words);
    RBG_CHECK_LAUNCH();
    const int64_t limit = flow_limit_ticks();
    const dim3 grid((unsigned)b->classes.total);
    int pf = prof_start(s);
    if (b->classes.dual)
      hipLaunchKernelGGL((k_bfw_flow<true>), grid, dim3(NT), 0, s, b->d_flow, b->d_ctl, b->d_fitems, b->classes,
                         limit);
    else
      hipLaunchKernelGGL((k_bfw_flow<false>), grid, dim3(NT), 0, s, b->d_flow, b->d_ctl, b->d_fitems, b->classes,
                         limit);
    RBG_CHECK_LAUNCH();
    prof_stop(4, pf, s, b->flow_work);
  }
  if (b->n_prewalk) {
    hipLaunchKernelGGL(k_bfw_prewalk, dim3((unsigned)b->n_prewalk), dim3(PANEL_NT), 0, s, b->d_comps, b->d_prewalk);
    RBG_CHECK_LAUNCH();
  }
  for (int K = 0; K < b->stepnb; K++) {
    if (b->panel_cnt[K]) {
      hipLaunchKernelGGL(k_bfw_panel, dim3(b->panel_cnt[K]), dim3(NT), 0, s, b->d_comps, panels, (int64_t)K);
      RBG_CHECK_LAUNCH();
    }
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

// the persistent launch's workspace without the counters (the launch keeps
// every component's in one array of its own)
int64_t rbg_batch_tri_snap_bytes(int32_t npad) { return flow_ctl_offset(npad, 4); }

int64_t rbg_batch_flow_tickets(const int32_t* npad, const int32_t* edge, int32_t count, int64_t* out, int64_t cap) {
  if (!npad || !edge || count <= 0) return RBG_E_ARG;
  std::vector<std::pair<int, int>> ne;
  for (int i = 0; i < count; i++) {
    if (npad[i] < CHUNK || npad[i] % CHUNK || (edge[i] != 64 && edge[i] != 128)) return RBG_E_ARG;
    if (i && npad[i] > npad[i - 1]) return RBG_E_ARG;  // launch order: largest first
    ne.push_back({npad[i], edge[i]});
  }
  BClasses k;
  if (!bclass_plan(ne, k)) return RBG_E_ARG;
  if (out && cap >= 2 * k.total) {
    for (int64_t x = 0; x < k.total; x++) {
      int j = 0, comp = 0;
      int64_t local = 0;
      bclass_locate(k, x, j, comp, local);
      out[2 * x] = comp;
      out[2 * x + 1] = local;
    }
  }
  return k.total;
}

int rbg_batch_create(const rbg_batch_comp* comps, int32_t count, rbg_batch** out) {
  if (!comps || count <= 0 || !out) return RBG_E_ARG;
  *out = nullptr;
  std::vector<BComp> c(count);
  for (int i = 0; i < count; i++) {
    const rbg_batch_comp& q = comps[i];
    if (!q.W || !q.snap || !q.st || q.npad < CHUNK || q.npad % CHUNK || ((uintptr_t)q.W & 15)) return RBG_E_ARG;
    if (q.tile != 128 && q.tile != 64) return RBG_E_ARG;
    if (q.tri < 0 || q.tri > 2) return RBG_E_ARG;
    if (q.snap_bytes < FwWork::bytes(q.npad, 4)) return RBG_E_WORKSPACE;
    // a triangular component's diagonal block K is final before step K (a
    // product into it needs an operand on the absent side), so its blocks
    // can be closed before the steps; the persistent launch takes lower ones
    const bool room = q.tri != 0 && q.snap_bytes >= flow_ctl_offset(q.npad, 4);
    const int flow = room && q.tri == 1 && q.persistent;
    c[i] = BComp{q.W, q.snap, q.st, q.npad, q.tile, q.tri, flow,
                 room && !flow ? (char*)q.snap + dsnap_offset(q.npad, 4) : nullptr};
  }
  // largest first: the components running at step K are a prefix
  std::stable_sort(c.begin(), c.end(), [](const BComp& a, const BComp& b) { return a.npad > b.npad; });
  // the persistent launch: lower triangular components with its workspace,
  // in classes of equal order and tile edge (else they take the per-step
  // launches)
  BClasses classes;
  {
    std::vector<std::pair<int, int>> ne;
    for (int i = 0; i < count; i++)
      if (c[i].flow) ne.push_back({c[i].npad, c[i].tile});
    if (!bclass_plan(ne, classes)) {
      for (auto& x : c) x.flow = 0;
      classes = BClasses{};
    }
  }
  auto* b = new rbg_batch();
  b->count = count;
  b->maxnb = c[0].npad / BLK;
  b->classes = classes;
  std::vector<BFlow> fl;
  std::vector<Item> walk;
  for (int i = 0; i < count; i++) {
    if (!c[i].flow) {
      b->stepnb = std::max(b->stepnb, c[i].npad / BLK);
      continue;
    }
    for (int K = 0; K < c[i].npad / BLK; K++) walk.push_back({(int)fl.size(), K});
    fl.push_back(BFlow{c[i].W, c[i].snap, (char*)c[i].snap + dsnap_offset(c[i].npad, 4), c[i].st, c[i].npad});
    b->flow_work += flow_relaxations(flow_ctl(c[i].npad, 4, nullptr, c[i].tile), c[i].npad);
  }
  b->n_flow = (int32_t)fl.size();
  b->n_walk = (int64_t)walk.size();
  std::vector<uint32_t> fitems;
  for (int j = 0; j < classes.count; j++) {
    const FlowCtl f = flow_ctl(classes.c[j].n, 4, nullptr, classes.c[j].edge);
    for (int64_t x = 0; x < f.total; x++) fitems.push_back(flow_pack(flow_item(f, x)));
  }
  std::vector<Item> chunks, panels, t128, t64, tri128, tri64;
  for (int i = 0; i < count; i++) {
    const int nc = c[i].npad / CHUNK, nt = c[i].npad / c[i].tile;
    for (int q = 0; q < nc * nc; q++) chunks.push_back({i, q});
    if (c[i].flow) continue;
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
    if (c[i].flow) continue;
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
      if (!c[i].tri || c[i].flow) continue;
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
  std::vector<Item> prewalk;
  for (int i = 0; i < count; i++)
    if (c[i].dsnap)
      for (int q = 0; q < c[i].npad / BLK; q++) prewalk.push_back({i, q});
  b->n_prewalk = (int64_t)prewalk.size();
  hipError_t e = hipMalloc((void**)&b->d_comps, sizeof(BComp) * count);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_items, sizeof(Item) * all.size());
  if (e == hipSuccess) e = hipMemcpy(b->d_comps, c.data(), sizeof(BComp) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_items, all.data(), sizeof(Item) * all.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && b->n_prewalk) {
    e = hipMalloc((void**)&b->d_prewalk, sizeof(Item) * prewalk.size());
    if (e == hipSuccess)
      e = hipMemcpy(b->d_prewalk, prewalk.data(), sizeof(Item) * prewalk.size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && b->n_flow) {
    e = hipMalloc((void**)&b->d_flow, sizeof(BFlow) * fl.size());
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_walk, sizeof(Item) * walk.size());
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_ctl, sizeof(int) * classes.ctl_words);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_fitems, sizeof(uint32_t) * fitems.size());
    if (e == hipSuccess)
      e = hipMemcpy(b->d_fitems, fitems.data(), sizeof(uint32_t) * fitems.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_flow, fl.data(), sizeof(BFlow) * fl.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_walk, walk.data(), sizeof(Item) * walk.size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    (void)hipFree(b->d_comps);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_prewalk);
    (void)hipFree(b->d_flow);
    (void)hipFree(b->d_walk);
    (void)hipFree(b->d_ctl);
    (void)hipFree(b->d_fitems);
    delete b;
    return (int)e;
  }
  *out = b;
  return 0;
}

int rbg_batch_info(const rbg_batch* b, int64_t* info) {
  if (!b || !info) return RBG_E_ARG;
  int64_t launches = 4 + (b->n_flow ? 3 : 0) + (b->n_prewalk ? 1 : 0);
  for (int K = 0; K < b->stepnb; K++)
    launches += (b->panel_cnt[K] > 0) + (b->t128_cnt[K] > 0) + (b->t64_cnt[K] > 0) +
                (b->tri128_off[K + 1] > b->tri128_off[K]) + (b->tri64_off[K + 1] > b->tri64_off[K]);
  info[0] = b->count;
  info[1] = b->maxnb;
  info[2] = launches;
  info[3] = b->n_panel;
  info[4] = b->n_t128 + b->n_tri128;
  info[5] = b->n_t64 + b->n_tri64;
  info[6] = b->n_chunk;
  info[7] = b->exec != nullptr;
  info[8] = b->n_flow;
  info[9] = b->classes.total;
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
  (void)hipFree(b->d_prewalk);
  (void)hipFree(b->d_flow);
  (void)hipFree(b->d_walk);
  (void)hipFree(b->d_ctl);
  (void)hipFree(b->d_fitems);
  delete b;
  return 0;
}

}  // extern "C"
