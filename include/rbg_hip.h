/*
 * librbg_hip.so -- C ABI of the MI355X (gfx950) AVOS closure library.
 *
 * Conventions
 *   - Every pointer argument is a device pointer owned by the caller unless the
 *     comment says "host".  No function allocates device memory unless it says
 *     so; workspace sizes come from the *_workspace_bytes queries.
 *   - Every call is asynchronous on `stream` and performs no host
 *     synchronisation, so it may be captured into a hipGraph -- except the
 *     ones whose comment says they wait (the DAG closure's completion test,
 *     the topological sort's round test, the components' round test, the
 *     batched plan's descriptor upload).
 *   - Return value: 0 on success, an RBG_E_* code for argument errors, or a
 *     hipError_t (>0) from a failed launch.
 *   - Arithmetic errors (cycles, overflow) are reported through the device
 *     resident rbg_status block; the caller reads it after synchronising.
 *   - Values are held in a widened signed domain: int32 for element widths
 *     8/16/32 and int64 for width 64, RED_ONE = -1.  `width` is the element
 *     width of the caller's dtype and `is_unsigned` its signedness; they set
 *     the overflow rules of the reference's per-dtype C arithmetic.
 */
#ifndef RBG_HIP_H
#define RBG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *rbg_stream_t; /* hipStream_t */

enum {
  RBG_OK = 0,
  RBG_E_ARG = -1,       /* bad size / pointer / width */
  RBG_E_WORKSPACE = -2, /* workspace too small */
};

/* event bits in rbg_status.events */
enum {
  RBG_EV_OVERFLOW_WIDTH = 1, /* MSB(x)+MSB(y) >= width  (OverflowError, rbg_math.c.in:104-111) */
  RBG_EV_OVERFLOW_RED = 2,   /* result collides with RED_ONE (rbg_math.c.in:113-119)          */
  RBG_EV_INVALID = 4,        /* operand < -1 (the reference's MSB loop does not terminate)     */
  RBG_EV_SIGN = 8,           /* signed result wraps negative                                   */
};

/* bits in rbg_status.flags */
enum {
  RBG_F_NEED_GENERAL = 1, /* fast kernels met an input they do not cover: rerun general mode */
  RBG_F_PRECOND_FAIL = 2, /* precheck: input outside the fast-path domain                    */
  RBG_F_CYCLE = 4,        /* DAG paths: a cycle was found                                    */
  RBG_F_CAPACITY = 8,     /* a device-side capacity was exceeded (caller must grow it)        */
  RBG_F_HAS_LOWER = 16,   /* precheck: some off-diagonal entry lies below the diagonal       */
  RBG_F_HAS_UPPER = 32,   /* precheck: some off-diagonal entry lies above the diagonal       */
  RBG_F_STALLED = 64,     /* a bounded wait of the persistent closure ran out (set with
                             RBG_F_NEED_GENERAL): rerun, e.g. with tri = 3                  */
};

/*
 * Device-resident status block.  Keys order events as the reference's naive
 * loop meets them: key = (k * n + i) * n + j (n = the padded order).
 */
typedef struct rbg_status {
  uint32_t events;
  uint32_t flags;
  uint64_t cycle_count;
  uint64_t cycle_first_key; /* UINT64_MAX if none */
  uint64_t cycle_last_key;
  uint64_t ov_first_key;    /* UINT64_MAX if none */
  uint64_t ov_last_key;
  int64_t cycle_first_a, cycle_first_b; /* operands W[i][k], W[k][i] at the detection */
  int64_t cycle_last_a, cycle_last_b;
  int64_t ov_last_a, ov_last_b;
  int64_t diameter_max;     /* max trajectory value (RED_ONE excluded)          */
  int64_t aux[8];           /* per-entry-point extras (documented per function) */
} rbg_status;

/* ---- status -------------------------------------------------------------- */
/* Reset a status block with a kernel (no memset node, safe under capture). */
int rbg_status_init(rbg_status *st, rbg_stream_t stream);

/* ---- element-wise AVOS (gpu/avos_kernels.py:179-297) ---------------------- */
/* z[i] = avos_sum(x[i], y[i]);  replaces avos_sum_gpu (gpu/avos_kernels.py:272-283) */
int rbg_avos_sum_i32(const int32_t *x, const int32_t *y, int32_t *z, int64_t n, rbg_stream_t stream);
int rbg_avos_sum_i64(const int64_t *x, const int64_t *y, int64_t *z, int64_t n, rbg_stream_t stream);
/* z[i] = avos_product(x[i], y[i]); replaces avos_product_gpu (gpu/avos_kernels.py:286-297);
 * events (overflow) are OR-ed into st->events when st is non-null. */
int rbg_avos_product_i32(const int32_t *x, const int32_t *y, int32_t *z, int64_t n,
                         rbg_status *st, rbg_stream_t stream);
int rbg_avos_product_i64(const int64_t *x, const int64_t *y, int64_t *z, int64_t n,
                         rbg_status *st, rbg_stream_t stream);

/* Host scalar arithmetic with the C core's 64-bit rules (c_avos_sum_impl /
 * c_avos_product_impl, redblackgraphmodule.c:490-534); *events receives RBG_EV_* bits. */
int64_t rbg_host_avos_sum_i64(int64_t a, int64_t b);
uint64_t rbg_host_avos_product_u64(int64_t a, int64_t b, uint32_t *events);

/* ---- dense Floyd-Warshall (warshall.c.in:50-107) -------------------------- */
/* Dense closure in place on an n x n row-major matrix with leading dimension ld.
 * n must be a multiple of 128 (the facade pads with zero rows/columns, which the
 * semiring ignores).  mode: 0 = fast (values must pass rbg_fw_precheck_*),
 * 1 = general (any input), 2 = fast on a matrix rbg_fw_load_* wrote (already
 * in the fast kernels' encoding; st still holding the load's flags, which the
 * closure checks against tri and clears).  On return st->diameter_max holds the max trajectory
 * value, st->cycle_* the first/last cycle detections, st->ov_* overflow events,
 * and st->flags & RBG_F_NEED_GENERAL asks for a general-mode rerun on the
 * original input.  rbg_fw_workspace_bytes is the workspace rbg_fw_dense_* needs. */
int64_t rbg_fw_workspace_bytes(int64_t n, int32_t elem_bytes);
/* The workspace for rbg_fw_dense_tri_* (no reference counterpart; the
 * triangular schedule's size query): with it a lower triangular fast closure
 * runs as one persistent launch (tri = 3: the look-ahead schedule, one launch
 * per pivot step); with only rbg_fw_workspace_bytes, one plain triangular
 * launch pair per pivot step. */
int64_t rbg_fw_tri_workspace_bytes(int64_t n, int32_t elem_bytes);
/* No counterpart in the reference (test hook): the longest wait of one
 * persistent work item, in device wall-clock ticks, for the closures
 * launched after the call (single-matrix and batched); 0 restores the
 * default of two seconds.  A wait that runs out ends every workgroup of that
 * closure and flags it RBG_F_STALLED | RBG_F_NEED_GENERAL.  Returns the
 * previous override. */
int64_t rbg_fw_flow_wait_limit(int64_t ticks);
/* No counterpart in the reference (schedule introspection, host only, for
 * tests): the persistent lower triangular closure's work items of order n in
 * ticket order (elem_bytes 4 or 8: the int32 or int64 closure; edge the tile
 * edge, 64 or 128, or 0 for the single-matrix closure's choice: 128 for int32
 * above n = 2048, else 64), four int32 per item --
 * (0, K, idx, 0) for item idx of pivot block K's panel -- with c = 1 when
 * block K has column-panel rows (K < n / 32 - 1), else 0: idx = 0 < c the
 * column panel's first 64 rows (from 32 (K + 1)), c <= idx < K + c row panel
 * (K, idx - c), idx >= K + c the column panel's rows from 32 (K + 1) +
 * 64 (idx - K) -- and (1, K, I, J) for tile (I, J) of pivot step K.  Returns the item count;
 * items are written when cap >= 4 * count. */
int64_t rbg_fw_flow_schedule(int64_t n, int32_t elem_bytes, int32_t edge, int32_t *items, int64_t cap);
/* No counterpart in the reference (host only, for tests): each item's waits,
 * ten (word, target) pairs per item in ticket order -- "counter word `word`
 * of the closure's counter block >= target", word = -1 for an unused slot.
 * The block: [4 + K] finished items of pivot block K's panel, [4 + nb + K]
 * finished tiles of step K, [4 + 2 nb + I nt + J] 1 + the last step tile
 * (I, J) finished, [4 + 2 nb + nt^2 + K nb + idx] panel item idx of block K
 * finished (nb = n / 32; nt = n / tile edge, edge as for the schedule).
 * Returns the item count; written when cap >= 20 * count. */
int64_t rbg_fw_flow_waits(int64_t n, int32_t elem_bytes, int32_t edge, int64_t *waits, int64_t cap);
/* Profiling (no reference counterpart; used by bench.py): while enabled, every
 * panel and tile launch of rbg_fw_dense_* is bracketed by hipEvents on its
 * stream.  rbg_fw_prof_read synchronises those events and returns the summed
 * durations (ms) and launch counts since the last read, then clears them. */
int rbg_fw_prof_enable(int32_t on);
/* Same facility for every timed launch kind: 0 FW panel, 1 FW tile, 5 FW small tile,
 * 6 FW persistent closure, 2 DAG level, 3 SpGEMM rows, 4 batched FW.  on = 1 starts recording afresh, 2
 * resumes and 0 pauses, both keeping what was recorded (so a caller can time
 * a sample of its launches).  rbg_prof_read returns (and clears) the summed
 * duration and count of one kind. */
int rbg_prof_enable(int32_t on);
int rbg_prof_read(int32_t kind, double *total_ms, int64_t *launches);
/* rbg_prof_read, also returning the work the launches recorded: relaxations
 * executed for the FW tile kinds (1: 128 x 128 tiles, 5: 64 x 64 int32 tiles
 * of the small-matrix and triangular schedules), 0 for the other kinds. */
int rbg_prof_read_work(int32_t kind, double *total_ms, int64_t *launches, double *work);
int rbg_fw_prof_read(double *panel_ms, int64_t *panel_launches, double *tile_ms, int64_t *tile_launches);
/* No counterpart in the reference (the fast path's copy-in, one pass over the
 * input where a copy, rbg_fw_precheck_* and the fast closure's encoding took
 * three): W (npad x npad, row stride npad, 16-byte aligned, npad a multiple
 * of 4) = A (n x n, row stride lda) zero-padded, in the fast kernels'
 * encoding, with rbg_fw_precheck_*'s flags in st and the largest input value
 * in st->aux[7].  Without RBG_F_PRECOND_FAIL, close W with mode 2 and the
 * schedule the flags allow; with it, W is unspecified (copy A again and run
 * mode 1). */
int rbg_fw_load_i32(const int32_t *A, int64_t n, int64_t lda, int32_t *W, int64_t npad, rbg_status *st,
                    rbg_stream_t stream);
int rbg_fw_load_i64(const int64_t *A, int64_t n, int64_t lda, int64_t *W, int64_t npad, rbg_status *st,
                    rbg_stream_t stream);
int rbg_fw_precheck_i32(const int32_t *W, int64_t n, int64_t ld, rbg_status *st, rbg_stream_t stream);
int rbg_fw_precheck_i64(const int64_t *W, int64_t n, int64_t ld, rbg_status *st, rbg_stream_t stream);
int rbg_fw_dense_i32(int32_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                     void *workspace, int64_t workspace_bytes, rbg_status *st, rbg_stream_t stream);
int rbg_fw_dense_i64(int64_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                     void *workspace, int64_t workspace_bytes, rbg_status *st, rbg_stream_t stream);
/* The same closure with the triangular schedule of the Cython FW's
 * assume_upper_triangular loop (_shortest_path.pyx:365-377), replacing
 * _floyd_warshall_avos(dist_matrix, assume_upper_triangular) (:340-389).
 * tri = 1: every off-diagonal entry of W lies below the diagonal (a pedigree
 * numbered parents first); tri = 2: above it; tri = 0: full schedule.  Per
 * pivot block only the tiles that block can change are launched (about 1/6
 * of the full schedule's relaxations), with W, the trajectory diameter and the
 * status identical to the full schedule's.  rbg_fw_precheck_* reports the
 * side(s) holding entries (RBG_F_HAS_LOWER / RBG_F_HAS_UPPER).  Fast mode
 * only: with mode = 1 the full schedule runs whatever tri says.  Unlike the
 * reference's loop, which reads only the assumed triangle, an entry on the
 * other side is detected: the fast run stops with RBG_F_NEED_GENERAL, and the
 * caller reruns mode 1 on the original input, as for any fast-mode miss.
 * tri = 1 with a workspace of rbg_fw_tri_workspace_bytes runs every
 * pivot step in one persistent launch whose work items wait for each other
 * under a time limit; should a wait run out (never expected), st->flags gets
 * RBG_F_STALLED | RBG_F_NEED_GENERAL and the caller reruns on the original
 * input, with tri = 3 (the same schedule, one launch per pivot step) or mode 1. */
int rbg_fw_dense_tri_i32(int32_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                         int32_t tri, void *workspace, int64_t workspace_bytes, rbg_status *st,
                         rbg_stream_t stream);
int rbg_fw_dense_tri_i64(int64_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                         int32_t tri, void *workspace, int64_t workspace_bytes, rbg_status *st,
                         rbg_stream_t stream);

/* No counterpart in the reference: rbg_fw_dense_* for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The snapshot workspace lives in a
 * file-scope device array of rbg_fw_noalloc_max_bytes() bytes (8.5 MiB, held
 * by each GPU the library runs on until the process ends), so these calls
 * allocate nothing; the caller passes W and st.  Orders whose workspace
 * (rbg_fw_workspace_bytes) exceeds the array -- int32 n > 17024, int64
 * n > 8576 -- return RBG_E_WORKSPACE before anything runs.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller.  Nothing in the
 * array needs to start cleared.  Results and status are those of
 * rbg_fw_dense_* with the same arguments (full schedule). */
int64_t rbg_fw_noalloc_max_bytes(void);
int rbg_fw_dense_noalloc_i32(int32_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                             rbg_status *st, rbg_stream_t stream);
int rbg_fw_dense_noalloc_i64(int64_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                             rbg_status *st, rbg_stream_t stream);

/* ---- dense AVOS product (einsum 'ij,jk', redblack.c.src:554-700) ----------- */
/* C[M x N] = A[M x K] (x) B[K x N]; row-major with leading dimensions. */
int rbg_matmul_i32(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C, int64_t ldc,
                   int64_t M, int64_t K, int64_t N, int32_t width, int32_t is_unsigned,
                   rbg_status *st, rbg_stream_t stream);
int rbg_matmul_i64(const int64_t *A, int64_t lda, const int64_t *B, int64_t ldb, int64_t *C, int64_t ldc,
                   int64_t M, int64_t K, int64_t N, int32_t width, int32_t is_unsigned,
                   rbg_status *st, rbg_stream_t stream);
/* The same product for operands in the fast domain (the replacement of the
 * 'ij,jk' loop for closures): every value >= -1, no off-diagonal +-1, every
 * diagonal value in {-1, 0, 1}, in A and in B; rbg_matmul_precheck_* sets
 * RBG_F_PRECOND_FAIL for an operand outside it.  The terms of two
 * off-diagonal entries run in the E domain on the phase-3 tile of the dense
 * closure (128 x 128 int32 / 64 x 64 int64 tiles of C, all of K staged in
 * 32-k blocks); the two terms with a diagonal operand are added with the full
 * arithmetic.  st->flags & RBG_F_NEED_GENERAL: a term could overflow (MSB
 * guard), C is unspecified and rbg_matmul_* must run instead.  Native widths
 * only (int32 / int64).  workspace: rbg_matmul_fast_workspace_bytes(M, K, N,
 * elem_bytes) bytes (the E-encoded operands). */
int64_t rbg_matmul_fast_workspace_bytes(int64_t M, int64_t K, int64_t N, int32_t elem_bytes);
int rbg_matmul_precheck_i32(const int32_t *X, int64_t ld, int64_t rows, int64_t cols, rbg_status *st,
                            rbg_stream_t stream);
int rbg_matmul_precheck_i64(const int64_t *X, int64_t ld, int64_t rows, int64_t cols, rbg_status *st,
                            rbg_stream_t stream);
int rbg_matmul_fast_i32(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C, int64_t ldc,
                        int64_t M, int64_t K, int64_t N, void *workspace, int64_t ws_bytes, rbg_status *st,
                        rbg_stream_t stream);
int rbg_matmul_fast_i64(const int64_t *A, int64_t lda, const int64_t *B, int64_t ldb, int64_t *C, int64_t ldc,
                        int64_t M, int64_t K, int64_t N, void *workspace, int64_t ws_bytes, rbg_status *st,
                        rbg_stream_t stream);

/* ---- relational composition (relational_composition.c.in:51-267) --------- */
/* Edge composition: out = closure R with edge alpha -> beta of value rel composed in
 * (edge_relational_composition).  out must not alias R. */
int rbg_edge_compose_i32(const int32_t *R, int64_t n, int64_t alpha, int64_t beta, int32_t rel,
                         int32_t *out, int32_t width, int32_t is_unsigned, rbg_status *st,
                         rbg_stream_t stream);
int rbg_edge_compose_i64(const int64_t *R, int64_t n, int64_t alpha, int64_t beta, int64_t rel,
                         int64_t *out, int32_t width, int32_t is_unsigned, rbg_status *st,
                         rbg_stream_t stream);
/* The same composition in place (no reference counterpart: it serves repeated
 * insertions into a resident closure).  Only the rows that reach alpha (W[m][alpha]
 * != 0) and, within them, the columns of the composed row change, so the call
 * costs those rows instead of the whole matrix; the results are those of
 * rbg_edge_compose_*.  scratch: rbg_edge_compose_scratch_bytes(n) bytes of device
 * memory (the composed row and the list of affected rows).  When st->flags gets
 * RBG_F_CYCLE, W no longer holds a meaningful closure: compose into a copy where W
 * must survive that error. */
int64_t rbg_edge_compose_scratch_bytes(int64_t n);
int rbg_edge_compose_inplace_i32(int32_t *W, int64_t n, int64_t alpha, int64_t beta, int32_t rel, void *scratch,
                                 int32_t width, int32_t is_unsigned, rbg_status *st, rbg_stream_t stream);
int rbg_edge_compose_inplace_i64(int64_t *W, int64_t n, int64_t alpha, int64_t beta, int64_t rel, void *scratch,
                                 int32_t width, int32_t is_unsigned, rbg_status *st, rbg_stream_t stream);
/* Vertex composition: out is (n+1) x (n+1).  On a cycle violation st->flags |=
 * RBG_F_CYCLE and st->aux[0..2] = {idx, u_idx, v_idx} of the first violation. */
int rbg_vertex_compose_i32(const int32_t *u, const int32_t *R, const int32_t *v, int32_t color,
                           int64_t n, int32_t *out, int32_t width, int32_t is_unsigned,
                           rbg_status *st, rbg_stream_t stream);
int rbg_vertex_compose_i64(const int64_t *u, const int64_t *R, const int64_t *v, int64_t color,
                           int64_t n, int64_t *out, int32_t width, int32_t is_unsigned,
                           rbg_status *st, rbg_stream_t stream);

/* ---- sparse CSR helpers ---------------------------------------------------- */
/* Device CSR on this ABI: int64 row pointers, int32 column indices, int32/int64 values. */
int64_t rbg_csr_tmp_bytes(int32_t n);
/* Sort rows, combine duplicate columns with the AVOS sum, drop explicit zeros
 * (rows are sorted in place in idx/val); writes the canonical CSR.  With
 * out_idx == NULL only out_indptr is produced (to size the outputs).
 * Replaces the canonicalisation rb_matrix does before its SpGEMM (sparse/rbm.py:49-56). */
int rbg_csr_canonicalize_i32(const int64_t *indptr, int32_t *idx, int32_t *val, int32_t n, void *tmp,
                             int64_t *out_indptr, int32_t *out_idx, int32_t *out_val, rbg_stream_t stream);
int rbg_csr_canonicalize_i64(const int64_t *indptr, int32_t *idx, int64_t *val, int32_t n, void *tmp,
                             int64_t *out_indptr, int32_t *out_idx, int64_t *out_val, rbg_stream_t stream);
/* Canonical-form test (no reference counterpart: it lets a caller skip
 * rbg_csr_canonicalize_* for a matrix that already is canonical): st->aux[0]
 * is increased by the number of rows whose columns are not strictly
 * increasing or that hold an explicit zero. */
int rbg_csr_check_canonical_i32(const int64_t *indptr, const int32_t *idx, const int32_t *val, int32_t n,
                                rbg_status *st, rbg_stream_t stream);
int rbg_csr_check_canonical_i64(const int64_t *indptr, const int32_t *idx, const int64_t *val, int32_t n,
                                rbg_status *st, rbg_stream_t stream);
/* CSRMatrixGPU._validate_csr (gpu/csr_gpu.py:93-147): monotone indptr, columns in
 * range; a violation sets RBG_F_CAPACITY in st->flags. */
int rbg_csr_validate(const int64_t *indptr, const int32_t *idx, int32_t n, int32_t ncols, rbg_status *st,
                     rbg_stream_t stream);

/* ---- sparse DAG closure (_transitive_closure_dag.pyx:168-351,
 *      gpu/transitive_closure.py:474-652) --------------------------------------- */
/* Workspace for n vertices and a matrix of nnz entries (nnz = 0 for the
 * SpGEMM entry points, which need no predecessor lists). */
int64_t rbg_dag_workspace_bytes(int32_t n, int64_t nnz);
/* Closure of a DAG (canonical CSR: int64 row pointers, sorted rows, no explicit
 * zeros).  Rounds r = 0, 1, ...: the vertices whose non-self successors all sit
 * in levels < r form level r, and each of their closure rows is merged by one
 * wavefront from its direct entries and A[v][w] (x) C[w] for each successor w.
 * Every level reserves its candidate count in the arena (rows are compacted
 * into the front of their slot), so the arena needs the total candidate count
 * sum_v (nnz(A[v]) + sum_w nnz(C[w])), not nnz(C).  All sizes live on the
 * device; the host synchronises once every `batch` rounds to test completion
 * (a graph of L levels takes L rounds).  batch < 0 runs exactly -batch rounds
 * and returns without synchronising: with -batch = L, known from an earlier
 * closure of the same graph, the whole closure is queued at once, and
 * st->aux[7] = 0 with neither RBG_F_CYCLE nor RBG_F_CAPACITY set says that
 * the rounds did not suffice (call again with batch > 0).  Either way the
 * call returns with the report still queued on `stream`: read st after it.
 * level_kinds: NULL, or with batch < 0 one byte per round as rbg_dag_level_kinds
 * returned it after an earlier closure of the same graph.
 * Levels come from the reversed edges (a vertex joins level r + 1 when its last
 * unplaced successor is placed in level r), whose lists live in the workspace:
 * size it with rbg_dag_workspace_bytes(n, nnz(A)) (a smaller one sets
 * RBG_F_CAPACITY).  Results: closure rows in the arena (see rbg_dag_gather_*); st->aux[2] max
 * |candidate| (Cython diameter rule), [3] max |final value| > 1 (GPU diameter
 * rule), [4] arena entries used, [6] levels, [7] 1 when every vertex was
 * placed.  A cycle sets RBG_F_CYCLE with
 * st->aux[0] = a vertex on the cycle; a full arena sets RBG_F_CAPACITY (grow
 * it and call again). */
int rbg_dag_closure_i32(const int64_t *indptr, const int32_t *indices, const int32_t *data, int32_t n,
                        void *workspace, int64_t ws_bytes, int32_t *arena_cols, int32_t *arena_vals,
                        int64_t arena_cap, int32_t batch, const uint8_t *level_kinds, int32_t width,
                        rbg_status *st, rbg_stream_t stream);
int rbg_dag_closure_i64(const int64_t *indptr, const int32_t *indices, const int64_t *data, int32_t n,
                        void *workspace, int64_t ws_bytes, int32_t *arena_cols, int64_t *arena_vals,
                        int64_t arena_cap, int32_t batch, const uint8_t *level_kinds, int32_t width,
                        rbg_status *st, rbg_stream_t stream);
/* No reference counterpart: the levels' kinds of the last closure in `workspace`
 * (count <= n bytes, one per level: 1 = merged by the unit kernel, else by the
 * pair kernel; | 2 = it had long rows), copied to host_dst (pinned) on `stream`.
 * Passed back as level_kinds with batch = -levels for the same graph, they
 * launch only the merge kernels each level needs.  A level that needs one the
 * hint left out stops the rounds (st->aux[7] = 0 without RBG_F_CYCLE or
 * RBG_F_CAPACITY: call again with batch > 0); a cycle is still reported. */
int rbg_dag_level_kinds(void *workspace, int32_t n, uint8_t *host_dst, int32_t count, rbg_stream_t stream);
/* Topological sort (_topological_sort.pyx:30-163, topological_sort), by the
 * closure's Kahn levels: level_out[v] = the longest path from v to a vertex
 * without non-self successors (NULL to skip), order_out = the vertices by
 * decreasing level, so every edge v -> w (a nonzero A[v][w], w != v) has v
 * before w, as the reference's order does; permuting A by it gives an upper
 * triangular matrix.  Within a level the order is unspecified (the reference's
 * DFS order is one valid topological order; this is another).  workspace:
 * rbg_dag_workspace_bytes(n, nnz).  A cycle sets RBG_F_CYCLE with st->aux[0]
 * = a vertex on it (the outputs are then not written); st->aux[6] = levels.
 * The host waits once every 8 rounds (a graph of L levels takes L rounds). */
int rbg_topological_sort(const int64_t *indptr, const int32_t *indices, int32_t n, void *workspace,
                         int64_t ws_bytes, int32_t *level_out, int32_t *order_out, rbg_status *st,
                         rbg_stream_t stream);
/* Entry counts by side of the diagonal, for is_upper_triangular
 * (_topological_sort.pyx:222-265): st->aux[0] += entries with column < row,
 * st->aux[1] += entries on the diagonal. */
int rbg_csr_triangles(const int64_t *indptr, const int32_t *indices, int32_t n, rbg_status *st,
                      rbg_stream_t stream);
/* Gather arena rows into CSR: out_indptr (n+1, int64) always; indices/data when
 * non-null, into buffers of out_cap entries.  When nnz > out_cap nothing is
 * written to them and st->flags gets RBG_F_CAPACITY (call again with buffers of
 * st->aux[1] entries).  Rows a failed closure did not merge have length 0, so the
 * call may be queued before the closure's status is read.  st->aux: [1] nnz,
 * [2] max |candidate|, [3] max |final value| > 1, [4] arena entries used, [5] temp
 * entries used. */
int rbg_dag_gather_i32(void *workspace, int32_t n, const int32_t *arena_cols, const int32_t *arena_vals,
                       int64_t *out_indptr, int32_t *out_indices, int32_t *out_data, int64_t out_cap,
                       rbg_status *st, rbg_stream_t stream);
int rbg_dag_gather_i64(void *workspace, int32_t n, const int32_t *arena_cols, const int64_t *arena_vals,
                       int64_t *out_indptr, int32_t *out_indices, int64_t *out_data, int64_t out_cap,
                       rbg_status *st, rbg_stream_t stream);

/* ---- AVOS SpGEMM (sparse/sparsetools/rbm_math.h:116-222, gpu/spgemm*.py) ---------- */
/* Rows of C = A (x) B into the row arena (rbg_dag_workspace_bytes(m) workspace,
 * sized for m = rows of A); then rbg_dag_gather_* writes the CSR.  A and B use
 * int64 row pointers, rows sorted.  mask_upper keeps columns >= row only.
 * Zero sums are dropped (the C++ pass 2 rule). */
int rbg_spgemm_rows_i32(const int64_t *ap, const int32_t *aj, const int32_t *ax, int32_t m, const int64_t *bp,
                        const int32_t *bj, const int32_t *bx, int32_t mask_upper, void *workspace,
                        int32_t *arena_cols, int32_t *arena_vals, int64_t arena_cap, int32_t *tmp_cols,
                        int32_t *tmp_vals, int64_t tmp_cap, int32_t width, int32_t is_unsigned, rbg_status *st,
                        rbg_stream_t stream);
int rbg_spgemm_rows_i64(const int64_t *ap, const int32_t *aj, const int64_t *ax, int32_t m, const int64_t *bp,
                        const int32_t *bj, const int64_t *bx, int32_t mask_upper, void *workspace,
                        int32_t *arena_cols, int64_t *arena_vals, int64_t arena_cap, int32_t *tmp_cols,
                        int64_t *tmp_vals, int64_t tmp_cap, int32_t width, int32_t is_unsigned, rbg_status *st,
                        rbg_stream_t stream);

/* Element-wise sum of two canonical CSR matrices (sparse_avos_sum_gpu,
 * gpu/transitive_closure.py:77-156; _sparse_avos_sum, sparse/csgraph/
 * transitive_closure.py:272-357).  combine: 0 = AVOS sum, 1 = smallest |value|
 * (the squaring merge of the reference).  cat_ptr = a_ptr + b_ptr; cat_idx /
 * cat_val hold nnz(A)+nnz(B) entries of scratch.  Two calls: first with
 * out_idx == NULL to size the output (out_indptr[n] = nnz), then with outputs. */
int rbg_csr_add_i32(const int64_t *ap, const int32_t *aj, const int32_t *ax, const int64_t *bp, const int32_t *bj,
                    const int32_t *bx, int32_t n, int32_t combine, const int64_t *cat_ptr, int32_t *cat_idx,
                    int32_t *cat_val, void *tmp, int64_t *out_indptr, int32_t *out_idx, int32_t *out_val,
                    rbg_stream_t stream);
int rbg_csr_add_i64(const int64_t *ap, const int32_t *aj, const int64_t *ax, const int64_t *bp, const int32_t *bj,
                    const int64_t *bx, int32_t n, int32_t combine, const int64_t *cat_ptr, int32_t *cat_idx,
                    int64_t *cat_val, void *tmp, int64_t *out_indptr, int32_t *out_idx, int64_t *out_val,
                    rbg_stream_t stream);

/* ---- batched dense closure -------------------------------------------------
 * Replaces closing components one at a time (sparse/csgraph/
 * transitive_closure.py:85-195 component_wise_closure, each through
 * warshall.c.in:50-107).  Every component is an int32 matrix of padded order
 * npad (multiple of 128, rows contiguous, zero padding, 16-byte aligned), closed in place with
 * the fast kernels' arithmetic; `snap` is its snapshot workspace of
 * `snap_bytes` bytes (at least rbg_batch_snap_bytes(npad)), `st` its status,
 * `tile` the phase-3 tile edge (128 or 64) for its size bucket.  Lower
 * triangular components (tri = 1) whose workspace holds
 * rbg_batch_tri_snap_bytes(npad) and that ask for it (`persistent`) are closed together by one persistent launch
 * (the work items of rbg_fw_dense_tri_*'s persistent closure, every such
 * component's in one grid); the others by one panel and tile launches per
 * pivot step.  A wait of the persistent launch that runs out (never seen)
 * flags its component RBG_F_STALLED | RBG_F_NEED_GENERAL.  After a run each
 * status holds what the
 * single-matrix fast path would report; RBG_F_PRECOND_FAIL, RBG_F_NEED_GENERAL
 * or a nonzero cycle_count mean the component must be closed again from its
 * input by rbg_fw_dense_i32 in general mode (its matrix is then unspecified).
 * rbg_batch_create copies the descriptors and builds the per-step work lists
 * (device memory owned by the handle); run launches the whole schedule on
 * `stream`; capture records it into a hipGraph owned by the handle (stream
 * must not be the legacy null stream) and replay launches that graph.
 * info (10 entries): [0] count [1] pivot-block steps [2] launches [3] panel
 *       items [4] 128-tiles [5] 64-tiles (the per-step launches) [6] 128x128
 *       chunks [7] graph captured [8] components of the persistent launch
 *       [9] its work items. */
typedef struct rbg_batch rbg_batch;
typedef struct {
  int32_t *W;
  void *snap;
  rbg_status *st;
  int32_t npad;
  int32_t tile;
  int32_t tri; /* 0: full schedule; 1 / 2: every off-diagonal entry lies below / above the diagonal,
                  and the component takes the triangular schedule of rbg_fw_dense_tri_* (a component
                  that is not triangular as declared is flagged RBG_F_PRECOND_FAIL) */
  int32_t persistent; /* tri = 1 with rbg_batch_tri_snap_bytes: 1 = the persistent launch, 0 = the per-step
                         launches; per step, a triangular component (tri = 1 or 2) with that workspace has
                         every diagonal block closed before the steps */
  int64_t snap_bytes;
} rbg_batch_comp;
int64_t rbg_batch_snap_bytes(int32_t npad);
int64_t rbg_batch_tri_snap_bytes(int32_t npad);
/* No counterpart in the reference (host only, for tests): the persistent
 * launch's tickets for components of padded orders npad[] (largest first)
 * with tile edges edge[] (64 or 128): out[2 x] the component and out[2 x + 1]
 * the component's own ticket (rbg_fw_flow_schedule's order) of ticket x.
 * Returns the ticket count; written when cap >= 2 * count. */
int64_t rbg_batch_flow_tickets(const int32_t *npad, const int32_t *edge, int32_t count, int64_t *out,
                               int64_t cap);
int rbg_batch_create(const rbg_batch_comp *comps, int32_t count, rbg_batch **out);
int rbg_batch_info(const rbg_batch *b, int64_t *info);
int rbg_batch_run(rbg_batch *b, rbg_stream_t stream);
int rbg_batch_capture(rbg_batch *b, rbg_stream_t stream);
int rbg_batch_replay(rbg_batch *b, rbg_stream_t stream);
int rbg_batch_destroy(rbg_batch *b);

/* ---- connected components ----------------------------------------------------
 * Replaces sparse/csgraph/_components.pyx:86-177 (find_components_sparse) and
 * :15-81 (find_components): components of the graph with edges taken in
 * both directions.  labels[v] = the smallest vertex of v's component (the
 * reference numbers components by that vertex, in increasing order).
 * flags: two int32 of device scratch.  The host waits once per hooking round
 * (*rounds of them). */
int rbg_components(const int64_t *indptr, const int32_t *indices, int32_t n, int32_t *labels, int32_t *flags,
                   int32_t *rounds, rbg_stream_t stream);

/* ---- CSR helpers around the closures -------------------------------------------
 * Device replacements for the host/scipy steps of the reference around its
 * closures.  temp: rbg_glue_temp_bytes(n) bytes of device memory (scan and
 * sort storage, per-row counts).  Calls with a two-pass protocol write the
 * n + 1 row pointers when out_indices / out_data are NULL, and the entries
 * (into buffers of out_indptr[n] entries) when they are given. */
int64_t rbg_glue_temp_bytes(int64_t n);
/* st->aux[0] += sum over A's entries (i, k) of |row k of B|: the candidate
 * count of A (x) B, an exact bound on its size (the symbolic count of
 * gpu/spgemm_symbolic.py:124-183).  m = rows of A. */
int rbg_csr_expand_size(const int64_t *ap, const int32_t *aj, int32_t m, const int64_t *bp, rbg_status *st,
                        rbg_stream_t stream);
/* st->aux[0] = max |value| of a CSR matrix (atomic max into a cleared
 * status): the final-value diameter of gpu/transitive_closure.py:43-50. */
int rbg_csr_max_abs_i32(const int64_t *ap, const int32_t *ax, int32_t n, rbg_status *st, rbg_stream_t stream);
int rbg_csr_max_abs_i64(const int64_t *ap, const int64_t *ax, int32_t n, rbg_status *st, rbg_stream_t stream);
/* The n x ncols dense matrix of a CSR (every cell written). */
int rbg_csr_to_dense_i32(const int64_t *ap, const int32_t *aj, const int32_t *ax, int32_t n, int64_t ncols,
                         int32_t *D, rbg_stream_t stream);
int rbg_csr_to_dense_i64(const int64_t *ap, const int32_t *aj, const int64_t *ax, int32_t n, int64_t ncols,
                         int64_t *D, rbg_stream_t stream);
/* The canonical CSR of a dense n x ncols matrix (zeros dropped), two-pass. */
int rbg_dense_to_csr_i32(const int32_t *D, int32_t n, int64_t ncols, int64_t *out_indptr, int32_t *out_indices,
                         int32_t *out_data, void *temp, int64_t temp_bytes, rbg_stream_t stream);
int rbg_dense_to_csr_i64(const int64_t *D, int32_t n, int64_t ncols, int64_t *out_indptr, int32_t *out_indices,
                         int64_t *out_data, void *temp, int64_t temp_bytes, rbg_stream_t stream);
/* Column `col` of a sorted CSR as an n x 1 CSR (one call: outputs of n entries
 * at most), for the edge composition's R[:, alpha] (core/redblack.py:74-97). */
int rbg_csr_column_i32(const int64_t *ap, const int32_t *aj, const int32_t *ax, int32_t n, int32_t col,
                       int64_t *out_indptr, int32_t *out_indices, int32_t *out_data, void *temp, int64_t temp_bytes,
                       rbg_stream_t stream);
int rbg_csr_column_i64(const int64_t *ap, const int32_t *aj, const int64_t *ax, int32_t n, int32_t col,
                       int64_t *out_indptr, int32_t *out_indices, int64_t *out_data, void *temp, int64_t temp_bytes,
                       rbg_stream_t stream);
/* Row `row` of a sorted CSR with the entry in column `col` set to val
 * (replaced, or inserted in column order): the composed row's seed
 * R[alpha] with R[alpha][beta] = relationship (core/redblack.py:82-85).
 * Outputs hold |row| + 1 entries; st->aux[0] = the new length. */
int rbg_csr_row_set_i32(const int64_t *ap, const int32_t *aj, const int32_t *ax, int64_t row, int32_t col,
                        int32_t val, int32_t *out_indices, int32_t *out_data, rbg_status *st, rbg_stream_t stream);
int rbg_csr_row_set_i64(const int64_t *ap, const int32_t *aj, const int64_t *ax, int64_t row, int32_t col,
                        int64_t val, int32_t *out_indices, int64_t *out_data, rbg_status *st, rbg_stream_t stream);
/* The entries on or above the diagonal of a sorted CSR (the triangular filter
 * of permute_sparse, _permutation.pyx:138-157), two-pass. */
int rbg_csr_upper_i32(const int64_t *ap, const int32_t *aj, const int32_t *ax, int32_t n, int64_t *out_indptr,
                      int32_t *out_indices, int32_t *out_data, void *temp, int64_t temp_bytes, rbg_stream_t stream);
int rbg_csr_upper_i64(const int64_t *ap, const int32_t *aj, const int64_t *ax, int32_t n, int64_t *out_indptr,
                      int32_t *out_indices, int64_t *out_data, void *temp, int64_t temp_bytes, rbg_stream_t stream);
/* Component layout from rbg_components' labels (each vertex's smallest
 * component vertex): comp[v] = its component id (components numbered by
 * their smallest vertex, _components.pyx:86-177), local[v] = its index
 * within the component (vertices keep their order), members = the vertices
 * grouped by component, sizes[c] (n entries, zero past the count), starts[c]
 * (n + 1 entries: the exclusive scan of sizes), *ncomp (device int64), and,
 * when vertex_size is not NULL, vertex_size[v] = sizes[comp[v]]. */
int rbg_components_layout(const int32_t *labels, int32_t n, int32_t *comp, int32_t *local, int32_t *members,
                          int32_t *sizes, int64_t *starts, int64_t *ncomp, int32_t *vertex_size, void *temp,
                          int64_t temp_bytes, rbg_stream_t stream);
/* extract_submatrix for every component at once (_components.pyx:180-251):
 * each entry (v, w) of the graph into cell (local[v], local[w]) of component
 * comp[v]'s padded int32 matrix at pool + comp_offset[c] (order comp_npad[c],
 * row-major).  The pool must be cleared first. */
int rbg_components_scatter(const int64_t *indptr, const int32_t *indices, const int32_t *data, int32_t n,
                           const int32_t *comp, const int32_t *local, const int64_t *comp_offset,
                           const int32_t *comp_npad, int32_t *pool, rbg_stream_t stream);
/* merge_component_matrices (_components.pyx:254-323): the closed components
 * back into one n x n CSR in global vertex order (rows sorted), two-pass. */
int rbg_components_merge(const int32_t *pool, const int64_t *comp_offset, const int32_t *comp_npad,
                         const int32_t *comp, const int32_t *local, const int32_t *members, const int64_t *starts,
                         const int32_t *sizes, int32_t n, int64_t *out_indptr, int32_t *out_indices,
                         int32_t *out_data, void *temp, int64_t temp_bytes, rbg_stream_t stream);

/* ---- canonical ordering and permutation -----------------------------------
 * Replaces sparse/csgraph/_ordering.pyx:17-157 (_get_permutation,
 * _get_permutation_sparse, _compute_ordering_metrics_sparse) and
 * _permutation.pyx:38-136 (permute, permute_sparse).
 * Metrics, one int64 per vertex: anc = sum of MSB(v) over the row's nonzero
 * entries (MSB(v) = floor(log2 v) for v > 1, else 0), maxrel = the column's
 * largest entry and at least 0, diag = the vertex's own entry (0 when a CSR
 * row holds none). */
int rbg_order_metrics_dense_i32(const int32_t *A, int64_t n, int64_t *anc, int64_t *maxrel, int64_t *diag,
                                rbg_stream_t stream);
int rbg_order_metrics_dense_i64(const int64_t *A, int64_t n, int64_t *anc, int64_t *maxrel, int64_t *diag,
                                rbg_stream_t stream);
int rbg_order_metrics_csr_i32(const int64_t *indptr, const int32_t *indices, const int32_t *data, int32_t n,
                              int64_t *anc, int64_t *maxrel, int64_t *diag, rbg_stream_t stream);
int rbg_order_metrics_csr_i64(const int64_t *indptr, const int32_t *indices, const int64_t *data, int32_t n,
                              int64_t *anc, int64_t *maxrel, int64_t *diag, rbg_stream_t stream);
/* perm = the vertices sorted on (-comp_size, comp, -anc, maxrel, -diag, vertex)
 * (comp: component id per vertex, comp_size: its component's vertex count);
 * temp: rbg_order_sort_temp_bytes(n) bytes of device memory. */
int64_t rbg_order_sort_temp_bytes(int32_t n);
int rbg_order_sort(int32_t n, const int32_t *comp, const int32_t *comp_size, const int64_t *anc,
                   const int64_t *maxrel, const int64_t *diag, int32_t *perm, void *temp, int64_t temp_bytes,
                   rbg_stream_t stream);
/* B[i][j] = A[p[i]][p[j]] (0 for j < i when upper != 0); B must not alias A. */
int rbg_permute_dense_i32(const int32_t *A, int64_t n, const int32_t *p, int32_t *B, int32_t upper,
                          rbg_stream_t stream);
int rbg_permute_dense_i64(const int64_t *A, int64_t n, const int32_t *p, int64_t *B, int32_t upper,
                          rbg_stream_t stream);
/* P A P^-1 of a CSR matrix (any nnz: 64-bit row pointers and offsets): row i
 * is row p[i] with every column c renamed to pinv[c], sorted (equal columns
 * keep their order).  Outputs: n + 1 row pointers and nnz entries;
 * workspace: rbg_permute_csr_workspace_bytes_*(n, nnz) bytes. */
int64_t rbg_permute_csr_workspace_bytes_i32(int32_t n, int64_t nnz);
int64_t rbg_permute_csr_workspace_bytes_i64(int32_t n, int64_t nnz);
int rbg_permute_csr_i32(const int64_t *indptr, const int32_t *indices, const int32_t *data, int32_t n, int64_t nnz,
                        const int32_t *p, int64_t *out_indptr, int32_t *out_indices, int32_t *out_data,
                        void *workspace, int64_t ws_bytes, rbg_stream_t stream);
int rbg_permute_csr_i64(const int64_t *indptr, const int32_t *indices, const int64_t *data, int32_t n, int64_t nnz,
                        const int32_t *p, int64_t *out_indptr, int32_t *out_indices, int64_t *out_data,
                        void *workspace, int64_t ws_bytes, rbg_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* RBG_HIP_H */
