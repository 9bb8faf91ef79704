/*
 * librbg_hip.so -- C ABI of the MI355X (gfx950) AVOS closure library.
 *
 * Conventions
 *   - Every pointer argument is a device pointer owned by the caller unless the
 *     comment says "host".  No function allocates device memory unless it says
 *     so; workspace sizes come from the *_workspace_bytes queries.
 *   - Every call is asynchronous on `stream` and performs no host
 *     synchronisation, so it may be captured into a hipGraph.
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
 * 1 = general (any input).  On return st->diameter_max holds the max trajectory
 * value, st->cycle_* the first/last cycle detections, st->ov_* overflow events,
 * and st->flags & RBG_F_NEED_GENERAL asks for a general-mode rerun on the
 * original input. */
int64_t rbg_fw_workspace_bytes(int64_t n, int32_t elem_bytes);
/* Profiling (no reference counterpart; used by bench.py): while enabled, every
 * panel and tile launch of rbg_fw_dense_* is bracketed by hipEvents on its
 * stream.  rbg_fw_prof_read synchronises those events and returns the summed
 * durations (ms) and launch counts since the last read, then clears them. */
int rbg_fw_prof_enable(int32_t on);
int rbg_fw_prof_read(double *panel_ms, int64_t *panel_launches, double *tile_ms, int64_t *tile_launches);
int rbg_fw_precheck_i32(const int32_t *W, int64_t n, int64_t ld, rbg_status *st, rbg_stream_t stream);
int rbg_fw_precheck_i64(const int64_t *W, int64_t n, int64_t ld, rbg_status *st, rbg_stream_t stream);
int rbg_fw_dense_i32(int32_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                     void *workspace, int64_t workspace_bytes, rbg_status *st, rbg_stream_t stream);
int rbg_fw_dense_i64(int64_t *W, int64_t n, int64_t ld, int32_t width, int32_t is_unsigned, int32_t mode,
                     void *workspace, int64_t workspace_bytes, rbg_status *st, rbg_stream_t stream);

/* ---- dense AVOS product (einsum 'ij,jk', redblack.c.src:554-700) ----------- */
/* C[M x N] = A[M x K] (x) B[K x N]; row-major with leading dimensions. */
int rbg_matmul_i32(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C, int64_t ldc,
                   int64_t M, int64_t K, int64_t N, int32_t width, int32_t is_unsigned,
                   rbg_status *st, rbg_stream_t stream);
int rbg_matmul_i64(const int64_t *A, int64_t lda, const int64_t *B, int64_t ldb, int64_t *C, int64_t ldc,
                   int64_t M, int64_t K, int64_t N, int32_t width, int32_t is_unsigned,
                   rbg_status *st, rbg_stream_t stream);

/* ---- relational composition (relational_composition.c.in:51-267) --------- */
/* Edge composition: out = closure R with edge alpha -> beta of value rel composed in
 * (edge_relational_composition).  out must not alias R. */
int rbg_edge_compose_i32(const int32_t *R, int64_t n, int64_t alpha, int64_t beta, int32_t rel,
                         int32_t *out, int32_t width, int32_t is_unsigned, rbg_status *st,
                         rbg_stream_t stream);
int rbg_edge_compose_i64(const int64_t *R, int64_t n, int64_t alpha, int64_t beta, int64_t rel,
                         int64_t *out, int32_t width, int32_t is_unsigned, rbg_status *st,
                         rbg_stream_t stream);
/* Vertex composition: out is (n+1) x (n+1).  On a cycle violation st->flags |=
 * RBG_F_CYCLE and st->aux[0..2] = {idx, u_idx, v_idx} of the first violation. */
int rbg_vertex_compose_i32(const int32_t *u, const int32_t *R, const int32_t *v, int32_t color,
                           int64_t n, int32_t *out, int32_t width, int32_t is_unsigned,
                           rbg_status *st, rbg_stream_t stream);
int rbg_vertex_compose_i64(const int64_t *u, const int64_t *R, const int64_t *v, int64_t color,
                           int64_t n, int64_t *out, int32_t width, int32_t is_unsigned,
                           rbg_status *st, rbg_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* RBG_HIP_H */
