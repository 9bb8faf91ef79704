/*
 * CPU ORACLE for the AVOS closure path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C and in the reference's own loop order, the
 * algorithms the HIP library re-implements.  It exists to check the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and never as the thing measured or shipped.
 *
 * Each function cites the reference file:line it restates.  Pinning (what
 * checks this restatement independently of itself) is listed in the header of
 * oracle/oracle.py next to the Python bindings.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* event flags (oracle-local numbering) */
#define OE_WIDTH 1   /* MSB(x)+MSB(y) >= bits  : rbg_math.c.in:104-111 */
#define OE_RED 2     /* result == all ones     : rbg_math.c.in:113-119 */
#define OE_HANG 4    /* signed operand < -1 reaches MSB(): the reference loops forever (rbg_math.c.in:42) */

/* ------------------------------------------------------------------------ */
/* Scalar AVOS arithmetic, one copy per element type (rbg_math.c.in:29-121)  */
/* ------------------------------------------------------------------------ */
#define DEFINE_AVOS(NAME, T, U, SIGNED)                                              \
  T orc_##NAME##_avos_sum(T a, T b) {                                                \
    if (a == 0 || (U)(~b) == 0) return b;                                            \
    if (b == 0 || (U)(~a) == 0) return a;                                            \
    if (((U)a) < ((U)b)) return a;                                                   \
    return b;                                                                        \
  }                                                                                  \
  /* MSB of the value as the signed/unsigned element type sees it */                \
  static int NAME##_msb(T x, int *ev) {                                              \
    if ((U)~x == 0) return 0;                                                        \
    if (SIGNED && x < 0) { *ev |= OE_HANG; return 0; }                               \
    int level = 0;                                                                   \
    while (x >>= 1) level++;                                                         \
    return level;                                                                    \
  }                                                                                  \
  U orc_##NAME##_avos_product(T lhs, T rhs, int *ev) {                               \
    U x = (U)lhs;                                                                    \
    U y = (U)rhs;                                                                    \
    if (x == 0 || y == 0) return 0;                                                  \
    if (lhs == (T)-1 && rhs == (T)-1) return x;                                      \
    if (lhs == 1 && rhs == 1) return 1;                                              \
    if (lhs == (T)-1 && rhs == 1) return 0;                                          \
    if (lhs == 1 && rhs == (T)-1) return 0;                                          \
    if (lhs == (T)-1) x = 1;                                                         \
    if (rhs == (T)-1) return (lhs & 1) ? 0 : x;                                      \
    if (rhs == 1) return (lhs & 1) ? x : 0;                                          \
    int bit_position = NAME##_msb((T)y, ev);                                         \
    int result_size = NAME##_msb((T)x, ev) + bit_position;                           \
    if (result_size >= (int)(sizeof(x) * 8)) *ev |= OE_WIDTH;                        \
    U result = ((y & (((U)1 << bit_position) - 1)) | (x << bit_position));           \
    if (result == (U)-1) *ev |= OE_RED;                                              \
    return result;                                                                   \
  }

DEFINE_AVOS(byte, int8_t, uint8_t, 1)
DEFINE_AVOS(ubyte, uint8_t, uint8_t, 0)
DEFINE_AVOS(short, int16_t, uint16_t, 1)
DEFINE_AVOS(ushort, uint16_t, uint16_t, 0)
DEFINE_AVOS(int, int32_t, uint32_t, 1)
DEFINE_AVOS(uint, uint32_t, uint32_t, 0)
DEFINE_AVOS(long, int64_t, uint64_t, 1)
DEFINE_AVOS(ulong, uint64_t, uint64_t, 0)

/* ------------------------------------------------------------------------ */
/* Naive in-place Floyd-Warshall with the C core's diameter and error rules  */
/* (warshall.c.in:50-107).                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t n_cycle;            /* number of cycle detections                      */
  int64_t first_k, first_i;   /* first detection (naive k,i order)               */
  int64_t first_a, first_b;   /* operands W[i][k], W[k][j] at that moment        */
  int64_t last_k, last_i, last_a, last_b;
  int64_t n_overflow;         /* number of products raising OverflowError        */
  int64_t ov_k, ov_i, ov_j;   /* last overflow position                          */
  int64_t ov_a, ov_b, ov_kind;
  int64_t last_is_overflow;   /* 1 if the error set last was an OverflowError    */
  int64_t hang;               /* 1 if the reference would not terminate          */
  int64_t diameter;           /* MSB(max trajectory value) as the C core returns */
} orc_fw_report;

#define DEFINE_WARSHALL(NAME, T, U)                                                  \
  void orc_##NAME##_warshall(T *W, int64_t n, orc_fw_report *rep) {                  \
    T diameter = 0;                                                                  \
    memset(rep, 0, sizeof(*rep));                                                    \
    for (int64_t k = 0; k < n; k++) {                                                \
      for (int64_t i = 0; i < n; i++) {                                              \
        for (int64_t j = 0; j < n; j++) {                                            \
          int ev = 0;                                                                \
          T a = W[i * n + k], b = W[k * n + j];                                      \
          T product = (T)orc_##NAME##_avos_product(a, b, &ev);                       \
          if (ev & OE_HANG) rep->hang = 1;                                           \
          if (ev & (OE_WIDTH | OE_RED)) {                                            \
            rep->n_overflow++;                                                       \
            rep->ov_k = k; rep->ov_i = i; rep->ov_j = j;                             \
            rep->ov_a = (int64_t)a; rep->ov_b = (int64_t)b;                          \
            rep->ov_kind = (ev & OE_RED) ? 2 : 1;                                    \
            rep->last_is_overflow = 1;                                               \
          }                                                                          \
          if (i == j && !(product == (T)-1 || product == 0 || product == 1)) {       \
            if (rep->n_cycle == 0) {                                                 \
              rep->first_k = k; rep->first_i = i;                                    \
              rep->first_a = (int64_t)a; rep->first_b = (int64_t)b;                  \
            }                                                                        \
            rep->n_cycle++;                                                          \
            rep->last_k = k; rep->last_i = i;                                        \
            rep->last_a = (int64_t)a; rep->last_b = (int64_t)b;                      \
            rep->last_is_overflow = 0;                                               \
          }                                                                          \
          T *op = &W[i * n + j];                                                     \
          *op = orc_##NAME##_avos_sum(*op, product);                                 \
          if ((U)(~*op) != 0 && *op > diameter) diameter = *op;                      \
        }                                                                            \
      }                                                                              \
    }                                                                                \
    { int ev = 0; rep->diameter = NAME##_msb(diameter, &ev); }                       \
  }

DEFINE_WARSHALL(byte, int8_t, uint8_t)
DEFINE_WARSHALL(ubyte, uint8_t, uint8_t)
DEFINE_WARSHALL(short, int16_t, uint16_t)
DEFINE_WARSHALL(ushort, uint16_t, uint16_t)
DEFINE_WARSHALL(int, int32_t, uint32_t)
DEFINE_WARSHALL(uint, uint32_t, uint32_t)
DEFINE_WARSHALL(long, int64_t, uint64_t)
DEFINE_WARSHALL(ulong, uint64_t, uint64_t)

/* ------------------------------------------------------------------------ */
/* Cython int32 Floyd-Warshall (_shortest_path.pyx:340-389): stops at the    */
/* first detection, no overflow checks, optional i<=k<=j triangular loop.    */
/* Returns 1 on a detection (CycleError), 0 otherwise.                       */
/* ------------------------------------------------------------------------ */
static int cy_msb(int32_t x) { int b = 0; while (x > 1) { x >>= 1; b++; } return b; }
static int32_t cy_sum(int32_t x, int32_t y) {
  if (x == 0) return y;
  if (y == 0) return x;
  return x < y ? x : y;
}
/* _rbg_math.pxi:33-87 */
int32_t orc_cy_avos_product(int32_t lhs, int32_t rhs) {
  uint32_t x = (uint32_t)lhs, y = (uint32_t)rhs;
  if (x == 0 || y == 0) return 0;
  if (lhs == -1 && rhs == -1) return -1;
  if (lhs == 1 && rhs == 1) return 1;
  if (lhs == -1 && rhs == 1) return 0;
  if (lhs == 1 && rhs == -1) return 0;
  if (lhs == -1) x = 1;
  if (rhs == -1) return (lhs & 1) ? 0 : (int32_t)x;
  if (rhs == 1) return (lhs & 1) ? (int32_t)x : 0;
  int bp = cy_msb((int32_t)y);
  return (int32_t)((y & ((1u << bp) - 1u)) | (x << bp));
}
int orc_cy_floyd_warshall(int32_t *dm, int64_t n, int upper, orc_fw_report *rep) {
  int32_t diameter = 0;
  memset(rep, 0, sizeof(*rep));
  for (int64_t k = 0; k < n; k++) {
    int64_t i_end = upper ? k + 1 : n;
    for (int64_t i = 0; i < i_end; i++) {
      for (int64_t j = upper ? k : 0; j < n; j++) {
        int32_t a = dm[i * n + k], b = dm[k * n + j];
        int32_t d = orc_cy_avos_product(a, b);
        if (i == j && !(d == -1 || d == 0 || d == 1)) {
          rep->n_cycle = 1;
          rep->first_k = rep->last_k = k;
          rep->first_i = rep->last_i = i;
          rep->first_a = rep->last_a = a;
          rep->first_b = rep->last_b = b;
          return 1;
        }
        dm[i * n + j] = cy_sum(dm[i * n + j], d);
        if (dm[i * n + j] > diameter) diameter = dm[i * n + j];
      }
    }
  }
  rep->diameter = cy_msb(diameter);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Dense AVOS matrix product C = A (x) B as the einsum 'ij,jk' loop computes */
/* it: accum = avos_sum(accum, avos_product(a_ij, b_jk)) (redblack.c.src:    */
/* 94-150, 554-700).  Returns OR of events.                                  */
/* ------------------------------------------------------------------------ */
#define DEFINE_MATMUL(NAME, T, U)                                                    \
  int orc_##NAME##_matmul(const T *A, const T *B, T *C, int64_t M, int64_t K,        \
                          int64_t N) {                                               \
    int ev = 0;                                                                      \
    for (int64_t i = 0; i < M; i++)                                                  \
      for (int64_t k = 0; k < N; k++) {                                              \
        T acc = 0;                                                                   \
        for (int64_t j = 0; j < K; j++)                                              \
          acc = orc_##NAME##_avos_sum(acc,                                           \
                   (T)orc_##NAME##_avos_product(A[i * K + j], B[j * N + k], &ev));   \
        C[i * N + k] = acc;                                                          \
      }                                                                              \
    return ev;                                                                       \
  }
DEFINE_MATMUL(byte, int8_t, uint8_t)
DEFINE_MATMUL(ubyte, uint8_t, uint8_t)
DEFINE_MATMUL(short, int16_t, uint16_t)
DEFINE_MATMUL(ushort, uint16_t, uint16_t)
DEFINE_MATMUL(int, int32_t, uint32_t)
DEFINE_MATMUL(uint, uint32_t, uint32_t)
DEFINE_MATMUL(long, int64_t, uint64_t)
DEFINE_MATMUL(ulong, uint64_t, uint64_t)

/* ------------------------------------------------------------------------ */
/* Relational composition, C core (relational_composition.c.in:51-267).      */
/* ------------------------------------------------------------------------ */
#define DEFINE_COMPOSE(NAME, T, U)                                                   \
  /* edge: out = R with edge alpha->beta (value np) composed in (:72-138) */         \
  int orc_##NAME##_edge_compose(const T *R, int64_t n, int64_t alpha, int64_t beta,  \
                                T np, T *out) {                                      \
    int ev = 0;                                                                      \
    for (int64_t p = 0; p < n; p++) {                                                \
      T r = 0;                                                                       \
      for (int64_t i = 0; i < n; i++) {                                              \
        T lhs = R[alpha * n + i];                                                    \
        if (np != 0 && i == beta) lhs = np;                                          \
        r = orc_##NAME##_avos_sum(r, (T)orc_##NAME##_avos_product(lhs, R[i * n + p], &ev)); \
      }                                                                              \
      out[alpha * n + p] = r;                                                        \
    }                                                                                \
    for (int64_t m = 0; m < n; m++) {                                                \
      if (m == alpha) continue;                                                      \
      T lhs = R[m * n + alpha];                                                      \
      for (int64_t p = 0; p < n; p++) {                                              \
        T u = out[alpha * n + p];                                                    \
        if (u != 0)                                                                  \
          out[m * n + p] = orc_##NAME##_avos_sum(R[m * n + p],                       \
                              (T)orc_##NAME##_avos_product(lhs, u, &ev));            \
        else                                                                         \
          out[m * n + p] = R[m * n + p];                                             \
      }                                                                              \
    }                                                                                \
    return ev;                                                                       \
  }                                                                                  \
  /* vertex: out (n+1)x(n+1); returns -1-idx on a cycle violation (:166-267) */      \
  int64_t orc_##NAME##_vertex_compose(const T *u, const T *R, const T *v, T color,   \
                                      int64_t n, T *out, int *evp) {                 \
    int ev = 0;                                                                      \
    int64_t m1 = n + 1;                                                              \
    for (int64_t c = 0; c < n; c++) {                                                \
      T r = 0;                                                                       \
      for (int64_t i = 0; i < n; i++)                                                \
        r = orc_##NAME##_avos_sum(r, (T)orc_##NAME##_avos_product(u[i], R[i * n + c], &ev)); \
      out[n * m1 + c] = r;                                                           \
    }                                                                                \
    out[n * m1 + n] = color;                                                         \
    for (int64_t rr = 0; rr < n; rr++) {                                             \
      T r = 0;                                                                       \
      for (int64_t i = 0; i < n; i++)                                                \
        r = orc_##NAME##_avos_sum(r, (T)orc_##NAME##_avos_product(R[rr * n + i], v[i], &ev)); \
      out[rr * m1 + n] = r;                                                          \
    }                                                                                \
    for (int64_t i = 0; i < n; i++) {                                                \
      if (out[i * m1 + n] != 0 && out[n * m1 + i] != 0) { *evp = ev; return -1 - i; } \
    }                                                                                \
    for (int64_t a = 0; a < n; a++)                                                  \
      for (int64_t b = 0; b < n; b++)                                                \
        out[a * m1 + b] = orc_##NAME##_avos_sum(R[a * n + b],                        \
            (T)orc_##NAME##_avos_product(out[a * m1 + n], out[n * m1 + b], &ev));    \
    *evp = ev;                                                                       \
    return 0;                                                                        \
  }
DEFINE_COMPOSE(int, int32_t, uint32_t)
DEFINE_COMPOSE(long, int64_t, uint64_t)

/* ------------------------------------------------------------------------ */
/* SpGEMM, C++ sparsetools (rbm_math.h:116-222), int64 values, int32 index.  */
/* pass1 fills Cp; pass2 fills Cj/Cx in linked-list order, dropping zeros.   */
/* ------------------------------------------------------------------------ */
int64_t orc_spgemm_pass1(int32_t n_row, int32_t n_col, const int32_t *Ap, const int32_t *Aj,
                         const int32_t *Bp, const int32_t *Bj, int64_t *Cp) {
  int32_t *mask = (int32_t *)malloc(sizeof(int32_t) * (n_col > 0 ? n_col : 1));
  for (int32_t c = 0; c < n_col; c++) mask[c] = -1;
  int64_t nnz = 0;
  Cp[0] = 0;
  for (int32_t i = 0; i < n_row; i++) {
    int64_t row_nnz = 0;
    for (int32_t jj = Ap[i]; jj < Ap[i + 1]; jj++) {
      int32_t j = Aj[jj];
      for (int32_t kk = Bp[j]; kk < Bp[j + 1]; kk++) {
        int32_t k = Bj[kk];
        if (mask[k] != i) { mask[k] = i; row_nnz++; }
      }
    }
    nnz += row_nnz;
    Cp[i + 1] = nnz;
  }
  free(mask);
  return nnz;
}
int orc_spgemm_pass2(int32_t n_row, int32_t n_col, const int32_t *Ap, const int32_t *Aj,
                     const int64_t *Ax, const int32_t *Bp, const int32_t *Bj, const int64_t *Bx,
                     int64_t *Cp, int32_t *Cj, int64_t *Cx) {
  int ev = 0;
  int32_t *next = (int32_t *)malloc(sizeof(int32_t) * (n_col > 0 ? n_col : 1));
  int64_t *sums = (int64_t *)calloc(n_col > 0 ? n_col : 1, sizeof(int64_t));
  for (int32_t c = 0; c < n_col; c++) next[c] = -1;
  int64_t nnz = 0;
  Cp[0] = 0;
  for (int32_t i = 0; i < n_row; i++) {
    int32_t head = -2, length = 0;
    for (int32_t jj = Ap[i]; jj < Ap[i + 1]; jj++) {
      int32_t j = Aj[jj];
      int64_t v = Ax[jj];
      for (int32_t kk = Bp[j]; kk < Bp[j + 1]; kk++) {
        int32_t k = Bj[kk];
        sums[k] = orc_long_avos_sum(sums[k], (int64_t)orc_long_avos_product(v, Bx[kk], &ev));
        if (next[k] == -1) { next[k] = head; head = k; length++; }
      }
    }
    for (int32_t jj = 0; jj < length; jj++) {
      if (sums[head] != 0) { Cj[nnz] = head; Cx[nnz] = sums[head]; nnz++; }
      int32_t temp = head;
      head = next[head];
      next[temp] = -1;
      sums[temp] = 0;
    }
    Cp[i + 1] = nnz;
  }
  free(next);
  free(sums);
  return ev;
}

/* ------------------------------------------------------------------------ */
/* Topological sort: iterative DFS of _topological_sort.pyx:30-163.          */
/* Writes order[] (sources first); returns -1 on success or the vertex the   */
/* reference's CycleError names (the GRAY neighbour of the first back edge). */
/* ------------------------------------------------------------------------ */
int64_t orc_topological_sort(int32_t n, const int32_t *indptr, const int32_t *indices,
                             int32_t *order) {
  int32_t *color = (int32_t *)calloc(n > 0 ? n : 1, sizeof(int32_t));
  int32_t *st_v = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t *st_i = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t order_idx = n - 1;
  int64_t bad = -1;
  for (int32_t s = 0; s < n && bad < 0; s++) {
    if (color[s] != 0) continue;
    int32_t top = 0;
    st_v[0] = s; st_i[0] = indptr[s]; color[s] = 1;
    while (top >= 0 && bad < 0) {
      int32_t vtx = st_v[top];
      int32_t idx = st_i[top];
      int32_t end = indptr[vtx + 1];
      int pushed = 0;
      while (idx < end) {
        int32_t nb = indices[idx++];
        if (nb == vtx) continue;
        if (color[nb] == 0) {
          st_i[top] = idx;
          top++;
          st_v[top] = nb; st_i[top] = indptr[nb]; color[nb] = 1;
          pushed = 1;
          break;
        } else if (color[nb] == 1) {
          bad = nb;
          break;
        }
      }
      if (bad >= 0) break;
      if (!pushed) {
        color[vtx] = 2;
        order[order_idx--] = vtx;
        top--;
      }
    }
  }
  free(color); free(st_v); free(st_i);
  return bad;
}

/* ------------------------------------------------------------------------ */
/* Sparse DAG closure (_transitive_closure_dag.pyx:168-351), generic value   */
/* width.  Rows are built in reverse topological order; each closure row is  */
/* a column-indexed map (a stamped dense position table replaces the        */
/* reference's per-row hash -- the row contents do not depend on it).        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t n, nnz;
  int64_t *indptr;
  int32_t *indices;
  int64_t *data;
  int64_t cand_max;   /* max |candidate| (direct edges and nonzero products) */
  int64_t final_max;  /* max |final value| over entries with |v| > 1           */
  int64_t cycle_vertex;
  int64_t events;
} orc_dag_result;

typedef struct { int32_t *cols; int64_t *vals; int32_t size, cap; } orc_row;

static void row_push(orc_row *r, int32_t c, int64_t v) {
  if (r->size == r->cap) {
    r->cap = r->cap ? r->cap * 2 : 8;
    r->cols = (int32_t *)realloc(r->cols, sizeof(int32_t) * r->cap);
    r->vals = (int64_t *)realloc(r->vals, sizeof(int64_t) * r->cap);
  }
  r->cols[r->size] = c; r->vals[r->size] = v; r->size++;
}

static int cmp_pair(const void *a, const void *b) {
  const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
  return (x[0] > y[0]) - (x[0] < y[0]);
}

/* width 32 uses the int32 Cython arithmetic; width 64 the C int64 rules */
orc_dag_result *orc_dag_closure(int32_t n, const int32_t *indptr, const int32_t *indices,
                                const int64_t *data, int width) {
  orc_dag_result *res = (orc_dag_result *)calloc(1, sizeof(orc_dag_result));
  res->n = n;
  res->cycle_vertex = -1;
  res->indptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  if (n == 0) return res;
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n);
  int64_t bad = orc_topological_sort(n, indptr, indices, order);
  if (bad >= 0) { res->cycle_vertex = bad; free(order); return res; }
  orc_row *rows = (orc_row *)calloc(n, sizeof(orc_row));
  int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t *stamp = (int32_t *)malloc(sizeof(int32_t) * n);
  for (int32_t i = 0; i < n; i++) stamp[i] = -1;
  int64_t max_value = 0;
  int ev = 0;
  for (int32_t t = n - 1; t >= 0; t--) {
    int32_t v = order[t];
    orc_row *rv = &rows[v];
    for (int32_t idx = indptr[v]; idx < indptr[v + 1]; idx++) {
      int32_t w = indices[idx];
      int64_t val = data[idx];
      /* sparse_row_add: direct edges are appended without a lookup */
      stamp[w] = v; pos[w] = rv->size;
      row_push(rv, w, val);
      int64_t av = val < 0 ? -val : val;
      if (av > max_value) max_value = av;
    }
    for (int32_t idx = indptr[v]; idx < indptr[v + 1]; idx++) {
      int32_t w = indices[idx];
      if (w == v) continue;
      int64_t vw = data[idx];
      orc_row *rw = &rows[w];
      for (int32_t q = 0; q < rw->size; q++) {
        int32_t x = rw->cols[q];
        int64_t wx = rw->vals[q];
        int64_t vx;
        if (width == 32) vx = orc_cy_avos_product((int32_t)vw, (int32_t)wx);
        else vx = (int64_t)orc_long_avos_product(vw, wx, &ev);
        if (vx == 0) continue;
        if (stamp[x] == v) {
          int64_t *cur = &rv->vals[pos[x]];
          if (width == 32) *cur = cy_sum((int32_t)*cur, (int32_t)vx);
          else *cur = orc_long_avos_sum(*cur, vx);
        } else {
          stamp[x] = v; pos[x] = rv->size;
          row_push(rv, x, vx);
        }
        int64_t av = vx < 0 ? -vx : vx;
        if (av > max_value) max_value = av;
      }
    }
  }
  int64_t total = 0;
  for (int32_t v = 0; v < n; v++) total += rows[v].size;
  res->nnz = total;
  res->indices = (int32_t *)malloc(sizeof(int32_t) * (total ? total : 1));
  res->data = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
  int64_t out = 0, fmax = 0;
  int64_t *pairs = NULL; int32_t pcap = 0;
  for (int32_t v = 0; v < n; v++) {
    res->indptr[v] = out;
    orc_row *r = &rows[v];
    if (r->size > pcap) { pcap = r->size; pairs = (int64_t *)realloc(pairs, sizeof(int64_t) * 2 * pcap); }
    for (int32_t q = 0; q < r->size; q++) { pairs[2 * q] = r->cols[q]; pairs[2 * q + 1] = r->vals[q]; }
    qsort(pairs, r->size, 2 * sizeof(int64_t), cmp_pair);
    for (int32_t q = 0; q < r->size; q++) {
      res->indices[out] = (int32_t)pairs[2 * q];
      res->data[out] = pairs[2 * q + 1];
      int64_t av = pairs[2 * q + 1] < 0 ? -pairs[2 * q + 1] : pairs[2 * q + 1];
      if (av > 1 && av > fmax) fmax = av;
      out++;
    }
    free(r->cols); free(r->vals);
  }
  res->indptr[n] = out;
  res->cand_max = max_value;
  res->final_max = fmax;
  res->events = ev;
  free(pairs); free(rows); free(pos); free(stamp); free(order);
  return res;
}

void orc_dag_free(orc_dag_result *r) {
  if (!r) return;
  free(r->indptr); free(r->indices); free(r->data); free(r);
}
