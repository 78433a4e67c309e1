/*
 * oracle.c -- CPU restatement of Poutre2's morphology / geodesy hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path, never part of it.
 * See oracle.h for the list of functions, the reference file:line each follows and
 * how each is pinned.  Written from the reference's documented behaviour; the
 * code is this repository's own (plain C99, one instantiation per element type
 * through the OR_DEFINE macro at the bottom).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_E_RANK (-1)
#define OR_E_SHAPE (-2)
#define OR_E_SE (-3)
#define OR_E_ARG (-4)
#define OR_E_TYPE (-5)
#define OR_E_NOMEM (-6)

/* ------------------------------------------------------------------------- */
/* Structuring elements: coordinates of the predefined neighbour lists,       */
/* neighbor_list_static_se_t.hpp:37-432 (centre first; {dz,dy,dx} order).     */
/* ------------------------------------------------------------------------- */
int or_se_offsets(int se, int64_t *offs, int *rank)
{
  int n = 0;
  switch (se) {
  case OR_SE_SEGX1D:
    *rank = 1;
    offs[0] = 0; offs[1] = -1; offs[2] = 1;
    return 3;
  case OR_SE_SQUARE2D:
    *rank = 2;
    offs[n * 2] = 0; offs[n * 2 + 1] = 0; n++;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (!dy && !dx) continue;
        offs[n * 2] = dy; offs[n * 2 + 1] = dx; n++;
      }
    return n;
  case OR_SE_CROSS2D: {
    static const int c[5][2] = {{0, 0}, {-1, 0}, {0, -1}, {0, 1}, {1, 0}};
    *rank = 2;
    for (n = 0; n < 5; n++) { offs[n * 2] = c[n][0]; offs[n * 2 + 1] = c[n][1]; }
    return 5;
  }
  case OR_SE_SEGX2D:
    *rank = 2;
    offs[0] = 0; offs[1] = 0; offs[2] = 0; offs[3] = -1; offs[4] = 0; offs[5] = 1;
    return 3;
  case OR_SE_SEGY2D:
    *rank = 2;
    offs[0] = 0; offs[1] = 0; offs[2] = -1; offs[3] = 0; offs[4] = 1; offs[5] = 0;
    return 3;
  case OR_SE_SEGX3D:
  case OR_SE_SEGY3D:
  case OR_SE_SEGZ3D: {
    int ax = (se == OR_SE_SEGX3D) ? 2 : (se == OR_SE_SEGY3D) ? 1 : 0;
    *rank = 3;
    memset(offs, 0, 9 * sizeof(int64_t));
    offs[3 + ax] = -1;
    offs[6 + ax] = 1;
    return 3;
  }
  case OR_SE_CROSS3D: {
    static const int c[7][3] = {{0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
    *rank = 3;
    for (n = 0; n < 7; n++) { offs[n * 3] = c[n][0]; offs[n * 3 + 1] = c[n][1]; offs[n * 3 + 2] = c[n][2]; }
    return 7;
  }
  case OR_SE_SQUARE3D:
    *rank = 3;
    offs[0] = offs[1] = offs[2] = 0; n = 1;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (!dz && !dy && !dx) continue;
          offs[n * 3] = dz; offs[n * 3 + 1] = dy; offs[n * 3 + 2] = dx; n++;
        }
    return n;
  default:
    return OR_E_SE;
  }
}

/* raster position of an offset relative to the centre: <0 upper (already visited
 * in a forward raster scan), >0 lower (neighbor_list_static_se_t.hpp:253-268,296-307
 * for 2D).  For 3D the reference orders the split with the last axis slowest
 * (:398-431), which is not raster-causal; this restatement uses the causal split. */
static int raster_sign(const int64_t *o, int rank)
{
  for (int a = 0; a < rank; a++) {
    if (o[a] < 0) return -1;
    if (o[a] > 0) return 1;
  }
  return 0;
}

static int64_t numel(int rank, const int64_t *shape)
{
  int64_t n = 1;
  for (int a = 0; a < rank; a++) n *= shape[a];
  return n;
}

/* in-bounds test of p + o (array_view.hpp:578-609 `contains`) and linear index */
static int shifted(const int64_t *p, const int64_t *o, int rank, const int64_t *shape, int64_t *lin)
{
  int64_t l = 0;
  for (int a = 0; a < rank; a++) {
    int64_t c = p[a] + o[a];
    if (c < 0 || c >= shape[a]) return 0;
    l = l * shape[a] + c;
  }
  *lin = l;
  return 1;
}

static void unravel(int64_t lin, int rank, const int64_t *shape, int64_t *p)
{
  for (int a = rank - 1; a >= 0; a--) {
    p[a] = lin % shape[a];
    lin /= shape[a];
  }
}

/* ------------------------------------------------------------------------- */
/* simple binary max-heap of (key = linear index) used by the leveling        */
/* restatement; std::priority_queue with lesserKey pops the largest key first */
/* (include/poutre/base/details/data_structures/pq.hpp:87-93,159).            */
/* ------------------------------------------------------------------------- */
typedef struct {
  int64_t *k;
  int64_t n, cap;
} heap_t;

static int heap_push(heap_t *h, int64_t key)
{
  if (h->n == h->cap) {
    int64_t nc = h->cap ? h->cap * 2 : 1024;
    int64_t *nk = (int64_t *)realloc(h->k, (size_t)nc * sizeof(int64_t));
    if (!nk) return OR_E_NOMEM;
    h->k = nk;
    h->cap = nc;
  }
  int64_t i = h->n++;
  h->k[i] = key;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h->k[p] >= h->k[i]) break;
    int64_t tmp = h->k[p]; h->k[p] = h->k[i]; h->k[i] = tmp;
    i = p;
  }
  return OR_OK;
}

static int64_t heap_pop(heap_t *h)
{
  int64_t top = h->k[0];
  h->k[0] = h->k[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->n && h->k[l] > h->k[m]) m = l;
    if (r < h->n && h->k[r] > h->k[m]) m = r;
    if (m == i) break;
    int64_t tmp = h->k[m]; h->k[m] = h->k[i]; h->k[i] = tmp;
    i = m;
  }
  return top;
}

/* ------------------------------------------------------------------------- */
/* Per-type implementation                                                    */
/* ------------------------------------------------------------------------- */
#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))
#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))

/* OR_DEFINE(SUFFIX, T, LOWEST, HIGHEST, TMIN, ACCU)
 *   LOWEST/HIGHEST = numeric_limits<T>::lowest()/max(): neutral of dilation/erosion
 *   TMIN           = numeric_limits<T>::min(): floor of saturated subtraction
 *                    (types_traits.hpp:114-263; FLT_MIN for float, arith_op_t.hpp:91-95)
 *   ACCU           = accumulation type of saturated arithmetic                      */
#define OR_DEFINE(S, T, LOWEST, HIGHEST, TMIN, ACCU)                                                          \
  /* van_herck_1d, ero_dil_line_se_t.hpp:49-89: f,g,h point at element 0 of buffers padded by r */            \
  static void van_herk_##S(T *f, T *g, T *h, int64_t n, int64_t r, int dil)                                   \
  {                                                                                                           \
    if (n == 0) return;                                                                                       \
    const int64_t alpha = 2 * r + 1, size = n + 2 * r;                                                        \
    for (int64_t cs = -r, rem = size; rem > 0; cs += alpha, rem -= alpha) {                                   \
      int64_t len = OR_MIN(rem, alpha);                                                                       \
      T acc = f[cs];                                                                                          \
      g[cs] = acc;                                                                                            \
      for (int64_t i = 1; i < len; i++) {                                                                     \
        acc = dil ? OR_MAX(acc, f[cs + i]) : OR_MIN(acc, f[cs + i]);                                          \
        g[cs + i] = acc;                                                                                      \
      }                                                                                                       \
    }                                                                                                         \
    for (int64_t cs = -r, rem = size; rem > 0; cs += alpha, rem -= alpha) {                                  \
      int64_t len = OR_MIN(rem, alpha);                                                                       \
      T acc = f[cs + len - 1];                                                                                \
      h[cs + len - 1] = acc;                                                                                  \
      for (int64_t i = len - 2; i >= 0; i--) {                                                                \
        acc = dil ? OR_MAX(acc, f[cs + i]) : OR_MIN(acc, f[cs + i]);                                          \
        h[cs + i] = acc;                                                                                      \
      }                                                                                                       \
    }                                                                                                         \
    for (int64_t i = 0; i < n; i++) f[i] = dil ? OR_MAX(h[i - r], g[i + r]) : OR_MIN(h[i - r], g[i + r]);     \
  }                                                                                                           \
                                                                                                              \
  /* t_ErodeDilateXOpLineDispatcher 1D/2D, ero_dil_line_se_t.hpp:103-190 */                                   \
  static int line_x_##S(const T *in, T *out, int rank, const int64_t *shape, int64_t r, int dil)              \
  {                                                                                                           \
    if (rank != 1 && rank != 2) return OR_E_RANK;                                                             \
    int64_t rows = rank == 1 ? 1 : shape[0], w = shape[rank - 1];                                             \
    if (w == 0 || rows == 0) return OR_OK;                                                                    \
    if (r <= 0) {                                                                                             \
      memcpy(out, in, (size_t)(rows * w) * sizeof(T));                                                        \
      return OR_OK;                                                                                           \
    }                                                                                                         \
    size_t padded = (size_t)(w + 2 * r);                                                                      \
    T *f = (T *)malloc(padded * sizeof(T)), *g = (T *)malloc(padded * sizeof(T)),                             \
      *h = (T *)malloc(padded * sizeof(T));                                                                   \
    if (!f || !g || !h) { free(f); free(g); free(h); return OR_E_NOMEM; }                                     \
    const T neutral = dil ? (LOWEST) : (HIGHEST);                                                             \
    for (int64_t y = 0; y < rows; y++) {                                                                      \
      for (size_t i = 0; i < padded; i++) f[i] = neutral;                                                     \
      memcpy(f + r, in + y * w, (size_t)w * sizeof(T));                                                       \
      van_herk_##S(f + r, g + r, h + r, w, r, dil);                                                           \
      memcpy(out + y * w, f + r, (size_t)w * sizeof(T));                                                      \
    }                                                                                                         \
    free(f); free(g); free(h);                                                                                \
    return OR_OK;                                                                                             \
  }                                                                                                           \
                                                                                                              \
  /* t_transposeDispatcher 2D, transpose_op_t.hpp:45-70: out[y*oH + x] = in[x*oW + y] */                       \
  static void transpose_##S(const T *in, T *out, int64_t H, int64_t W)                                        \
  {                                                                                                           \
    const int64_t oH = W, oW = H;                                                                             \
    for (int64_t y = 0; y < oW; y++)                                                                          \
      for (int64_t x = 0; x < oH; x++) out[y * oH + x] = in[x * oW + y];                                      \
  }                                                                                                           \
                                                                                                              \
  /* t_DilateY / t_ErodeY, ero_dil_line_se_t.hpp:248-280: the temporaries take the input's   */               \
  /* shape, so t_transpose rejects H != W (transpose_op_t.hpp:54-60)                          */               \
  static int line_y_##S(const T *in, T *out, int rank, const int64_t *shape, int64_t r, int dil)              \
  {                                                                                                           \
    if (rank != 2) return OR_E_RANK;                                                                          \
    if (shape[0] != shape[1]) return OR_E_SHAPE;                                                              \
    int64_t n = shape[0] * shape[1];                                                                          \
    T *t1 = (T *)malloc((size_t)n * sizeof(T)), *t2 = (T *)malloc((size_t)n * sizeof(T));                     \
    if (!t1 || !t2) { free(t1); free(t2); return OR_E_NOMEM; }                                                \
    transpose_##S(in, t1, shape[0], shape[1]);                                                                \
    int rc = line_x_##S(t1, t2, 2, shape, r, dil);                                                            \
    transpose_##S(t2, out, shape[0], shape[1]);                                                               \
    free(t1); free(t2);                                                                                       \
    return rc;                                                                                                \
  }                                                                                                           \
                                                                                                              \
  /* max/min over an offset list with out-of-bounds taps skipped,                     */                      \
  /* ero_dil_runtime_nl_se_t.hpp:48-97 (the static 3x3 paths compute the same value,  */                      \
  /* their line buffers being padded with the neutral, ero_dil_static_se_t.hpp:148-212) */                    \
  static int nl_##S(const T *in, T *out, int rank, const int64_t *shape, const int64_t *offs, int n, int dil)  \
  {                                                                                                           \
    if (rank < 1 || rank > 3) return OR_E_RANK;                                                               \
    int64_t N = numel(rank, shape), p[3], lin;                                                                \
    for (int64_t i = 0; i < N; i++) {                                                                         \
      unravel(i, rank, shape, p);                                                                             \
      T acc = dil ? (LOWEST) : (HIGHEST);                                                                     \
      for (int k = 0; k < n; k++) {                                                                           \
        if (!shifted(p, offs + k * rank, rank, shape, &lin)) continue;                                        \
        acc = dil ? OR_MAX(acc, in[lin]) : OR_MIN(acc, in[lin]);                                              \
      }                                                                                                       \
      out[i] = acc;                                                                                           \
    }                                                                                                         \
    return OR_OK;                                                                                             \
  }                                                                                                           \
                                                                                                              \
  static int static_##S(const T *in, T *out, int rank, const int64_t *shape, int se, int dil)                 \
  {                                                                                                           \
    int64_t offs[27 * 3];                                                                                     \
    int serank = 0;                                                                                           \
    int n = or_se_offsets(se, offs, &serank);                                                                 \
    if (n < 0) return OR_E_SE;                                                                                \
    if (serank != rank) return OR_E_RANK;                                                                     \
    return nl_##S(in, out, rank, shape, offs, n, dil);                                                        \
  }                                                                                                           \
                                                                                                              \
  /* Erode/Dilate(iter), src/low_level_morpho/ero_dil_static.cpp:229-305 */                                   \
  static int static_iter_##S(const T *in, T *out, int rank, const int64_t *shape, int se, int iter, int dil)  \
  {                                                                                                           \
    if (iter < 0) return OR_E_ARG;                                                                            \
    int64_t N = numel(rank, shape);                                                                           \
    if (iter == 0) { memcpy(out, in, (size_t)N * sizeof(T)); return OR_OK; }                                  \
    if (iter == 1) return static_##S(in, out, rank, shape, se, dil);                                          \
    T *tmp = (T *)malloc((size_t)N * sizeof(T));                                                              \
    if (!tmp) return OR_E_NOMEM;                                                                              \
    memcpy(tmp, in, (size_t)N * sizeof(T));                                                                   \
    int rc = OR_OK;                                                                                           \
    int odd = iter % 2, even = iter - odd;                                                                    \
    for (int k = 0; k < even && rc == OR_OK; k += 2) {                                                        \
      rc = static_##S(tmp, out, rank, shape, se, dil);                                                        \
      if (rc == OR_OK) rc = static_##S(out, tmp, rank, shape, se, dil);                                       \
    }                                                                                                         \
    if (rc == OR_OK) {                                                                                        \
      if (odd == 0) memcpy(out, tmp, (size_t)N * sizeof(T));                                                  \
      else rc = static_##S(tmp, out, rank, shape, se, dil);                                                   \
    }                                                                                                         \
    free(tmp);                                                                                                \
    return rc;                                                                                                \
  }                                                                                                           \
                                                                                                              \
  /* t_DilateOctagon / t_DilateRhombicuboctahedron, ero_dil_compound_static_se_t.hpp:33-99 */                 \
  static int compound_##S(const T *in, T *out, int rank, const int64_t *shape, int compound, int size,        \
                          int dil)                                                                            \
  {                                                                                                           \
    int cross, square;                                                                                        \
    if (compound == 1) { if (rank != 2) return OR_E_RANK; cross = OR_SE_CROSS2D; square = OR_SE_SQUARE2D; }   \
    else if (compound == 2) { if (rank != 3) return OR_E_RANK; cross = OR_SE_CROSS3D; square = OR_SE_SQUARE3D; } \
    else return OR_E_SE;                                                                                      \
    if (size < 0) return OR_E_ARG;                                                                            \
    int64_t N = numel(rank, shape);                                                                           \
    if (size == 0) { memcpy(out, in, (size_t)N * sizeof(T)); return OR_OK; }                                  \
    /* size==1 on 3D applies SECross2D to a 3D image and throws (:77-79) */                                   \
    if (size == 1) return compound == 2 ? OR_E_RANK : static_##S(in, out, rank, shape, cross, dil);           \
    double q = (double)size / (1.0 + sqrt(2.0)), fl = floor(q);                                               \
    int nsq = (int)((q - fl) < 0.5 ? fl : fl + 1.0);                                                          \
    T *tmp = (T *)malloc((size_t)N * sizeof(T));                                                              \
    if (!tmp) return OR_E_NOMEM;                                                                              \
    int rc = static_##S(in, out, rank, shape, cross, dil);                                                    \
    for (int i = 1; i < size - nsq && rc == OR_OK; i++) {                                                     \
      rc = static_##S(out, tmp, rank, shape, cross, dil);                                                     \
      memcpy(out, tmp, (size_t)N * sizeof(T));                                                                \
    }                                                                                                         \
    for (int i = 0; i < nsq && rc == OR_OK; i++) {                                                            \
      rc = static_##S(out, tmp, rank, shape, square, dil);                                                    \
      memcpy(out, tmp, (size_t)N * sizeof(T));                                                                \
    }                                                                                                         \
    free(tmp);                                                                                                \
    return rc;                                                                                                \
  }                                                                                                           \
                                                                                                              \
  /* t_ReconstructionViewDispatcher, mreconstruct_t.hpp:87-194 (dil: OpRecDilate :39-55,  */                 \
  /* erosion: OpRecErode :57-73)                                                           */                 \
  static int reconstruct_##S(int dil, const T *marker, const T *mask, T *out, int rank, const int64_t *shape, \
                             int se)                                                                          \
  {                                                                                                           \
    int64_t offs[27 * 3], up[13 * 3], lo[13 * 3], nb[26 * 3];                                                 \
    int serank = 0, nup = 0, nlo = 0, nnb = 0;                                                                \
    int n = or_se_offsets(se, offs, &serank);                                                                 \
    if (n < 0) return OR_E_SE;                                                                                \
    if (serank != rank) return OR_E_RANK;                                                                     \
    for (int k = 0; k < n; k++) {                                                                             \
      int s = raster_sign(offs + k * rank, rank);                                                             \
      if (s == 0) continue;                                                                                   \
      memcpy(nb + nnb++ * rank, offs + k * rank, (size_t)rank * sizeof(int64_t));                             \
      if (s < 0) memcpy(up + nup++ * rank, offs + k * rank, (size_t)rank * sizeof(int64_t));                  \
      else memcpy(lo + nlo++ * rank, offs + k * rank, (size_t)rank * sizeof(int64_t));                        \
    }                                                                                                         \
    int64_t N = numel(rank, shape), p[3], q;                                                                  \
    /* select_marker: inf (dilate) / sup (erode) of marker and mask */                                        \
    for (int64_t i = 0; i < N; i++)                                                                           \
      out[i] = dil ? OR_MIN(marker[i], mask[i]) : OR_MAX(marker[i], mask[i]);                                 \
    /* forward raster scan over the upper half */                                                             \
    for (int64_t i = 0; i < N; i++) {                                                                         \
      unravel(i, rank, shape, p);                                                                             \
      T v = out[i];                                                                                           \
      for (int k = 0; k < nup; k++)                                                                           \
        if (shifted(p, up + k * rank, rank, shape, &q)) v = dil ? OR_MAX(v, out[q]) : OR_MIN(v, out[q]);      \
      out[i] = dil ? OR_MIN(v, mask[i]) : OR_MAX(v, mask[i]);                                                 \
    }                                                                                                         \
    /* backward scan over the lower half, seeding the FIFO */                                                 \
    int64_t *fifo = NULL, qh = 0, qt = 0, qcap = 0;                                                           \
    for (int64_t i = N - 1; i >= 0; i--) {                                                                    \
      unravel(i, rank, shape, p);                                                                             \
      T v = out[i];                                                                                           \
      for (int k = 0; k < nlo; k++)                                                                           \
        if (shifted(p, lo + k * rank, rank, shape, &q)) v = dil ? OR_MAX(v, out[q]) : OR_MIN(v, out[q]);      \
      out[i] = dil ? OR_MIN(v, mask[i]) : OR_MAX(v, mask[i]);                                                 \
      for (int k = 0; k < nlo; k++) {                                                                         \
        if (!shifted(p, lo + k * rank, rank, shape, &q)) continue;                                            \
        int enq = dil ? (out[q] < out[i] && out[q] < mask[q]) : (out[q] > out[i] && out[q] > mask[q]);        \
        if (enq) {                                                                                            \
          if (qt == qcap) {                                                                                   \
            qcap = qcap ? qcap * 2 : 4096;                                                                    \
            int64_t *nf = (int64_t *)realloc(fifo, (size_t)qcap * sizeof(int64_t));                          \
            if (!nf) { free(fifo); return OR_E_NOMEM; }                                                       \
            fifo = nf;                                                                                        \
          }                                                                                                   \
          fifo[qt++] = i;                                                                                     \
        }                                                                                                     \
      }                                                                                                       \
    }                                                                                                         \
    /* FIFO propagation over the full neighbourhood */                                                        \
    while (qh < qt) {                                                                                         \
      int64_t i = fifo[qh++];                                                                                 \
      unravel(i, rank, shape, p);                                                                             \
      for (int k = 0; k < nnb; k++) {                                                                         \
        if (!shifted(p, nb + k * rank, rank, shape, &q)) continue;                                            \
        int grow = dil ? (out[q] < out[i]) : (out[q] > out[i]);                                               \
        if (mask[q] != out[q] && grow) {                                                                      \
          out[q] = dil ? OR_MIN(out[i], mask[q]) : OR_MAX(out[i], mask[q]);                                   \
          if (qt == qcap) {                                                                                   \
            /* compact the consumed head before growing */                                                    \
            if (qh > 0) { memmove(fifo, fifo + qh, (size_t)(qt - qh) * sizeof(int64_t)); qt -= qh; qh = 0; }  \
            if (qt == qcap) {                                                                                 \
              qcap = qcap ? qcap * 2 : 4096;                                                                  \
              int64_t *nf = (int64_t *)realloc(fifo, (size_t)qcap * sizeof(int64_t));                        \
              if (!nf) { free(fifo); return OR_E_NOMEM; }                                                     \
              fifo = nf;                                                                                      \
            }                                                                                                 \
          }                                                                                                   \
          fifo[qt++] = q;                                                                                     \
        }                                                                                                     \
      }                                                                                                       \
    }                                                                                                         \
    free(fifo);                                                                                               \
    return OR_OK;                                                                                             \
  }                                                                                                           \
                                                                                                              \
  /* brute force: x0 = marker ^ mask; x <- mask ^ delta_SE(x) until nothing changes */                        \
  static int reconstruct_bf_##S(int dil, const T *marker, const T *mask, T *out, int rank,                    \
                                const int64_t *shape, int se)                                                 \
  {                                                                                                           \
    int64_t offs[27 * 3];                                                                                     \
    int serank = 0;                                                                                           \
    int n = or_se_offsets(se, offs, &serank);                                                                 \
    if (n < 0) return OR_E_SE;                                                                                \
    if (serank != rank) return OR_E_RANK;                                                                     \
    int64_t N = numel(rank, shape);                                                                           \
    T *tmp = (T *)malloc((size_t)N * sizeof(T));                                                              \
    if (!tmp) return OR_E_NOMEM;                                                                              \
    for (int64_t i = 0; i < N; i++) out[i] = dil ? OR_MIN(marker[i], mask[i]) : OR_MAX(marker[i], mask[i]);   \
    for (;;) {                                                                                                \
      nl_##S(out, tmp, rank, shape, offs, n, dil);                                                            \
      int changed = 0;                                                                                        \
      for (int64_t i = 0; i < N; i++) {                                                                       \
        T v = dil ? OR_MIN(tmp[i], mask[i]) : OR_MAX(tmp[i], mask[i]);                                        \
        if (v != out[i]) { out[i] = v; changed = 1; }                                                         \
      }                                                                                                       \
      if (!changed) break;                                                                                    \
    }                                                                                                         \
    free(tmp);                                                                                                \
    return OR_OK;                                                                                             \
  }                                                                                                           \
                                                                                                              \
  /* t_low_levelingHelper (low=1) / t_high_levelingHelper (low=0), leveling_t.hpp:34-202 */                   \
  static int leveling_half_##S(int low, const T *ref, const T *marker, T *out, int rank, const int64_t *shape,\
                               int se)                                                                        \
  {                                                                                                           \
    int64_t offs[27 * 3], nb[26 * 3];                                                                         \
    int serank = 0, nnb = 0;                                                                                  \
    int n = or_se_offsets(se, offs, &serank);                                                                 \
    if (n < 0) return OR_E_SE;                                                                                \
    if (serank != rank) return OR_E_RANK;                                                                     \
    for (int k = 0; k < n; k++)                                                                               \
      if (raster_sign(offs + k * rank, rank) != 0)                                                            \
        memcpy(nb + nnb++ * rank, offs + k * rank, (size_t)rank * sizeof(int64_t));                           \
    int64_t N = numel(rank, shape), p[3], q;                                                                  \
    memcpy(out, marker, (size_t)N * sizeof(T));                                                               \
    heap_t pq = {0, 0, 0};                                                                                    \
    for (int64_t i = 0; i < N; i++) {                                                                         \
      int below = low ? (out[i] < ref[i]) : (out[i] > ref[i]);                                                \
      if (!below) continue;                                                                                   \
      unravel(i, rank, shape, p);                                                                             \
      T ext = out[i];                                                                                         \
      for (int k = 0; k < nnb; k++)                                                                           \
        if (shifted(p, nb + k * rank, rank, shape, &q)) ext = low ? OR_MAX(ext, out[q]) : OR_MIN(ext, out[q]);\
      if (ext != out[i]) {                                                                                    \
        out[i] = low ? OR_MIN(ref[i], ext) : OR_MAX(ref[i], ext);                                             \
        if (heap_push(&pq, i)) { free(pq.k); return OR_E_NOMEM; }                                             \
      }                                                                                                       \
    }                                                                                                         \
    while (pq.n > 0) {                                                                                        \
      int64_t i = heap_pop(&pq);                                                                              \
      unravel(i, rank, shape, p);                                                                             \
      for (int k = 0; k < nnb; k++) {                                                                         \
        if (!shifted(p, nb + k * rank, rank, shape, &q)) continue;                                            \
        int go = low ? (out[q] < out[i] && out[q] < ref[q]) : (out[q] > out[i] && out[q] > ref[q]);           \
        if (go) {                                                                                             \
          T nv = low ? OR_MIN(ref[q], out[i]) : OR_MAX(ref[q], out[i]);                                       \
          if (out[q] != nv) {                                                                                 \
            out[q] = nv;                                                                                      \
            if (heap_push(&pq, q)) { free(pq.k); return OR_E_NOMEM; }                                         \
          }                                                                                                   \
        }                                                                                                     \
      }                                                                                                       \
    }                                                                                                         \
    free(pq.k);                                                                                               \
    /* post-processing clamp (:115 low: inf with ref; :200 high: sup with ref) */                             \
    for (int64_t i = 0; i < N; i++) out[i] = low ? OR_MIN(out[i], ref[i]) : OR_MAX(out[i], ref[i]);           \
    return OR_OK;                                                                                             \
  }                                                                                                           \
                                                                                                              \
  /* op_Saturated_Sub_Constant / op_Saturated_Add_Constant / op_Saturated_Sub,          */                    \
  /* arith_op_t.hpp:84-343: accumulate in ACCU, clamp to numeric_limits min (sub) / max (add) */               \
  static T ssub_##S(T a, T b)                                                                                 \
  {                                                                                                           \
    ACCU r = (ACCU)a - (ACCU)b;                                                                               \
    if (r < (ACCU)(TMIN)) return (TMIN);                                                                      \
    return (T)r;                                                                                              \
  }                                                                                                           \
  static T sadd_##S(T a, T b)                                                                                 \
  {                                                                                                           \
    ACCU r = (ACCU)a + (ACCU)b;                                                                               \
    if (r > (ACCU)(HIGHEST)) return (HIGHEST);                                                                \
    return (T)r;                                                                                              \
  }

OR_DEFINE(u8, uint8_t, 0, UINT8_MAX, 0, int32_t)
OR_DEFINE(u16, uint16_t, 0, UINT16_MAX, 0, int32_t)
OR_DEFINE(i32, int32_t, INT32_MIN, INT32_MAX, INT32_MIN, int64_t)
OR_DEFINE(i64, int64_t, INT64_MIN, INT64_MAX, INT64_MIN, int64_t)
OR_DEFINE(f32, float, -FLT_MAX, FLT_MAX, FLT_MIN, double)
OR_DEFINE(f64, double, -DBL_MAX, DBL_MAX, DBL_MIN, double)

/* ------------------------------------------------------------------------- */
/* type dispatch                                                              */
/* ------------------------------------------------------------------------- */
#define OR_DISPATCH(t, CALL)                                                                                  \
  switch (t) {                                                                                                \
  case OR_T_U8: { typedef uint8_t T; (void)sizeof(T); return CALL(u8); }                                      \
  case OR_T_U16: { typedef uint16_t T; (void)sizeof(T); return CALL(u16); }                                   \
  case OR_T_I32: { typedef int32_t T; (void)sizeof(T); return CALL(i32); }                                    \
  case OR_T_I64: { typedef int64_t T; (void)sizeof(T); return CALL(i64); }                                    \
  case OR_T_F32: { typedef float T; (void)sizeof(T); return CALL(f32); }                                      \
  case OR_T_F64: { typedef double T; (void)sizeof(T); return CALL(f64); }                                     \
  default: return OR_E_TYPE;                                                                                  \
  }

int or_line_x(int t, const void *in, void *out, int rank, const int64_t *shape, int64_t r, int dilate)
{
#define C(S) line_x_##S((const T *)in, (T *)out, rank, shape, r, dilate)
  OR_DISPATCH(t, C)
#undef C
}

int or_line_y(int t, const void *in, void *out, int rank, const int64_t *shape, int64_t r, int dilate)
{
#define C(S) line_y_##S((const T *)in, (T *)out, rank, shape, r, dilate)
  OR_DISPATCH(t, C)
#undef C
}

int or_transpose(int t, const void *in, void *out, const int64_t *shape)
{
#define C(S) (transpose_##S((const T *)in, (T *)out, shape[0], shape[1]), OR_OK)
  OR_DISPATCH(t, C)
#undef C
}

int or_static(int t, const void *in, void *out, int rank, const int64_t *shape, int se, int dilate)
{
#define C(S) static_##S((const T *)in, (T *)out, rank, shape, se, dilate)
  OR_DISPATCH(t, C)
#undef C
}

int or_static_iter(int t, const void *in, void *out, int rank, const int64_t *shape, int se, int iter, int dilate)
{
#define C(S) static_iter_##S((const T *)in, (T *)out, rank, shape, se, iter, dilate)
  OR_DISPATCH(t, C)
#undef C
}

int or_nl(int t, const void *in, void *out, int rank, const int64_t *shape, const int64_t *offs, int n, int dilate)
{
#define C(S) nl_##S((const T *)in, (T *)out, rank, shape, offs, n, dilate)
  OR_DISPATCH(t, C)
#undef C
}

int or_compound(int t, const void *in, void *out, int rank, const int64_t *shape, int compound, int size, int dilate)
{
#define C(S) compound_##S((const T *)in, (T *)out, rank, shape, compound, size, dilate)
  OR_DISPATCH(t, C)
#undef C
}

int or_reconstruct(int t, int rec_type, const void *marker, const void *mask, void *out, int rank,
                   const int64_t *shape, int se)
{
#define C(S) reconstruct_##S(rec_type == 1, (const T *)marker, (const T *)mask, (T *)out, rank, shape, se)
  OR_DISPATCH(t, C)
#undef C
}

int or_reconstruct_bf(int t, int rec_type, const void *marker, const void *mask, void *out, int rank,
                      const int64_t *shape, int se)
{
#define C(S) reconstruct_bf_##S(rec_type == 1, (const T *)marker, (const T *)mask, (T *)out, rank, shape, se)
  OR_DISPATCH(t, C)
#undef C
}

int or_low_leveling(int t, const void *ref, const void *marker, void *out, int rank, const int64_t *shape, int se)
{
#define C(S) leveling_half_##S(1, (const T *)ref, (const T *)marker, (T *)out, rank, shape, se)
  OR_DISPATCH(t, C)
#undef C
}

int or_high_leveling(int t, const void *ref, const void *marker, void *out, int rank, const int64_t *shape, int se)
{
#define C(S) leveling_half_##S(0, (const T *)ref, (const T *)marker, (T *)out, rank, shape, se)
  OR_DISPATCH(t, C)
#undef C
}

static size_t tsize(int t)
{
  static const size_t s[] = {1, 2, 4, 8, 4, 8};
  return (t >= 0 && t <= 5) ? s[t] : 0;
}

int or_leveling(int t, const void *ref, const void *marker, void *out, int rank, const int64_t *shape, int se)
{
  size_t bytes = (size_t)numel(rank, shape) * tsize(t);
  void *tmp = malloc(bytes ? bytes : 1);
  if (!tmp) return OR_E_NOMEM;
  int rc = or_high_leveling(t, ref, marker, tmp, rank, shape, se);
  if (rc == OR_OK) rc = or_low_leveling(t, tmp, marker, out, rank, shape, se);
  free(tmp);
  return rc;
}

/* elementwise helpers */
#define OR_EW(name, EXPR)                                                                                     \
  int name(int t, const void *a_, const void *b_, void *o_, int64_t n)                                        \
  {                                                                                                           \
    switch (t) {                                                                                              \
    case OR_T_U8: { const uint8_t *a = a_, *b = b_; uint8_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = EXPR(u8); return OR_OK; } \
    case OR_T_U16: { const uint16_t *a = a_, *b = b_; uint16_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = EXPR(u16); return OR_OK; } \
    case OR_T_I32: { const int32_t *a = a_, *b = b_; int32_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = EXPR(i32); return OR_OK; } \
    case OR_T_I64: { const int64_t *a = a_, *b = b_; int64_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = EXPR(i64); return OR_OK; } \
    case OR_T_F32: { const float *a = a_, *b = b_; float *o = o_; for (int64_t i = 0; i < n; i++) o[i] = EXPR(f32); return OR_OK; } \
    case OR_T_F64: { const double *a = a_, *b = b_; double *o = o_; for (int64_t i = 0; i < n; i++) o[i] = EXPR(f64); return OR_OK; } \
    default: return OR_E_TYPE;                                                                                \
    }                                                                                                         \
  }
#define E_SSUB(S) ssub_##S(a[i], b[i])
#define E_SADD(S) sadd_##S(a[i], b[i])
#define E_SUP(S) OR_MAX(a[i], b[i])
#define E_INF(S) OR_MIN(a[i], b[i])
OR_EW(or_arith_ssub_image, E_SSUB)
OR_EW(or_arith_sadd_image, E_SADD)
OR_EW(or_arith_sup, E_SUP)
OR_EW(or_arith_inf, E_INF)

/* op_Invert (arith_op_t.hpp:40-57): -a in the pixel type; integers negate modulo 2^bits
   (what -a converted back to the pixel type gives), floats flip the sign */
int or_arith_invert(int t, const void *a_, void *o_, int64_t n)
{
  switch (t) {
  case OR_T_U8: { const uint8_t *a = a_; uint8_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = (uint8_t)(0u - a[i]); return OR_OK; }
  case OR_T_U16: { const uint16_t *a = a_; uint16_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = (uint16_t)(0u - a[i]); return OR_OK; }
  case OR_T_I32: { const int32_t *a = a_; int32_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = (int32_t)(0u - (uint32_t)a[i]); return OR_OK; }
  case OR_T_I64: { const int64_t *a = a_; int64_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = (int64_t)(0ull - (uint64_t)a[i]); return OR_OK; }
  case OR_T_F32: { const float *a = a_; float *o = o_; for (int64_t i = 0; i < n; i++) o[i] = -a[i]; return OR_OK; }
  case OR_T_F64: { const double *a = a_; double *o = o_; for (int64_t i = 0; i < n; i++) o[i] = -a[i]; return OR_OK; }
  default: return OR_E_TYPE;
  }
}

#define OR_EWC(name, FN)                                                                                      \
  int name(int t, const void *a_, double c, void *o_, int64_t n)                                              \
  {                                                                                                           \
    switch (t) {                                                                                              \
    case OR_T_U8: { const uint8_t *a = a_; uint8_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = FN##_u8(a[i], (uint8_t)c); return OR_OK; } \
    case OR_T_U16: { const uint16_t *a = a_; uint16_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = FN##_u16(a[i], (uint16_t)c); return OR_OK; } \
    case OR_T_I32: { const int32_t *a = a_; int32_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = FN##_i32(a[i], (int32_t)c); return OR_OK; } \
    case OR_T_I64: { const int64_t *a = a_; int64_t *o = o_; for (int64_t i = 0; i < n; i++) o[i] = FN##_i64(a[i], (int64_t)c); return OR_OK; } \
    case OR_T_F32: { const float *a = a_; float *o = o_; for (int64_t i = 0; i < n; i++) o[i] = FN##_f32(a[i], (float)c); return OR_OK; } \
    case OR_T_F64: { const double *a = a_; double *o = o_; for (int64_t i = 0; i < n; i++) o[i] = FN##_f64(a[i], c); return OR_OK; } \
    default: return OR_E_TYPE;                                                                                \
    }                                                                                                         \
  }
OR_EWC(or_arith_ssub_const, ssub)
OR_EWC(or_arith_sadd_const, sadd)

/* ------------------------------------------------------------------------- */
/* extrema chain, src/geodesy/extrema.cpp:27-114                              */
/* ------------------------------------------------------------------------- */
int or_h_maxima(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se)
{
  int64_t n = numel(rank, shape);
  void *tmp = malloc((size_t)n * tsize(t) + 1);
  if (!tmp) return OR_E_NOMEM;
  int rc = or_arith_ssub_const(t, in, h, tmp, n);
  if (rc == OR_OK) rc = or_reconstruct(t, 1, tmp, in, out, rank, shape, se);
  free(tmp);
  return rc;
}

int or_h_minima(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se)
{
  int64_t n = numel(rank, shape);
  void *tmp = malloc((size_t)n * tsize(t) + 1);
  if (!tmp) return OR_E_NOMEM;
  int rc = or_arith_sadd_const(t, in, h, tmp, n);
  if (rc == OR_OK) rc = or_reconstruct(t, 0, tmp, in, out, rank, shape, se);
  free(tmp);
  return rc;
}

int or_h_convex(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se)
{
  int64_t n = numel(rank, shape);
  void *tmp = malloc((size_t)n * tsize(t) + 1);
  if (!tmp) return OR_E_NOMEM;
  int rc = or_h_maxima(t, in, h, tmp, rank, shape, se);
  if (rc == OR_OK) rc = or_arith_ssub_image(t, in, tmp, out, n);
  free(tmp);
  return rc;
}

int or_h_concave(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se)
{
  int64_t n = numel(rank, shape);
  void *tmp = malloc((size_t)n * tsize(t) + 1);
  if (!tmp) return OR_E_NOMEM;
  int rc = or_h_minima(t, in, h, tmp, rank, shape, se);
  if (rc == OR_OK) rc = or_arith_ssub_image(t, tmp, in, out, n);
  free(tmp);
  return rc;
}

/* CompareImage(img, op, comp, valtrue, valfalse, out) with per-pixel operands,
 * specialised to the two calls made by dynamic_pseudo_* (extrema.cpp:90-93,110-113) */
#define OR_DPO(S, T, EXTREME)                                                                                 \
  static void dpo_marker_##S(const T *ext, const T *tmp, const T *img, T *mask, T *otmp, int64_t n)           \
  {                                                                                                           \
    for (int64_t i = 0; i < n; i++) mask[i] = (ext[i] == tmp[i]) ? img[i] : (EXTREME);                        \
    for (int64_t i = 0; i < n; i++) otmp[i] = (ext[i] != (EXTREME)) ? mask[i] : (EXTREME);                    \
  }
OR_DPO(lo_u8, uint8_t, 0)
OR_DPO(lo_u16, uint16_t, 0)
OR_DPO(lo_i32, int32_t, INT32_MIN)
OR_DPO(lo_i64, int64_t, INT64_MIN)
OR_DPO(lo_f32, float, -FLT_MAX)
OR_DPO(lo_f64, double, -DBL_MAX)
OR_DPO(hi_u8, uint8_t, UINT8_MAX)
OR_DPO(hi_u16, uint16_t, UINT16_MAX)
OR_DPO(hi_i32, int32_t, INT32_MAX)
OR_DPO(hi_i64, int64_t, INT64_MAX)
OR_DPO(hi_f32, float, FLT_MAX)
OR_DPO(hi_f64, double, DBL_MAX)

static void dpo_marker(int t, int opening, const void *ext, const void *tmp, const void *img, void *mask,
                       void *otmp, int64_t n)
{
#define DPO_CASE(code, S, T)                                                                                  \
  case code:                                                                                                  \
    if (opening) dpo_marker_lo_##S((const T *)ext, (const T *)tmp, (const T *)img, (T *)mask, (T *)otmp, n);  \
    else dpo_marker_hi_##S((const T *)ext, (const T *)tmp, (const T *)img, (T *)mask, (T *)otmp, n);          \
    break;
  switch (t) {
    DPO_CASE(OR_T_U8, u8, uint8_t)
    DPO_CASE(OR_T_U16, u16, uint16_t)
    DPO_CASE(OR_T_I32, i32, int32_t)
    DPO_CASE(OR_T_I64, i64, int64_t)
    DPO_CASE(OR_T_F32, f32, float)
    DPO_CASE(OR_T_F64, f64, double)
  default: break;
  }
#undef DPO_CASE
}

static int dynamic_pseudo(int t, int opening, const void *in, double h, void *out, int rank, const int64_t *shape,
                          int se)
{
  int64_t n = numel(rank, shape);
  size_t b = (size_t)n * tsize(t) + 1;
  void *tmp = malloc(b), *ext = malloc(b), *mask = malloc(b), *tmp2 = malloc(b);
  int rc = OR_E_NOMEM;
  if (tmp && ext && mask && tmp2) {
    rc = opening ? or_arith_ssub_const(t, in, h, tmp, n) : or_arith_sadd_const(t, in, h, tmp, n);
    if (rc == OR_OK) rc = or_reconstruct(t, opening ? 1 : 0, tmp, in, ext, rank, shape, se);
    if (rc == OR_OK) {
      dpo_marker(t, opening, ext, tmp, in, mask, tmp2, n);
      rc = or_reconstruct(t, opening ? 1 : 0, tmp2, in, out, rank, shape, se);
    }
  }
  free(tmp); free(ext); free(mask); free(tmp2);
  return rc;
}

int or_dynamic_pseudo_opening(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se)
{
  return dynamic_pseudo(t, 1, in, h, out, rank, shape, se);
}

int or_dynamic_pseudo_closing(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se)
{
  return dynamic_pseudo(t, 0, in, h, out, rank, shape, se);
}

/* ------------------------------------------------------------------------- */
/* Connected-component labelling (label_t.hpp t_label_helper, see oracle.h)   */
/* ------------------------------------------------------------------------- */
static int64_t lab_find(int64_t *par, int64_t x)
{
  int64_t r = x;
  while (par[r] != r) r = par[r];
  while (par[x] != r) { int64_t n = par[x]; par[x] = r; x = n; }
  return r;
}
/* disjoint_min_sets::union_sets: the smaller root becomes the parent */
static int64_t lab_union(int64_t *par, int64_t a, int64_t b)
{
  a = lab_find(par, a);
  b = lab_find(par, b);
  if (a < b) { par[b] = a; return a; }
  if (b < a) { par[a] = b; return b; }
  return a;
}
/* SE neighbours that precede the centre in raster order (split_coordinates_upper_lower) */
static int lab_upper(int se, int rank, int64_t *up)
{
  int64_t offs[27 * 3];
  int serank = 0, n = or_se_offsets(se, offs, &serank), m = 0;
  if (n < 0) return OR_E_SE;
  if (serank != rank) return OR_E_RANK;
  for (int k = 0; k < n; k++) {
    int sign = 0;
    for (int a = 0; a < rank && !sign; a++) sign = offs[k * rank + a] < 0 ? -1 : offs[k * rank + a] > 0 ? 1 : 0;
    if (sign < 0) { memcpy(up + m * rank, offs + k * rank, sizeof(int64_t) * (size_t)rank); m++; }
  }
  return m;
}

#define OR_LABEL_DEFINE(S, T, ACCEPT, POS)                                                                    \
  static int label_##S(int flat, const T *in, int64_t *out, int rank, const int64_t *shape, int se,           \
                       int64_t *count)                                                                        \
  {                                                                                                           \
    if (rank < 1 || rank > 3) return OR_E_RANK;                                                               \
    int64_t up[13 * 3], p[3], q;                                                                              \
    const int nu = lab_upper(se, rank, up);                                                                   \
    if (nu < 0) return nu;                                                                                    \
    const int64_t N = numel(rank, shape);                                                                     \
    int64_t *par = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));                              \
    if (!par) return OR_E_NOMEM;                                                                              \
    int64_t nlab = 0;                                                                                         \
    for (int64_t i = 0; i < N; i++) out[i] = 0;                                                               \
    for (int64_t i = 0; i < N; i++) {                                                                         \
      const T v = in[i];                                                                                      \
      if (!flat && !(ACCEPT(v))) continue;                                                                    \
      unravel(i, rank, shape, p);                                                                             \
      int64_t label = -1;                                                                                     \
      for (int k = 0; k < nu; k++) {                                                                          \
        if (!shifted(p, up + k * rank, rank, shape, &q)) continue;                                            \
        const T w = in[q];                                                                                    \
        const int rel = flat ? (w == v) : (POS(w) && (ACCEPT(w)));                                            \
        if (!rel) continue;                                                                                   \
        if (label == -1 || label == out[q]) label = out[q];                                                   \
        else label = lab_union(par, label, out[q]);                                                           \
      }                                                                                                       \
      if (label == -1) { par[nlab] = nlab; label = nlab++; }                                                  \
      out[i] = label;                                                                                         \
    }                                                                                                         \
    /* disjoint_min_sets::min_label: roots numbered 1.. in creation order */                                  \
    int64_t *num = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nlab > 0 ? nlab : 1));                       \
    if (!num) { free(par); return OR_E_NOMEM; }                                                               \
    int64_t cnt = 0;                                                                                          \
    for (int64_t l = 0; l < nlab; l++) { if (par[l] == l) cnt++; num[l] = cnt; }                              \
    for (int64_t i = 0; i < N; i++)                                                                           \
      if (flat || (ACCEPT(in[i]))) out[i] = num[lab_find(par, out[i])];                                       \
    *count = cnt;                                                                                             \
    free(num);                                                                                                \
    free(par);                                                                                                \
    return OR_OK;                                                                                             \
  }

#define LAB_NZ(v) ((v) != 0)
#define LAB_NZF(v) (fabsf(v) > 1E-10F)
#define LAB_NZD(v) (fabs(v) > 1E-10)
#define LAB_POS(v) ((v) > 0)
OR_LABEL_DEFINE(u8, uint8_t, LAB_NZ, LAB_POS)
OR_LABEL_DEFINE(u16, uint16_t, LAB_NZ, LAB_POS)
OR_LABEL_DEFINE(i32, int32_t, LAB_NZ, LAB_POS)
OR_LABEL_DEFINE(i64, int64_t, LAB_NZ, LAB_POS)
OR_LABEL_DEFINE(f32, float, LAB_NZF, LAB_POS)
OR_LABEL_DEFINE(f64, double, LAB_NZD, LAB_POS)

int or_label(int t, int flat, const void *in, int64_t *out, int rank, const int64_t *shape, int se, int64_t *count)
{
#define C(S) label_##S(flat, (const T *)in, out, rank, shape, se, count)
  OR_DISPATCH(t, C)
#undef C
}
