/*
 * oracle.h -- CPU restatement of Poutre2's morphology / geodesy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is compared
 * against (tests/, __graft_entry__.smoke(), and bench.py's cpu_baseline leg).  The
 * product path (poutre2_amd) never links, loads or calls it.
 *
 * Every function restates one reference algorithm; the reference file:line it
 * follows is cited next to it (paths relative to the Poutre2 source root).
 * Images are dense row-major arrays, rank 1..3, shape[0] slowest / shape[rank-1]
 * fastest (include/poutre/base/details/data_structures/image_t.hpp:79-85,
 * include/poutre/base/details/data_structures/array_view.hpp:751-766).
 *
 * Element type codes (argument `t`): see OR_T_* below.  All functions return 0 on
 * success and a negative code on a contract violation the reference throws on.
 *
 * Pinning status of each function (what checks it besides its own formula):
 *   or_line_x / or_line_y      KATs test/low_level_morpho/ero_dil_line_se.cpp + brute-force window
 *   or_static                  KATs ero_dil_static_se_t.cpp + ero_dil_runtime_se.cpp + or_nl
 *   or_static_iter             composition identity vs or_line (square) and or_nl (diamond)
 *   or_nl                      KATs ero_dil_runtime_se.cpp
 *   or_compound                one or_nl dilation by the explicitly enumerated Minkowski sum
 *                              diamond(n-m) + square(m) (3D: octahedron + cube), m from a
 *                              hand-written table of the round-half-up rule, n = 0..12; errors
 *   or_reconstruct             KATs test/geodesy/mreconstruct.cpp + or_reconstruct_bf
 *   or_reconstruct_bf          KATs + invariants (marker^mask <= R <= mask, idempotence)
 *   or_low/high_leveling       KATs test/geodesy/leveling.cpp + closed forms via or_reconstruct
 *   or_arith_*                 closed forms (saturating arithmetic) + KATs through h_* chain
 *   extrema chain              KATs test/geodesy/extrema.cpp
 *   3D SESquare3D / SECross3D  or_static computes the mathematical definition (= the runtime
 *                              neighbour-list path, ero_dil_runtime_nl_se_t.hpp:48-97), pinned by
 *                              the z-constant 3D KATs and the brute-force 3D window.  The
 *                              reference's static slice composites (ero_dil_static_se_t.hpp:
 *                              761-915) return something else on z-varying input; a literal
 *                              restatement of them (tests/test_static3d_composite.py) reproduces
 *                              the reference's 3D KATs and shows where they diverge.  The
 *                              library keeps the definition, as or_static does.
 */
#ifndef POUTRE_ORACLE_H
#define POUTRE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OR_T_U8 = 0,
  OR_T_U16 = 1,
  OR_T_I32 = 2,
  OR_T_I64 = 3,
  OR_T_F32 = 4,
  OR_T_F64 = 5,
};

/* Common_NL_SE values, include/poutre/structuring_element/se_types_and_tags.hpp:45-57 */
enum {
  OR_SE_UNDEF = 0,
  OR_SE_SEGX1D = 1,
  OR_SE_SQUARE2D = 2,
  OR_SE_CROSS2D = 3,
  OR_SE_SEGX2D = 4,
  OR_SE_SEGY2D = 5,
  OR_SE_SEGX3D = 6,
  OR_SE_SEGY3D = 7,
  OR_SE_SEGZ3D = 8,
  OR_SE_CROSS3D = 9,
  OR_SE_SQUARE3D = 10,
};

/* number of taps / rank of a Common_NL_SE; fills offs (up to 27*3 int64, centre first) */
int or_se_offsets(int se, int64_t *offs, int *rank);

/* DilateX/ErodeX: van Herk line pass, ero_dil_line_se_t.hpp:49-190 */
int or_line_x(int t, const void *in, void *out, int rank, const int64_t *shape, int64_t r, int dilate);
/* DilateY/ErodeY: transpose, X pass, transpose back; square 2D only, ero_dil_line_se_t.hpp:248-280 */
int or_line_y(int t, const void *in, void *out, int rank, const int64_t *shape, int64_t r, int dilate);
/* t_transpose, transpose_op_t.hpp:45-70 (in shape (H,W), out shape (W,H)) */
int or_transpose(int t, const void *in, void *out, const int64_t *shape);

/* One application of a Common_NL_SE: max/min over in-bounds taps (ero_dil_static_se_t.hpp:917-1152) */
int or_static(int t, const void *in, void *out, int rank, const int64_t *shape, int se, int dilate);
/* iter driver, src/low_level_morpho/ero_dil_static.cpp:229-305 */
int or_static_iter(int t, const void *in, void *out, int rank, const int64_t *shape, int se, int iter, int dilate);
/* runtime neighbour list brute force, ero_dil_runtime_nl_se_t.hpp:48-97 */
int or_nl(int t, const void *in, void *out, int rank, const int64_t *shape, const int64_t *offs, int n, int dilate);
/* compound SE (octagon / rhombicuboctahedron), ero_dil_compound_static_se_t.hpp:33-232 */
int or_compound(int t, const void *in, void *out, int rank, const int64_t *shape, int compound, int size, int dilate);

/* geo::Reconstruction hybrid algorithm, mreconstruct_t.hpp:39-195 (rec_type: 0 erode, 1 dilate) */
int or_reconstruct(int t, int rec_type, const void *marker, const void *mask, void *out, int rank,
                   const int64_t *shape, int se);
/* brute-force fixed point of the same reconstruction: x <- mask ^ delta(x) until stable */
int or_reconstruct_bf(int t, int rec_type, const void *marker, const void *mask, void *out, int rank,
                      const int64_t *shape, int se);

/* leveling_t.hpp:34-202 (priority queue keyed by pixel index) */
int or_low_leveling(int t, const void *ref, const void *marker, void *out, int rank, const int64_t *shape, int se);
int or_high_leveling(int t, const void *ref, const void *marker, void *out, int rank, const int64_t *shape, int se);
/* src/geodesy/leveling.cpp:231-237 */
int or_leveling(int t, const void *ref, const void *marker, void *out, int rank, const int64_t *shape, int se);

/* arith_op_t.hpp saturated ops (clamp at numeric_limits<T>::min()/max()) and sup/inf */
int or_arith_ssub_const(int t, const void *in, double c, void *out, int64_t n);
int or_arith_sadd_const(int t, const void *in, double c, void *out, int64_t n);
int or_arith_ssub_image(int t, const void *a, const void *b, void *out, int64_t n);
/* ArithSaturatedAddImage (arith.hpp:44; op_Saturated_Add arith_op_t.hpp:146-190) and
   ArithInvertImage (arith.hpp:33; op_Invert arith_op_t.hpp:40-57) */
int or_arith_sadd_image(int t, const void *a, const void *b, void *out, int64_t n);
int or_arith_invert(int t, const void *in, void *out, int64_t n);
int or_arith_sup(int t, const void *a, const void *b, void *out, int64_t n);
int or_arith_inf(int t, const void *a, const void *b, void *out, int64_t n);

/* src/geodesy/extrema.cpp:27-114 */
int or_h_maxima(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se);
int or_h_minima(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se);
int or_h_convex(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se);
int or_h_concave(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se);
int or_dynamic_pseudo_opening(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se);
int or_dynamic_pseudo_closing(int t, const void *in, double h, void *out, int rank, const int64_t *shape, int se);

/* Connected-component labelling, include/poutre/label/details/label_t.hpp t_label_helper:
 * one raster pass over the accepted pixels with a union-find of provisional labels (new
 * labels 0, 1, 2... in raster order, union keeps the smaller root), linking a pixel to its
 * raster-earlier ("upper") SE neighbours that satisfy the relation, then roots renumbered
 * 1..count in creation order; background 0.
 *  flat = 0: label_binary (src/label/label_binary.cpp): accept v != 0 (|v| > 1e-10 for
 *            float), relation: neighbour > 0 (label_t.hpp not_background / sup_zero);
 *  flat = 1: label_flat_zones (src/label/label_flat_zones.cpp): accept all, relation ==.
 * Positive float neighbours at or below the 1e-10 background threshold make the reference
 * read an unassigned provisional label; here they are not linked (not pinned by the
 * reference).  out is int64 (the reference's GINT64 output image). */
int or_label(int t, int flat, const void *in, int64_t *out, int rank, const int64_t *shape, int se, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
