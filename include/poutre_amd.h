/*
 * poutre_amd.h -- C ABI of the MI355X (gfx950) morphology / geodesy library.
 *
 * Drop-in boundary for Poutre2's hot path.  Each entry point replaces one reference
 * C++ API function (cited below, paths relative to the Poutre2 source root); the
 * C++ shim (poutre_amd.hpp) and the Python bindings (poutre2_amd) both call these.
 *
 * Conventions
 *  - Images are described by pm_image: a DEVICE pointer (HBM, from hipMalloc or a
 *    torch tensor), the reference PType bit value, rank 1..3 and the shape in the
 *    reference's order ({W}, {H, W}, {D, H, W}; last axis fastest, dense row-major).
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are stream-ordered:
 *    they enqueue work and return without synchronising, except the geodesic
 *    fixed-point calls, which read a convergence flag back each global round.
 *  - Return value: PM_OK (0) or a negative pm_status; pm_last_error() then returns
 *    a thread-local message.  Checks happen in the reference's order: sizes, types,
 *    aliasing, rank, SE (src/low_level_morpho/ero_dil_static.cpp:229-236).
 *  - Temporaries come from the stream-ordered HIP allocator; the caller never
 *    allocates scratch.
 */
#ifndef POUTRE_AMD_H
#define POUTRE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PType bit values, include/poutre/base/types.hpp:517-527.  PM_PTYPE_U16 has no
 * reference counterpart (uint16 is an extension of this library). */
enum pm_ptype {
  PM_PTYPE_U8 = 1 << 2,
  PM_PTYPE_I32 = 1 << 3,
  PM_PTYPE_F32 = 1 << 4,
  PM_PTYPE_I64 = 1 << 5,
  PM_PTYPE_D64 = 1 << 6,
  PM_PTYPE_U16 = 1 << 7,
};

/* se::Common_NL_SE values, include/poutre/structuring_element/se_types_and_tags.hpp:45-57 */
enum pm_common_se {
  PM_SE_UNDEF = 0,
  PM_SE_SEGMENT_X_1D = 1,
  PM_SE_SQUARE_2D = 2,
  PM_SE_CROSS_2D = 3,
  PM_SE_SEGMENT_X_2D = 4,
  PM_SE_SEGMENT_Y_2D = 5,
  PM_SE_SEGMENT_X_3D = 6,
  PM_SE_SEGMENT_Y_3D = 7,
  PM_SE_SEGMENT_Z_3D = 8,
  PM_SE_CROSS_3D = 9,
  PM_SE_SQUARE_3D = 10,
};

/* se::Compound_NL_SE, se_types_and_tags.hpp:59-63 */
enum pm_compound_se { PM_COMPOUND_UNDEF = 0, PM_COMPOUND_OCTAGON = 1, PM_COMPOUND_RHOMBICUBOCTAHEDRON = 2 };

/* geo::reconstruction_type, include/poutre/geodesy/mreconstruct.hpp:30-34 */
enum pm_reconstruction_type { PM_REC_ERODE = 0, PM_REC_DILATE = 1 };

enum pm_status {
  PM_OK = 0,
  PM_E_SIZE = -1,        /* AssertSizesCompatible, image_interface.hpp:95 */
  PM_E_PTYPE = -2,       /* AssertAsTypesCompatible / unsupported PType */
  PM_E_ALIAS = -3,       /* AssertImagesAreDifferent */
  PM_E_RANK = -4,        /* unsupported rank or SE/rank mismatch */
  PM_E_SE = -5,          /* unknown structuring element */
  PM_E_ARG = -6,         /* iter < 0, bad offsets, NULL pointers ... */
  PM_E_HIP = -7,         /* HIP runtime error */
  PM_E_UNSUPPORTED = -8, /* the reference throws for this combination too */
  PM_E_NOCONV = -9,      /* geodesic iteration cap reached */
};

typedef struct pm_image {
  void *data;
  int32_t ptype;
  int32_t rank;
  int64_t shape[3];
} pm_image;

/* thread-local message of the last failing call on this thread ("" if none) */
const char *pm_last_error(void);
/* library version string */
const char *pm_version(void);

/* ---- low-level morphology: poutre::Erode/Dilate by Common_NL_SE with iterations
 * include/poutre/low_level_morpho/ero_dil.hpp:32,35 -> src/low_level_morpho/ero_dil_static.cpp:229-305 */
int pm_erode_static(const pm_image *in, int se, int iter, pm_image *out, void *stream);
int pm_dilate_static(const pm_image *in, int se, int iter, pm_image *out, void *stream);

/* poutre::Erode/Dilate by Compound_NL_SE, ero_dil.hpp:38,41 -> src/low_level_morpho/ero_dil_compound_static.cpp:58-188 */
int pm_erode_compound(const pm_image *in, int compound, int size, pm_image *out, void *stream);
int pm_dilate_compound(const pm_image *in, int compound, int size, pm_image *out, void *stream);

/* poutre::Erode/Dilate by a runtime neighbour list (se::neighbor_list_t<R>), ero_dil.hpp:44,47 ->
 * src/low_level_morpho/ero_dil.cpp:59-243.  `offsets` holds n rows of `rank` int64 {dz,dy,dx}. */
int pm_erode_nl(const pm_image *in, const int64_t *offsets, int n, pm_image *out, void *stream);
int pm_dilate_nl(const pm_image *in, const int64_t *offsets, int n, pm_image *out, void *stream);

/* poutre::ErodeX/DilateX/ErodeY/DilateY, include/poutre/low_level_morpho/ero_dil_line.hpp:32-47 ->
 * src/low_level_morpho/ero_dil_line.cpp:80-295.  X: rank 1 and 2; Y: rank 2.  The reference's
 * Y pass only accepts square images (its transposed temporary keeps the input shape,
 * ero_dil_line_se_t.hpp:258-259); these entry points accept any H, W and the C++ shim
 * poutre::ErodeY/DilateY restores the reference's error for H != W. */
int pm_erode_x(const pm_image *in, int64_t half, pm_image *out, void *stream);
int pm_dilate_x(const pm_image *in, int64_t half, pm_image *out, void *stream);
int pm_erode_y(const pm_image *in, int64_t half, pm_image *out, void *stream);
int pm_dilate_y(const pm_image *in, int64_t half, pm_image *out, void *stream);

/* Flat erosion/dilation by a (2ry+1) x (2rx+1) rectangle (rank 2, or rank 3 slice-wise), the
 * separable building block of SESquare2D iterations (no reference counterpart of its own:
 * equals DilateX then DilateY, ero_dil_line_se_t.hpp:192-280). */
int pm_erode_rect(const pm_image *in, int64_t ry, int64_t rx, pm_image *out, void *stream);
int pm_dilate_rect(const pm_image *in, int64_t ry, int64_t rx, pm_image *out, void *stream);

/* Opening / closing / morphological gradient by a Common_NL_SE iterated `iter` times.
 * No reference symbol (only mentioned at include/poutre/low_level_morpho/low_level_morpho.hpp:15):
 * defined as Dilate(Erode(f)), Erode(Dilate(f)) and the saturated difference Dilate(f) - Erode(f),
 * each operand computed with pm_*_static semantics. */
int pm_opening(const pm_image *in, int se, int iter, pm_image *out, void *stream);
int pm_closing(const pm_image *in, int se, int iter, pm_image *out, void *stream);
int pm_gradient(const pm_image *in, int se, int iter, pm_image *out, void *stream);

/* ---- geodesy
 * poutre::geo::Reconstruction, include/poutre/geodesy/mreconstruct.hpp:37-41 ->
 * src/geodesy/mreconstruct.cpp:48-132 (mreconstruct_t.hpp:75-512).  Returns after
 * synchronising the stream (it checks convergence on the host).  For uint8 2D images its
 * device scratch is kept between calls (up to 8 idle buffers per process, thread-safe). */
int pm_reconstruct(int rec_type, const pm_image *marker, const pm_image *mask, int se, pm_image *out,
                   void *stream);
/* ---- testing hooks (no reference counterpart): force the error paths from the host, so that
 * tests reach them with every kernel exiting normally.
 * pm_debug_set_max_rounds(n): reconstructions (and the geodesy built on them) called afterwards
 *   stop with PM_E_NOCONV after n global rounds when not converged; 0 restores the default caps.
 * pm_debug_fail_allocations(n): the next n scratch allocations fail with PM_E_HIP, as on an
 *   out-of-memory device, without calling the allocator; 0 clears.
 * Both settings are process-wide. */
int pm_debug_set_max_rounds(int n);
int pm_debug_fail_allocations(int n);

/* pm_reconstruct for uint8 1D/2D images and 3x3-class SEs (square, cross, segments) that
 * allocates nothing.  No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is a 1 MiB array per GPU held until the process ends, so images are limited to
 * 196608 tiles of 62 x 62 pixels (about 755 Mpixel); larger ones return PM_E_ARG, other pixel
 * types or SEs PM_E_UNSUPPORTED.  Calls on one GPU share that array: calls that overlap (on
 * two streams, or from two threads) race and must be serialised by the caller. */
int pm_reconstruct_noalloc(int rec_type, const pm_image *marker, const pm_image *mask, int se, pm_image *out,
                           void *stream);
/* poutre::geo::low_leveling / high_leveling / leveling, include/poutre/geodesy/leveling.hpp:37-52 ->
 * src/geodesy/leveling.cpp:67-237 */
int pm_low_leveling(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream);
int pm_high_leveling(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream);
int pm_leveling(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream);
/* poutre::geo::h_minima / h_maxima / h_concave / h_convex / dynamic_pseudo_opening / _closing,
 * include/poutre/geodesy/extrema.hpp:33-64 -> src/geodesy/extrema.cpp:27-114.  `h` is the
 * ScalarTypeVariant value converted to double. */
int pm_h_maxima(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_h_minima(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_h_convex(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_h_concave(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_dynamic_pseudo_opening(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_dynamic_pseudo_closing(const pm_image *in, double h, int se, pm_image *out, void *stream);

/* ---- elementwise helpers on the path (include/poutre/pixel_processing/arith.hpp:31-53):
 * ArithSupImage / ArithInfImage / ArithSaturatedSubImage / ArithSaturatedAddImage /
 * ArithSaturatedSub/AddConstant (accumulated in a wider type, clamped at the type's minimum /
 * maximum), ArithInvertImage (-x; integers modulo 2^bits) */
int pm_arith_sup(const pm_image *a, const pm_image *b, pm_image *out, void *stream);
int pm_arith_inf(const pm_image *a, const pm_image *b, pm_image *out, void *stream);
int pm_arith_ssub_image(const pm_image *a, const pm_image *b, pm_image *out, void *stream);
int pm_arith_sadd_image(const pm_image *a, const pm_image *b, pm_image *out, void *stream);
int pm_arith_ssub_const(const pm_image *in, double c, pm_image *out, void *stream);
int pm_arith_sadd_const(const pm_image *in, double c, pm_image *out, void *stream);
int pm_arith_invert(const pm_image *in, pm_image *out, void *stream);

/* ---- ConvertInto / CopyInto (include/poutre/pixel_processing/copy_convert.hpp:49-52): out =
 * in converted pixel by pixel to out's PType (a C cast, as the reference's static_cast; float to
 * integer is defined only for values the integer type holds), a copy when the PTypes agree.
 * Same shape; out may be in only when the PTypes agree. */
int pm_convert_into(const pm_image *in, pm_image *out, void *stream);

/* ---- poutre::CompareImage, all eight overloads (include/poutre/pixel_processing/compare.hpp:31-84,
 * src/pixel_processing/compare.cpp:601-): out = (in OP comp) ? vtrue : vfalse, pixel by pixel,
 * the input pixel on the left.  Each operand is an image (image != NULL, same shape as in) or
 * the scalar value.  comp has in's PType; vtrue, vfalse and out share a PType that may differ
 * from in's.  Scalars are converted to those PTypes (integers: truncated toward zero and
 * clamped to the type's range, NaN -> 0).  op: one CompOpType bit (types.hpp:443-453).  out
 * may be one of the inputs. */
enum pm_compare_op {
  PM_CMP_EQUAL = 1 << 0,
  PM_CMP_DIFF = 1 << 1,
  PM_CMP_SUP = 1 << 2,
  PM_CMP_SUP_EQUAL = 1 << 3,
  PM_CMP_INF = 1 << 4,
  PM_CMP_INF_EQUAL = 1 << 5
};
typedef struct pm_operand {
  const pm_image *image; /* NULL: the operand is `value` */
  double value;
} pm_operand;
int pm_compare(const pm_image *in, int op, const pm_operand *comp, const pm_operand *vtrue, const pm_operand *vfalse,
               pm_image *out, void *stream);

/* ---- differentiable float32 dilation / erosion (no reference counterpart; extends the flat
 * Dilate/Erode of include/poutre/low_level_morpho/ero_dil.hpp:32-47 to a structuring function).
 * x, y, gy, gx: B contiguous float32 images of H x W (device pointers).  w: K x K float32
 * structuring function (row-major, tap (dy,dx) at [(dy+R)*K + dx+R]) or NULL for the flat
 * K x K square.  K in {3,5,7,9}.  dilate != 0: y[q] = max_t x[q+t] + w[t]; else
 * y[q] = min_t x[q+t] - w[t]; taps outside the image are skipped.  The backward routes gy[q]
 * to the first maximiser (minimiser) in raster order of t and, if w != NULL, writes the
 * gradient of w (K*K floats) to gw.  Deterministic (no atomics). */
int pm_maxplus_fwd(const float *x, int64_t B, int64_t H, int64_t W, const float *w, int K, int dilate, float *y,
                   void *stream);
int pm_maxplus_bwd(const float *x, const float *gy, int64_t B, int64_t H, int64_t W, const float *w, int K,
                   int dilate, float *gx, float *gw, void *stream);
/* the same operator with the selected taps kept: the forward also writes, per output, the
 * raster index dy*K + dx of its selected tap (B*H*W bytes); the backward reads those instead of
 * x and w (with_w != 0: also write gw).  For training, where the forward's input need not be
 * kept: 1 byte per pixel instead of 4, and no tap re-selection in the backward.  gx is summed
 * over tap rows, then tap columns (gw: per-wave partials, then over waves), in a fixed order:
 * deterministic, and equal to pm_maxplus_bwd's up to float rounding of the order. */
int pm_maxplus_fwd_sel(const float *x, int64_t B, int64_t H, int64_t W, const float *w, int K, int dilate, float *y,
                       uint8_t *sel, void *stream);
int pm_maxplus_bwd_sel(const uint8_t *sel, const float *gy, int64_t B, int64_t H, int64_t W, int K, int dilate,
                       int with_w, float *gx, float *gw, void *stream);

/* ---- connected-component labelling (include/poutre/label/).  out is an int64 (PM_PTYPE_I64)
 * image of the input's shape; background pixels get 0 and the components 1..count, numbered
 * by the raster position of their first pixel (label_t.hpp t_label_helper).  Input PTypes:
 * uint8, uint16, int32, int64, float32, float64; SEs of the input's rank.  *count (may be
 * NULL) receives the number of labels; reading it synchronises the stream. */
/* label_binary.hpp:35 -- foreground = value != 0 (|v| > 1e-10 for floats), linked to a
 * raster-earlier neighbour whose value is > 0 */
int pm_label_binary(const pm_image *in, int se, pm_image *out, int64_t *count, void *stream);
/* label_flat_zones.hpp:35 -- every pixel labelled, linked to equal-valued neighbours */
int pm_label_flat_zones(const pm_image *in, int se, pm_image *out, int64_t *count, void *stream);
/* label_extrema.hpp:34,39 -- label_binary of h_concave(in, 1) / h_convex(in, 1) */
int pm_label_minima(const pm_image *in, int se, pm_image *out, void *stream);
int pm_label_maxima(const pm_image *in, int se, pm_image *out, void *stream);

/* ---- batched launches (no reference counterpart: the reference processes one image per call).
 * pm_batch_dilate_open: for each i < n, dil_outs[i] = Dilate(ins[i], SESquare2D, radii[i]) and
 * open_outs[i] = Dilate(Erode(ins[i], SESquare2D, radii[i]), SESquare2D, radii[i]), all images
 * uint8 2D of any sizes, radii in [1, 32]: one persistent launch for all images with radii
 * up to 15 (dilation tiles and fused-opening tiles of every radius in one work list), two per
 * distinct radius above, whatever the number of images.
 * pm_batch_reconstruct: n reconstructions (geo::Reconstruction semantics, uint8 2D, 3x3 SEs)
 * sharing one sequence of global rounds. */
int pm_batch_dilate_open(int n, const pm_image *ins, const int *radii, pm_image *dil_outs, pm_image *open_outs,
                         void *stream);
int pm_batch_reconstruct(int n, int rec_type, const pm_image *markers, const pm_image *masks, int se, pm_image *outs,
                         void *stream);

/* ---- batch plans (no reference counterpart): pm_batch_dilate_open captured once into a HIP
 * graph, with its scratch and descriptor table allocated at creation; a launch is one
 * hipGraphLaunch on the caller's stream, with no allocation and no per-image host work.  The
 * plan keeps the image pointers it was created with; launches on one stream are ordered,
 * launches of one plan on different streams at the same time race on its scratch. */
typedef struct pm_batch_plan pm_batch_plan;
int pm_batch_plan_create(int n, const pm_image *ins, const int *radii, pm_image *dil_outs, pm_image *open_outs,
                         pm_batch_plan **plan);
int pm_batch_plan_launch(pm_batch_plan *plan, void *stream);
int pm_batch_plan_destroy(pm_batch_plan *plan);

/* ---- autotuner (no reference counterpart).  The fused uint8 square kernel has three tile
 * widths (128, 192 and 256 columns) per radius; pm_autotune times each for every radius 1..32
 * on 8192 x 8192 images with 16-byte aligned rows and on 8192 x 8189 ones with unaligned rows
 * (device memory allocated and freed by the call; the stream is synchronised) and keeps the
 * fastest in a per-device table used by every later call.  pm_tuning_get returns the width in
 * use for (aligned != 0, r); pm_tuning_set sets it (128, 192, 256, or 0 = built-in rule). */
int pm_autotune(void *stream);
int pm_tuning_get(int aligned, int r);
int pm_tuning_set(int aligned, int r, int tile_width);

#ifdef __cplusplus
}
#endif
#endif /* POUTRE_AMD_H */
