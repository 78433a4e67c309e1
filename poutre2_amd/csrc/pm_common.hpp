// pm_common.hpp -- shared host/device helpers of the gfx950 morphology library:
// element traits (neutral values of dilation/erosion), the PType dispatch, argument
// validation in the reference's order, the thread-local error message and the
// stream-ordered scratch allocation.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <string>

#include "../../include/poutre_amd.h"

namespace pm {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
void clear_error();

struct Status {
  int code;
};

#define PM_FAIL(code, ...)                                                                                   \
  do {                                                                                                       \
    ::pm::set_error(__VA_ARGS__);                                                                            \
    return (code);                                                                                           \
  } while (0)

#define PM_HIP(expr)                                                                                         \
  do {                                                                                                       \
    hipError_t e_ = (expr);                                                                                  \
    if (e_ != hipSuccess) PM_FAIL(PM_E_HIP, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__,   \
                                  __LINE__, #expr);                                                          \
  } while (0)

#define PM_TRY(expr)                                                                                         \
  do {                                                                                                       \
    int rc_ = (expr);                                                                                        \
    if (rc_ != PM_OK) return rc_;                                                                            \
  } while (0)

// ---------------------------------------------------------------- element traits
// Neutral of dilation = numeric_limits<T>::lowest(), of erosion = max()
// (ero_dil_line_se_t.hpp:37,44; types_traits.hpp:114-263).  Float neutrals are +-FLT_MAX.
template <typename T> struct Lim;
template <> struct Lim<uint8_t> { static constexpr __host__ __device__ uint8_t lowest() { return 0; } static constexpr __host__ __device__ uint8_t highest() { return 255; } static constexpr __host__ __device__ uint8_t tmin() { return 0; } };
template <> struct Lim<uint16_t> { static constexpr __host__ __device__ uint16_t lowest() { return 0; } static constexpr __host__ __device__ uint16_t highest() { return 65535; } static constexpr __host__ __device__ uint16_t tmin() { return 0; } };
template <> struct Lim<int32_t> { static constexpr __host__ __device__ int32_t lowest() { return INT32_MIN; } static constexpr __host__ __device__ int32_t highest() { return INT32_MAX; } static constexpr __host__ __device__ int32_t tmin() { return INT32_MIN; } };
template <> struct Lim<int64_t> { static constexpr __host__ __device__ int64_t lowest() { return INT64_MIN; } static constexpr __host__ __device__ int64_t highest() { return INT64_MAX; } static constexpr __host__ __device__ int64_t tmin() { return INT64_MIN; } };
template <> struct Lim<float> { static constexpr __host__ __device__ float lowest() { return -FLT_MAX; } static constexpr __host__ __device__ float highest() { return FLT_MAX; } static constexpr __host__ __device__ float tmin() { return FLT_MIN; } };
template <> struct Lim<double> { static constexpr __host__ __device__ double lowest() { return -DBL_MAX; } static constexpr __host__ __device__ double highest() { return DBL_MAX; } static constexpr __host__ __device__ double tmin() { return DBL_MIN; } };

// The lattice operation of a dilation (sup) or an erosion (inf).  Written as a select
// so that integer and float types share one definition (`a > b ? a : b` is the
// reference's scalar form, arith_op_t.hpp:363,414).
template <typename T, bool DIL> struct Lat {
  static __host__ __device__ __forceinline__ T op(T a, T b) { return DIL ? (a > b ? a : b) : (a < b ? a : b); }
  static __host__ __device__ __forceinline__ T neutral() { return DIL ? Lim<T>::lowest() : Lim<T>::highest(); }
  // "a is strictly better than b" in the direction of the operation
  static __host__ __device__ __forceinline__ bool better(T a, T b) { return DIL ? (a > b) : (a < b); }
};

inline size_t ptype_size(int ptype)
{
  switch (ptype) {
  case PM_PTYPE_U8: return 1;
  case PM_PTYPE_U16: return 2;
  case PM_PTYPE_I32: return 4;
  case PM_PTYPE_F32: return 4;
  case PM_PTYPE_I64: return 8;
  case PM_PTYPE_D64: return 8;
  default: return 0;
  }
}

template <typename T> struct Tag { using type = T; };

// Calls f(Tag<T>{}) for the image's element type (f is typically a generic lambda).
template <typename F> int dispatch_ptype(int ptype, F &&f)
{
  switch (ptype) {
  case PM_PTYPE_U8: return f(Tag<uint8_t>{});
  case PM_PTYPE_U16: return f(Tag<uint16_t>{});
  case PM_PTYPE_I32: return f(Tag<int32_t>{});
  case PM_PTYPE_F32: return f(Tag<float>{});
  case PM_PTYPE_I64: return f(Tag<int64_t>{});
  case PM_PTYPE_D64: return f(Tag<double>{});
  default: PM_FAIL(PM_E_PTYPE, "Unsupported PType %d", ptype);
  }
}

// ---------------------------------------------------------------- images
inline int64_t numel(const pm_image &im)
{
  int64_t n = 1;
  for (int a = 0; a < im.rank; a++) n *= im.shape[a];
  return n;
}

// Logical 3D view {D, H, W} of a rank-1..3 image (rank 1: D = H = 1).
struct Dims {
  int64_t D, H, W;
};
inline Dims dims3(const pm_image &im)
{
  if (im.rank == 1) return {1, 1, im.shape[0]};
  if (im.rank == 2) return {1, im.shape[0], im.shape[1]};
  return {im.shape[0], im.shape[1], im.shape[2]};
}

// AssertSizesCompatible / AssertAsTypesCompatible / AssertImagesAreDifferent
// (include/poutre/base/image_interface.hpp:95-102), then rank 1..3.
int check_pair(const pm_image *a, const pm_image *b, const char *what);
bool same_shape(const pm_image *a, const pm_image *b);
// one descriptor: rank 1..3, a known PType, extents >= 0, data for a non-empty image
int check_image(const pm_image *im, const char *what);
int check_unary(const pm_image *in, const pm_image *out, const char *what);
int check_binary(const pm_image *a, const pm_image *b, const pm_image *out, const char *what);

// ---------------------------------------------------------------- scratch
// Stream-ordered device scratch: allocated with hipMallocAsync on the call's stream and
// released with hipFreeAsync on the same stream, so that it never races with the kernels
// of that stream that use it.
struct Scratch {
  void *ptr = nullptr;
  hipStream_t stream = nullptr;
  Scratch() = default;
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
  int alloc(size_t bytes, hipStream_t s);
  ~Scratch();
};

// 16 and 4 bytes at any byte address.  gfx950 under ROCm runs with unaligned memory
// access enabled (the compiler emits these as single global_load/store_dwordx4 / _dwordx3 /
// _dword / _short instructions at the unaligned address); each access reads or writes exactly its own
// bytes, so no neighbouring byte of another tile or row is touched.
typedef unsigned int u32x4_a1 __attribute__((ext_vector_type(4), aligned(1)));
typedef unsigned int u32_a1 __attribute__((aligned(1)));

// 16 bytes at any address p, all of them inside the image
__device__ __forceinline__ uint4 load16_any(const uint8_t *p)
{
  const u32x4_a1 v = *reinterpret_cast<const u32x4_a1 *>(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store16_any(uint8_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
  const u32x4_a1 v = {a, b, c, d};
  *reinterpret_cast<u32x4_a1 *>(p) = v;
}
__device__ __forceinline__ void store4_any(uint8_t *p, uint32_t a) { *reinterpret_cast<u32_a1 *>(p) = a; }
typedef unsigned int u32x3_a1 __attribute__((ext_vector_type(3), aligned(1)));
typedef unsigned short u16_a1 __attribute__((aligned(1)));
__device__ __forceinline__ void store12_any(uint8_t *p, uint32_t a, uint32_t b, uint32_t c)
{
  const u32x3_a1 v = {a, b, c};
  *reinterpret_cast<u32x3_a1 *>(p) = v;
}
__device__ __forceinline__ void store2_any(uint8_t *p, uint32_t a) { *reinterpret_cast<u16_a1 *>(p) = (unsigned short)a; }

// testing hooks (pm_debug_set_max_rounds / pm_debug_fail_allocations): the round cap a
// reconstruction uses (its own cap when no hook is set), and whether the next scratch
// allocation must fail
int round_cap(int own_cap);
bool take_forced_alloc_failure();

// Pinned host memory for PM_COUNTERS round counters and one more int (index PM_COUNTERS: a
// guard flag read with them), allocated once per host thread (the reconstructions read their
// convergence counters into it once per batch of rounds).
constexpr int PM_COUNTERS = 64;
int pinned_counters(int **h);

// A roctx range around each C ABI call, named after the entry point, so that kernel traces
// (rocprofv3 --marker-trace) attribute kernels to API calls; without a tracer attached a
// push / pop is a call into an empty table.
struct TraceRange {
  explicit TraceRange(const char *name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};
#define PM_TRACE() ::pm::TraceRange pm_trace_range_(__func__)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// A scalar operand (h of the extrema, the constant of the saturated arithmetic) converted to
// the pixel type T.  The reference takes it as the ScalarTypeVariant alternative of the
// image's type (std::get, arith.cpp:38,55), so it is always representable there; here an
// integer T needs a whole number inside T's range and a float T a value inside its range
// (or an infinity / NaN), anything else is PM_E_ARG rather than a silent wrap.
template <typename T> int scalar_as(double v, T *out, const char *what)
{
  if constexpr (std::numeric_limits<T>::is_integer) {
    const double lo = (double)std::numeric_limits<T>::lowest();
    const double hi_excl = std::ldexp(1.0, std::numeric_limits<T>::digits);  // max + 1, exact
    if (!(v == std::floor(v)) || v < lo || v >= hi_excl)
      PM_FAIL(PM_E_ARG, "%s: value %.17g is not a value of the image's pixel type", what, v);
  } else {
    if (std::isfinite(v) && std::fabs(v) > (double)std::numeric_limits<T>::max())
      PM_FAIL(PM_E_ARG, "%s: value %.17g is outside the image's pixel type", what, v);
  }
  *out = static_cast<T>(v);
  return PM_OK;
}

constexpr int PM_MAX_DEVICES = 64;

// Fill n bytes at device pointer p with v (a kernel; safe under graph capture).
int fill_bytes(void *p, uint8_t v, size_t n, hipStream_t s);

inline unsigned grid1d(int64_t n, int block, int64_t cap = 1 << 20)
{
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace pm
