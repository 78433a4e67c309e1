// arith.hip -- elementwise helpers on the path (include/poutre/pixel_processing/arith.hpp:31-53,
// arith_op_t.hpp:40-445, compare.hpp:31-84): sup/inf of two images, saturated addition and
// subtraction of images and of a constant (clamped at numeric_limits<T>::min()/max(),
// accumulated in a wider type), negation, the comparison-select of CompareImage, plus the
// opening / closing / gradient compositions.
#include <type_traits>

#include "arith.hpp"
#include "erodil.hpp"
#include "sq_open.hpp"
#include "sq_u8.hpp"

namespace pm {

// accumulation type of the saturated operations (TypeTraits<T>::accu_type)
template <typename T> struct Accu { using type = T; };
template <> struct Accu<uint8_t> { using type = int32_t; };
template <> struct Accu<uint16_t> { using type = int32_t; };
template <> struct Accu<int32_t> { using type = int64_t; };
template <> struct Accu<float> { using type = double; };

template <typename T> __device__ __forceinline__ T ssub_sat(T a, T b)
{
  using A = typename Accu<T>::type;
  const A r = (A)a - (A)b;
  return r < (A)Lim<T>::tmin() ? Lim<T>::tmin() : (T)r;
}
template <typename T> __device__ __forceinline__ T sadd_sat(T a, T b)
{
  using A = typename Accu<T>::type;
  const A r = (A)a + (A)b;
  return r > (A)Lim<T>::highest() ? Lim<T>::highest() : (T)r;
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void ew2_kernel(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ o,
                                                  int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const T x = a[i], y = b[i];
    T r;
    if (OP == EW_SUP) r = x > y ? x : y;
    else if (OP == EW_INF) r = x < y ? x : y;
    else if (OP == EW_SADD) r = sadd_sat<T>(x, y);
    else r = ssub_sat<T>(x, y);
    o[i] = r;
  }
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void ewc_kernel(const T *__restrict__ a, T c, T *__restrict__ o, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    o[i] = (OP == EWC_SSUB) ? ssub_sat<T>(a[i], c) : sadd_sat<T>(a[i], c);
}

// op_Invert (arith_op_t.hpp:40-57): -a in the pixel type (integers modulo 2^bits)
template <typename T> __device__ __forceinline__ T negate(T a)
{
  if constexpr (std::is_integral<T>::value) {
    using U = typename std::make_unsigned<T>::type;
    return (T)(U)(U(0) - (U)a);
  } else {
    return -a;
  }
}
template <typename T>
__global__ __launch_bounds__(256) void inv_kernel(const T *__restrict__ a, T *__restrict__ o, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) o[i] = negate<T>(a[i]);
}

// CompareImage (compare_op_t.hpp:40-165, select :305-351): o = (in OP comp) ? vtrue : vfalse,
// the input pixel on the left; an operand image pointer that is null selects its scalar.
// op is one CompOpType bit; ">=" is "> or ==" (the same for every value, NaN included).  The
// output may be one of the inputs (each pixel is read before it is written), so no restrict.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void cmp_kernel(const TI *in, int op, const TI *ci, TI cv, const TO *ti, TO tv,
                                                  const TO *fi, TO fv, TO *o, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const TI a = in[i], c = ci ? ci[i] : cv;
    const bool eq = a == c, gt = a > c, lt = a < c;
    bool p;
    switch (op) {
    case PM_CMP_EQUAL: p = eq; break;
    case PM_CMP_DIFF: p = !eq; break;
    case PM_CMP_SUP: p = gt; break;
    case PM_CMP_SUP_EQUAL: p = gt || eq; break;
    case PM_CMP_INF: p = lt; break;
    default: p = lt || eq; break;  // PM_CMP_INF_EQUAL (op is checked on the host)
    }
    o[i] = p ? (ti ? ti[i] : tv) : (fi ? fi[i] : fv);
  }
}

// ConvertInto / CopyInto (copy_convert_t.hpp:40-80): o = static_cast<TO>(a) per pixel
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void cvt_kernel(const TI *__restrict__ a, TO *__restrict__ o, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) o[i] = (TO)a[i];
}

// marker of dynamic_pseudo_opening/closing (extrema.cpp:90-93,110-113) in one pass:
// mask = (ext == tmp) ? img : E;  marker = (ext != E) ? mask : E, E = lowest (opening) / highest
template <typename T, bool OPENING>
__global__ __launch_bounds__(256) void dpo_kernel(const T *__restrict__ ext, const T *__restrict__ tmp,
                                                  const T *__restrict__ img, T *__restrict__ o, int64_t n)
{
  const T E = OPENING ? Lim<T>::lowest() : Lim<T>::highest();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const T e = ext[i];
    const T m = (e == tmp[i]) ? img[i] : E;
    o[i] = (e != E) ? m : E;
  }
}

template <typename T, int OP> int ew2(const T *a, const T *b, T *o, int64_t n, hipStream_t s)
{
  if (n == 0) return PM_OK;
  hipLaunchKernelGGL((ew2_kernel<T, OP>), dim3(grid1d(n, 256, 8192)), dim3(256), 0, s, a, b, o, n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <typename T, int OP> int ewc(const T *a, T c, T *o, int64_t n, hipStream_t s)
{
  if (n == 0) return PM_OK;
  hipLaunchKernelGGL((ewc_kernel<T, OP>), dim3(grid1d(n, 256, 8192)), dim3(256), 0, s, a, c, o, n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <typename T, bool OPENING> int dpo_marker(const T *ext, const T *tmp, const T *img, T *o, int64_t n, hipStream_t s)
{
  if (n == 0) return PM_OK;
  hipLaunchKernelGGL((dpo_kernel<T, OPENING>), dim3(grid1d(n, 256, 8192)), dim3(256), 0, s, ext, tmp, img, o, n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

#define PM_INST(T)                                                                                           \
  template int ew2<T, EW_SADD>(const T *, const T *, T *, int64_t, hipStream_t);                             \
  template int ew2<T, EW_SUP>(const T *, const T *, T *, int64_t, hipStream_t);                              \
  template int ew2<T, EW_INF>(const T *, const T *, T *, int64_t, hipStream_t);                              \
  template int ew2<T, EW_SSUB>(const T *, const T *, T *, int64_t, hipStream_t);                             \
  template int ewc<T, EWC_SSUB>(const T *, T, T *, int64_t, hipStream_t);                                    \
  template int ewc<T, EWC_SADD>(const T *, T, T *, int64_t, hipStream_t);                                    \
  template int dpo_marker<T, true>(const T *, const T *, const T *, T *, int64_t, hipStream_t);              \
  template int dpo_marker<T, false>(const T *, const T *, const T *, T *, int64_t, hipStream_t);
PM_INST(uint8_t)
PM_INST(uint16_t)
PM_INST(int32_t)
PM_INST(int64_t)
PM_INST(float)
PM_INST(double)
#undef PM_INST

template <int OP> static int arith2(const pm_image *a, const pm_image *b, pm_image *out, void *stream, const char *what)
{
  clear_error();
  PM_TRY(check_image(a, what));
  PM_TRY(check_pair(a, b, what));
  PM_TRY(check_pair(b, out, what));
  const int64_t n = numel(*a);
  auto f = [a = a, b = b, o = out, n = n, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type; return ew2<T, OP>((const T *)a->data, (const T *)b->data, (T *)o->data, n, s); };
  return dispatch_ptype(a->ptype, f);
}

template <int OP> static int arithc(const pm_image *a, double c, pm_image *out, void *stream, const char *what)
{
  clear_error();
  PM_TRY(check_image(a, what));
  PM_TRY(check_pair(a, out, what));
  const int64_t n = numel(*a);
  auto f = [a = a, c = c, o = out, n = n, s = as_stream(stream), what = what](auto tag) -> int {
      using T = typename decltype(tag)::type;
      T cv{};
      PM_TRY(scalar_as<T>(c, &cv, what));
      return ewc<T, OP>((const T *)a->data, cv, (T *)o->data, n, s); };
  return dispatch_ptype(a->ptype, f);
}

// ---------------------------------------------------------------- open / close / gradient
enum { C_OPEN, C_CLOSE, C_GRAD };

template <int K> static int compose(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{
  const char *what = K == C_OPEN ? "Opening" : K == C_CLOSE ? "Closing" : "Gradient";
  PM_TRY(check_unary(in, out, what));
  if (iter < 0) PM_FAIL(PM_E_ARG, "%s iter must be >= 0", what);
  hipStream_t s = as_stream(stream);
  const int64_t n = numel(*in);
  if (n == 0) return PM_OK;
  if (iter == 0) {
    if (K == C_GRAD) return fill_bytes(out->data, 0, (size_t)n * ptype_size(in->ptype), s);
    else PM_HIP(hipMemcpyAsync(out->data, in->data, (size_t)n * ptype_size(in->ptype), hipMemcpyDeviceToDevice, s));
    return PM_OK;
  }
  PM_TRY(check_se_rank(se, in->rank));
  const Dims d = dims3(*in);
  auto f = [in = in, out = out, d = d, rank = in->rank, se = se, iter = iter, s = s, n = n](auto tag) -> int {
      using T = typename decltype(tag)::type;
      const T *src = (const T *)in->data;
      T *dst = (T *)out->data;
      if constexpr (std::is_same<T, uint8_t>::value) {
        // uint8 gradient by a square, r <= 32: one fused launch (no scratch image, no
        // subtraction pass); 8192^2: r = 7 0.053 vs 0.105 ms for the composition, r = 31
        // 0.109 vs 0.137 ms (per-tile thread indices keep two workgroups per CU at r >= 8).
        if (K == C_GRAD && se == PM_SE_SQUARE_2D && rank == 2 && iter <= SQ_U8_GRAD_RMAX)
          return sq_u8_gradient(src, dst, 1, d.H, d.W, iter, s);
        // opening / closing by a square: one fused launch, the eroded (dilated) image stays
        // in LDS
        if (K != C_GRAD && se == PM_SE_SQUARE_2D && rank == 2 && iter <= SQ_U8_OPEN_RMAX &&
            d.H <= INT32_MAX / 2 && d.W <= INT32_MAX / 2)
          return sq_u8_open_close(K == C_OPEN, src, dst, 1, d.H, d.W, iter, s);
      }
      Scratch tmp;
      PM_TRY(tmp.alloc((size_t)n * sizeof(T), s));
      T *t = (T *)tmp.ptr;
      if (K == C_OPEN) {
        PM_TRY((static_iter<T, false>(src, t, d, rank, se, iter, s)));
        return static_iter<T, true>(t, dst, d, rank, se, iter, s);
      } else if (K == C_CLOSE) {
        PM_TRY((static_iter<T, true>(src, t, d, rank, se, iter, s)));
        return static_iter<T, false>(t, dst, d, rank, se, iter, s);
      } else {
        Scratch tmp2;
        PM_TRY(tmp2.alloc((size_t)n * sizeof(T), s));
        T *e = (T *)tmp2.ptr;
        PM_TRY((static_iter<T, true>(src, t, d, rank, se, iter, s)));
        PM_TRY((static_iter<T, false>(src, e, d, rank, se, iter, s)));
        return ew2<T, EW_SSUB>(t, e, dst, n, s);
      }
    };
  return dispatch_ptype(in->ptype, f);
}

// a scalar operand in the pixel type T: integers truncated toward zero and clamped to T's
// range (NaN -> 0), floats converted
template <typename T> static T to_pixel(double v)
{
  if constexpr (std::is_integral<T>::value) {
    if (!(v == v)) return T(0);
    if (v <= (double)std::numeric_limits<T>::lowest()) return std::numeric_limits<T>::lowest();
    if (v >= (double)std::numeric_limits<T>::max()) return std::numeric_limits<T>::max();
    return (T)v;
  } else {
    return (T)v;
  }
}

static int cmp_operand(const pm_operand *o, const pm_image *like, int ptype, const char *name)
{
  if (!o) PM_FAIL(PM_E_ARG, "CompareImage: null %s operand", name);
  if (!o->image) return PM_OK;
  PM_TRY(check_image(o->image, "CompareImage"));
  if (!same_shape(o->image, like)) PM_FAIL(PM_E_SIZE, "CompareImage images have not compatible sizes");
  if (o->image->ptype != ptype) PM_FAIL(PM_E_PTYPE, "CompareImage: the %s image has not the expected PType", name);
  return PM_OK;
}

static int compare_entry(const pm_image *in, int op, const pm_operand *comp, const pm_operand *vt, const pm_operand *vf,
                         pm_image *out, hipStream_t s)
{
  clear_error();
  PM_TRY(check_image(in, "CompareImage"));
  PM_TRY(check_image(out, "CompareImage"));
  if (!same_shape(in, out)) PM_FAIL(PM_E_SIZE, "CompareImage images have not compatible sizes");
  if (op != PM_CMP_EQUAL && op != PM_CMP_DIFF && op != PM_CMP_SUP && op != PM_CMP_SUP_EQUAL && op != PM_CMP_INF &&
      op != PM_CMP_INF_EQUAL)
    PM_FAIL(PM_E_ARG, "CompareImage: unsupported CompOpType %d", op);
  PM_TRY(cmp_operand(comp, in, in->ptype, "comparand"));
  PM_TRY(cmp_operand(vt, in, out->ptype, "true"));
  PM_TRY(cmp_operand(vf, in, out->ptype, "false"));
  const int64_t n = numel(*in);
  if (n == 0) return PM_OK;
  return dispatch_ptype(in->ptype, [&](auto ti) -> int {
    using TI = typename decltype(ti)::type;
    return dispatch_ptype(out->ptype, [&](auto to) -> int {
      using TO = typename decltype(to)::type;
      const TI *cimg = comp->image ? (const TI *)comp->image->data : nullptr;
      const TO *timg = vt->image ? (const TO *)vt->image->data : nullptr;
      const TO *fimg = vf->image ? (const TO *)vf->image->data : nullptr;
      hipLaunchKernelGGL((cmp_kernel<TI, TO>), dim3(grid1d(n, 256, 8192)), dim3(256), 0, s, (const TI *)in->data, op,
                         cimg, to_pixel<TI>(comp->value), timg, to_pixel<TO>(vt->value), fimg, to_pixel<TO>(vf->value),
                         (TO *)out->data, n);
      PM_HIP(hipGetLastError());
      return PM_OK;
    });
  });
}

static int convert_entry(const pm_image *in, pm_image *out, hipStream_t s)
{
  clear_error();
  PM_TRY(check_image(in, "ConvertInto"));
  PM_TRY(check_image(out, "ConvertInto"));
  if (!same_shape(in, out)) PM_FAIL(PM_E_SIZE, "ConvertInto images have not compatible sizes");
  const int64_t n = numel(*in);
  if (n == 0 || (in->data == out->data && in->ptype == out->ptype)) return PM_OK;
  if (in->data == out->data) PM_FAIL(PM_E_ALIAS, "ConvertInto: input and output of different types share memory");
  return dispatch_ptype(in->ptype, [&](auto ti) -> int {
    using TI = typename decltype(ti)::type;
    return dispatch_ptype(out->ptype, [&](auto to) -> int {
      using TO = typename decltype(to)::type;
      hipLaunchKernelGGL((cvt_kernel<TI, TO>), dim3(grid1d(n, 256, 8192)), dim3(256), 0, s, (const TI *)in->data,
                         (TO *)out->data, n);
      PM_HIP(hipGetLastError());
      return PM_OK;
    });
  });
}

static int invert_entry(const pm_image *in, pm_image *out, hipStream_t s)
{
  clear_error();
  PM_TRY(check_image(in, "ArithInvertImage"));
  PM_TRY(check_pair(in, out, "ArithInvertImage"));
  const int64_t n = numel(*in);
  if (n == 0) return PM_OK;
  return dispatch_ptype(in->ptype, [&](auto tag) -> int {
    using T = typename decltype(tag)::type;
    hipLaunchKernelGGL(inv_kernel<T>, dim3(grid1d(n, 256, 8192)), dim3(256), 0, s, (const T *)in->data, (T *)out->data, n);
    PM_HIP(hipGetLastError());
    return PM_OK;
  });
}

}  // namespace pm

using namespace pm;

extern "C" {
int pm_arith_sadd_image(const pm_image *a, const pm_image *b, pm_image *out, void *stream)
{ PM_TRACE(); return arith2<EW_SADD>(a, b, out, stream, "ArithSaturatedAddImage"); }
int pm_arith_invert(const pm_image *in, pm_image *out, void *stream) { PM_TRACE(); return invert_entry(in, out, as_stream(stream)); }
int pm_convert_into(const pm_image *in, pm_image *out, void *stream) { PM_TRACE(); return convert_entry(in, out, as_stream(stream)); }
int pm_compare(const pm_image *in, int op, const pm_operand *comp, const pm_operand *vtrue, const pm_operand *vfalse,
               pm_image *out, void *stream)
{ PM_TRACE();
  return compare_entry(in, op, comp, vtrue, vfalse, out, as_stream(stream));
}
int pm_arith_sup(const pm_image *a, const pm_image *b, pm_image *out, void *stream)
{ PM_TRACE(); return arith2<EW_SUP>(a, b, out, stream, "ArithSupImage"); }
int pm_arith_inf(const pm_image *a, const pm_image *b, pm_image *out, void *stream)
{ PM_TRACE(); return arith2<EW_INF>(a, b, out, stream, "ArithInfImage"); }
int pm_arith_ssub_image(const pm_image *a, const pm_image *b, pm_image *out, void *stream)
{ PM_TRACE(); return arith2<EW_SSUB>(a, b, out, stream, "ArithSaturatedSubImage"); }
int pm_arith_ssub_const(const pm_image *in, double c, pm_image *out, void *stream)
{ PM_TRACE(); return arithc<EWC_SSUB>(in, c, out, stream, "ArithSaturatedSubConstant"); }
int pm_arith_sadd_const(const pm_image *in, double c, pm_image *out, void *stream)
{ PM_TRACE(); return arithc<EWC_SADD>(in, c, out, stream, "ArithSaturatedAddConstant"); }
int pm_opening(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{ PM_TRACE(); return compose<C_OPEN>(in, se, iter, out, stream); }
int pm_closing(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{ PM_TRACE(); return compose<C_CLOSE>(in, se, iter, out, stream); }
int pm_gradient(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{ PM_TRACE(); return compose<C_GRAD>(in, se, iter, out, stream); }
}
