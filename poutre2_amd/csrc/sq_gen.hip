// sq_gen.hip -- launcher of the fused square kernels for uint16 / int32 / int64 / float32 /
// float64 images (kernels: sq_u16.hpp for uint16, instantiated in sq_u16_k*.hip; sq_gen.hpp
// for the others, instantiated in sq_gen_<type>.hip).
#include <type_traits>

#include "sq_gen.hpp"
#include "sq_u16.hpp"

namespace pm {

static_assert(SQ_U16_RMAX == SQ_GEN_RMAX, "one chaining rule for every type");

template <bool DIL>
static int sq_u16_once(const uint16_t *in, uint16_t *out, int64_t B, int64_t H, int64_t W, int r, hipStream_t s)
{
  const bool aligned = (W % 8) == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0;
  const int th = u16_th(r);
  const int64_t tiles_x = (W + U16_TW - 1) / U16_TW, tiles_y = (H + th - 1) / th;
  const int64_t per = tiles_x * tiles_y, ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused square kernel");
  const u16_kernel_t k = aligned ? sq_u16_kernel_for<DIL, true>(r) : sq_u16_kernel_for<DIL, false>(r);
  const size_t lds = sq_u16_lds_bytes(r);
  int dev = 0, cus = 0, occ = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, FTHREADS, lds));
  int64_t g = (int64_t)(occ > 0 ? occ : 1) * cus;
  if (g >= 8) g &= ~(int64_t)7;
  if (g > ntiles) g = ntiles;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(FTHREADS), lds, s, in, out, (int)H, (int)W, (int)tiles_x, (int)per,
                     (int)ntiles);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <typename T, bool DIL>
static int sq_gen_generic(const T *in, T *out, int64_t B, int64_t H, int64_t W, int r, hipStream_t s)
{
  constexpr int TWE = gen_twe<T>();
  const int64_t tiles_x = (W + TWE - 1) / TWE, tiles_y = (H + GEN_FT - 1) / GEN_FT;
  const int64_t per = tiles_x * tiles_y, ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused square kernel");
  const gen_kernel_t<T, DIL> k = sq_gen_kernel_for<T, DIL>(r);
  const size_t lds = gen_lds_bytes<T>(r);
  int dev = 0, cus = 0, occ = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, 256, lds));
  int64_t g = (int64_t)(occ > 0 ? occ : 1) * cus;
  if (g >= 8) g &= ~(int64_t)7;
  if (g > ntiles) g = ntiles;
  // rows of whole, 16-byte aligned chunks: the region is read chunk by chunk
  const int al = (W % (16 / (int64_t)sizeof(T)) == 0 && ((uintptr_t)in & 15) == 0) ? 1 : 0;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(256), lds, s, in, out, (int)H, (int)W, (int)tiles_x, (int)per,
                     (int)ntiles, al);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <typename T, bool DIL>
static int sq_gen_once(const T *in, T *out, int64_t B, int64_t H, int64_t W, int r, hipStream_t s)
{
  if constexpr (std::is_same<T, uint16_t>::value) return sq_u16_once<DIL>(in, out, B, H, W, r, s);
  else return sq_gen_generic<T, DIL>(in, out, B, H, W, r, s);
}

template <typename T, bool DIL> int sq_gen(const T *in, T *out, int64_t B, int64_t H, int64_t W, int64_t r, hipStream_t s)
{
  const int64_t n = B * H * W;
  if (n == 0) return PM_OK;
  if (r < 1) PM_FAIL(PM_E_ARG, "square radius %lld < 1", (long long)r);
  if (!sq_gen_applicable(H, W)) PM_FAIL(PM_E_ARG, "image too large for the fused square kernel");
  // a window wider than the image covers it whole: clamp (same result, fewer launches)
  const int64_t rmax = (H > W ? H : W);
  if (r > rmax) r = rmax;
  if (r <= SQ_GEN_RMAX) return sq_gen_once<T, DIL>(in, out, B, H, W, (int)r, s);
  // chain: radii of at most SQ_GEN_RMAX summing to r, ping-ponging so the last writes out
  const int64_t steps = (r + SQ_GEN_RMAX - 1) / SQ_GEN_RMAX;
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * sizeof(T), s));
  T *bufs[2] = {out, reinterpret_cast<T *>(tmp.ptr)};
  int dst = (steps % 2 == 1) ? 0 : 1;
  const T *src = in;
  for (int64_t i = 0; i < steps; i++) {
    const int ri = i == 0 ? (int)(r - SQ_GEN_RMAX * (steps - 1)) : SQ_GEN_RMAX;  // 1 .. SQ_GEN_RMAX
    PM_TRY((sq_gen_once<T, DIL>(src, bufs[dst], B, H, W, ri, s)));
    src = bufs[dst];
    dst ^= 1;
  }
  return PM_OK;
}

#define PM_GEN_INST(T)                                                                                       \
  template int sq_gen<T, true>(const T *, T *, int64_t, int64_t, int64_t, int64_t, hipStream_t);             \
  template int sq_gen<T, false>(const T *, T *, int64_t, int64_t, int64_t, int64_t, hipStream_t);
PM_GEN_INST(uint16_t)
PM_GEN_INST(int32_t)
PM_GEN_INST(int64_t)
PM_GEN_INST(float)
PM_GEN_INST(double)
#undef PM_GEN_INST

}  // namespace pm
