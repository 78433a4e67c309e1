// sq_open.hip -- launcher of the fused uint8 opening / closing (kernels: sq_open.hpp,
// instantiated in sq_open_{o,c}{a,u}.hip).
#include "sq_open.hpp"
#include "sq_u8.hpp"

namespace pm {

int sq_u8_open_close(bool open, const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int64_t r,
                     hipStream_t s)
{
  const int64_t n = B * H * W;
  if (n == 0) return PM_OK;
  if (r < 1 || r > SQ_U8_OPEN_RMAX)
    PM_FAIL(PM_E_ARG, "fused opening/closing radius %lld outside [1, %d]", (long long)r, (int)SQ_U8_OPEN_RMAX);
  if (H > INT32_MAX / 2 || W > INT32_MAX / 2) PM_FAIL(PM_E_ARG, "image too large for the fused opening/closing");
  // the 3 x 3 square on rows of a multiple of 16 bytes: streamed through registers
  if (r == 1 && sq_u8_rs_applicable(in, out, H, W)) return sq_u8_rs_compose(open ? 0 : 1, in, out, B, H, W, s);
  const int fo = open_fo((int)r);
  const int64_t tiles_x = (W + OPEN_TW - 1) / OPEN_TW, tiles_y = (H + fo - 1) / fo;
  const int64_t per = tiles_x * tiles_y, ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused opening/closing");
  const bool aligned = (W % 16) == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0;
  const open_kernel_t k = open ? (aligned ? sq_open_kernel_for<true, true>((int)r) : sq_open_kernel_for<true, false>((int)r))
                               : (aligned ? sq_open_kernel_for<false, true>((int)r) : sq_open_kernel_for<false, false>((int)r));
  const size_t lds = sq_open_lds_bytes((int)r);
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

}  // namespace pm
