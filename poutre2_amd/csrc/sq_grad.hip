// sq_grad.hip -- fused morphological gradient of uint8 images by a (2r+1)^2 square, r <= 32:
// out = Dilate(in, SESquare2D, r) - Erode(in, SESquare2D, r) in one launch (the composition
// of ero_dil_static.cpp:229-305 with arith_saturated_sub_image; max >= min, so the
// subtraction never saturates).  One read of the input, one write of the output; the tile
// code is that of the fused square kernel (sq_u8_tile.hpp) run for max and min on one
// region with clamped borders.  Up to r = 15 the max and min go to two transposed buffers
// (128-column tiles); from r = 16 (runs of 32 rows) one transposed buffer serves both in
// turn (sq_tile_grad_passes1), which halves the LDS and allows 192-column tiles: 8192x8192,
// us, two buffers / 128 columns vs one buffer / 192 columns: r = 16 92.8 / 77.6, r = 24
// 100.5 / 82.4, r = 31 108.8 / 88.5 (r = 15: 66.0 / 63.8, r = 8: 57.5 / 57.2).
#include <utility>

#include "pm_common.hpp"
#include "sq_u8.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

// tile width: 128 columns with two transposed buffers (runs of up to 16 rows), 192 with one
constexpr int grad_tw(int K) { return K == 32 ? 192 : 128; }
inline size_t grad_lds_bytes(int r)
{
  return sq_k_open(r) == 32 ? sq_tile_lds_bytes(r, grad_tw(32)) : sq_grad_lds_bytes(r, grad_tw(16));
}

struct GradSched {
  int first, step, end;
};

template <int K, bool ALIGNED, int RC>
__global__ __launch_bounds__(FTHREADS) void sq_grad_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                           int H, int W, int tiles_x, int tiles_per_img, int ntiles)
{
  extern __shared__ __align__(16) uint8_t smem[];
  // XCD-aware persistent schedule (see sq_u8.hip)
  const int G = gridDim.x, b = blockIdx.x;
  GradSched ts{b, G, ntiles};
  if ((G & 7) == 0 && ntiles >= 8) {
    const int x = b & 7, j = b >> 3, gx = G >> 3;
    ts = {(int)((int64_t)x * ntiles / 8) + j, gx, (int)((int64_t)(x + 1) * ntiles / 8)};
  }
  constexpr int TW = grad_tw(K);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int img = tile / tiles_per_img;
    const int t_in = tile - img * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * FT, x0 = tx * TW;
    const int64_t ibase = (int64_t)img * H * W;
    // K >= 16 (r >= 8): indices recomputed per tile (fewer registers) are faster; below, the
    // hoisted ones (see tile_tid)
    sq_tile_load_clamped<K, ALIGNED, RC, TW, (K >= 16)>(smem, in + ibase, H, W, RC, y0, x0);
    __syncthreads();
    if constexpr (K == 32) sq_tile_grad_passes1<K, ALIGNED, RC, TW, true>(smem, out + ibase, H, W, RC, y0, x0);
    else sq_tile_grad_passes<K, ALIGNED, RC, TW, (K >= 16)>(smem, out + ibase, H, W, RC, y0, x0);
    __syncthreads();
  }
}

typedef void (*grad_kernel_t)(const uint8_t *, uint8_t *, int, int, int, int, int);

template <int... RS> static grad_kernel_t grad_pick(int r, bool aligned, std::integer_sequence<int, RS...>)
{
  static constexpr grad_kernel_t al[] = {sq_grad_kernel<sq_k_open(RS + 1), true, RS + 1>...};
  static constexpr grad_kernel_t any[] = {sq_grad_kernel<sq_k_open(RS + 1), false, RS + 1>...};
  return aligned ? al[r - 1] : any[r - 1];
}

int sq_u8_gradient(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int64_t r, hipStream_t s)
{
  const int64_t n = B * H * W;
  if (n == 0) return PM_OK;
  if (r < 1 || r > SQ_U8_GRAD_RMAX)
    PM_FAIL(PM_E_ARG, "fused gradient radius %lld outside [1, %d]", (long long)r, (int)SQ_U8_GRAD_RMAX);
  if (H > INT32_MAX / 2 || W > INT32_MAX / 2) PM_FAIL(PM_E_ARG, "image too large for the fused gradient");
  // the 3 x 3 square on rows of a multiple of 16 bytes: streamed through registers
  if (r == 1 && sq_u8_rs_applicable(in, out, H, W)) return sq_u8_rs_compose(2, in, out, B, H, W, s);
  const int tw = grad_tw(sq_k_open((int)r));
  const int tiles_x = (int)((W + tw - 1) / tw), tiles_y = (int)((H + FT - 1) / FT);
  const int64_t per = (int64_t)tiles_x * tiles_y, ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused gradient");
  const bool aligned = (W % 16) == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0;
  const grad_kernel_t k = grad_pick((int)r, aligned, std::make_integer_sequence<int, (int)SQ_U8_GRAD_RMAX>{});
  const size_t lds = grad_lds_bytes((int)r);
  int dev = 0, cus = 0, occ = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, FTHREADS, lds));
  int64_t g = (int64_t)(occ > 0 ? occ : 1) * cus;
  if (g >= 8) g &= ~(int64_t)7;
  if (g > ntiles) g = ntiles;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(FTHREADS), lds, s, in, out, (int)H, (int)W, tiles_x, (int)per,
                     (int)ntiles);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

}  // namespace pm
