// sq_u8_kernel.hpp -- the persistent fused square kernel (see sq_u8.hip), instantiated
// per operation and alignment class in sq_u8_k*.hip so that they compile in parallel.
#pragma once
#include "pm_common.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

// Persistent tile schedule.  Workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b on XCD b % 8); XCD x owns the contiguous tile range [x*n/8, (x+1)*n/8) so
// that neighbouring tiles, which share halo rows, are read through the same L2.
struct TileSched {
  int first, step, end;
};
__device__ __forceinline__ TileSched tile_sched(int b, int G, int ntiles)
{
  if ((G & 7) != 0 || ntiles < 8) return {b, G, ntiles};
  const int x = b & 7, j = b >> 3, gx = G >> 3;
  const int lo = (int)((int64_t)x * ntiles / 8), hi = (int)((int64_t)(x + 1) * ntiles / 8);
  return {lo + j, gx, hi};
}

template <int K, bool DIL, bool ALIGNED, int RC, int TW>
__global__ __launch_bounds__(FTHREADS) void sq_u8_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                         int H, int W, int r_rt, int tiles_x, int tiles_per_img,
                                                         int ntiles)
{
  extern __shared__ __align__(16) uint8_t smem[];
  static_assert(RC > 0, "the tile height follows the compile-time radius");
  constexpr int TH = sq_th(RC);  // tile height (the host sizes the grid and LDS with sq_th)
  // A tile's region is read with all its loads in flight, then copied to LDS.  The
  // workgroups resident on a CU overlap each other's loads and compute; prefetching the
  // next tile into registers instead was slower (r = 1: 34.9 vs 33.5 us, r = 7: 40.2 vs
  // 36.5 us, r = 31: 55.7 vs 52.1 us on 8192x8192), its registers costing occupancy.
  const TileSched ts = tile_sched(blockIdx.x, gridDim.x, ntiles);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int img = tile / tiles_per_img;
    const int t_in = tile - img * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * TH, x0 = tx * TW;
    const int64_t ibase = (int64_t)img * H * W;
    sq_tile_load<K, DIL, ALIGNED, RC, TW, sq_r16<K>(), false, TH>(smem, in + ibase, H, W, r_rt, y0, x0);
    __syncthreads();
    sq_tile_passes<K, DIL, ALIGNED, RC, TW, TH>(smem, out + ibase, H, W, r_rt, y0, x0);
    __syncthreads();  // A/T are rewritten by the next tile
  }
}

typedef void (*sq_kernel_t)(const uint8_t *, uint8_t *, int, int, int, int, int, int);

// Kernel of radius r (1..32) and tile width tw (128, 192 or 256), specialised on both, for
// 16-byte aligned rows (ALIGNED) or rows of any alignment (sq_u8_k*.hip).
template <bool DIL, bool ALIGNED> sq_kernel_t sq_kernel_for(int r, int tw);
template <> sq_kernel_t sq_kernel_for<true, true>(int r, int tw);
template <> sq_kernel_t sq_kernel_for<true, false>(int r, int tw);
template <> sq_kernel_t sq_kernel_for<false, true>(int r, int tw);
template <> sq_kernel_t sq_kernel_for<false, false>(int r, int tw);

}  // namespace pm
