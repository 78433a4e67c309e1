// batch_tile.hpp -- device code of the batched dilation + opening (batch.hip): the image
// descriptor table, the per-item tile code and the persistent kernel, instantiated per
// radius range in batch_r*.hip so that the instantiations compile in parallel.
#pragma once
#include "pm_common.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

struct BatchImage {
  const uint8_t *in;
  uint8_t *dil, *open, *tmp;
  int H, W, r, tiles_x, tile0;
};

__device__ __forceinline__ int batch_find(const BatchImage *im, int n, int t)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (im[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int R, bool DIL>
__device__ __forceinline__ void batch_tile(uint8_t *smem, const uint8_t *src, uint8_t *dst, const BatchImage &b,
                                           int y0, int x0)
{
  constexpr int K = sq_k_of(R), TW = sq_tw(K, R);
  sq_tile_load<K, DIL, false, R, TW>(smem, src, b.H, b.W, R, y0, x0);
  __syncthreads();
  sq_tile_passes<K, DIL, false, R, TW>(smem, dst, b.H, b.W, R, y0, x0);
  __syncthreads();  // LDS is rewritten by the next item
}

// PHASE 0: items 2t (dilation of tile t into dil) and 2t+1 (erosion into tmp);
// PHASE 1: item t = dilation of tile t of tmp into open.
template <int R, int PHASE>
__global__ __launch_bounds__(FTHREADS) void batch_sq_kernel(const BatchImage *__restrict__ im, int nimg, int nitems)
{
  constexpr int TW = sq_tw(sq_k_of(R), R);
  extern __shared__ __align__(16) uint8_t smem[];
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int t = PHASE == 0 ? item >> 1 : item;
    const BatchImage b = im[batch_find(im, nimg, t)];
    const int tl = t - b.tile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
    const int y0 = ty * FT, x0 = tx * TW;
    if (PHASE == 0) {
      if (item & 1) batch_tile<R, false>(smem, b.in, b.tmp, b, y0, x0);
      else batch_tile<R, true>(smem, b.in, b.dil, b, y0, x0);
    } else {
      batch_tile<R, true>(smem, b.tmp, b.open, b, y0, x0);
    }
  }
}

typedef void (*batch_kernel_t)(const BatchImage *, int, int);

// kernel of radius r in [LO, LO + 7] and phase 0 / 1 (batch_r*.hip)
template <int LO> batch_kernel_t batch_kernel_range(int r, int phase);
template <> batch_kernel_t batch_kernel_range<1>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<9>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<17>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<25>(int r, int phase);

}  // namespace pm
