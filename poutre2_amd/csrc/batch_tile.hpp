// batch_tile.hpp -- device code of the batched dilation + opening (batch.hip): the image
// descriptor table, the per-item tile code and the persistent kernel, instantiated per
// radius range in batch_r*.hip so that the instantiations compile in parallel.
#pragma once
#include "pm_common.hpp"
#include "sq_open.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

// One image of a batch.  tile0: index of its first dilation tile (128 x TW) in its radius
// group's work list; otile0: index of its first fused-opening tile (open_fo(r) x OPEN_TW,
// radius groups up to BATCH_FUSED_RMAX; 0 otherwise); tmp: its eroded image (two-phase groups).
struct BatchImage {
  const uint8_t *in;
  uint8_t *dil, *open, *tmp;
  int H, W, r, tiles_x, tile0, otile0;
};

// the image whose tile range (by tile0, or by otile0 when OPEN) holds tile t
template <bool OPEN = false> __device__ __forceinline__ int batch_find(const BatchImage *im, int n, int t)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((OPEN ? im[mid].otile0 : im[mid].tile0) <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int R, bool DIL>
__device__ __forceinline__ void batch_tile(uint8_t *smem, const uint8_t *src, uint8_t *dst, const BatchImage &b,
                                           int y0, int x0)
{
  constexpr int K = sq_k_of(R), TW = sq_tw(K, R);
  // thread-index-derived offsets recomputed per item (tile_tid): the batched kernels hold
  // two or three tile variants, whose hoisted offsets would all stay live
  sq_tile_load<K, DIL, false, R, TW, sq_r16<K>(), true>(smem, src, b.H, b.W, R, y0, x0);
  __syncthreads();
  sq_tile_passes<K, DIL, false, R, TW, FT, true>(smem, dst, b.H, b.W, R, y0, x0);
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

// Radius groups up to this radius run the one-launch form (batch_fused_kernel).  The fused
// opening tile is used up to r = 15 here although single images switch to the two-launch
// composition above r = 7 (sq_open.hpp): in a batch the composition's second launch per
// radius group costs more than the fused tile's larger halo.
constexpr int BATCH_FUSED_RMAX = 15;

// Radius groups up to BATCH_FUSED_RMAX: one launch over ndil dilation tiles and nitems - ndil
// fused-opening tiles (sq_open.hpp) of the same images, so no eroded image goes through
// memory and no second launch is needed.  Workgroups take items from a queue (*queue: the
// next item, zero at launch): a workgroup that becomes resident late, when other groups'
// kernels share the GPU, takes fewer items instead of a fixed share.  The opening tiles, the
// heavier items, come first.
template <int R>
__global__ __launch_bounds__(FTHREADS) void batch_fused_kernel(const BatchImage *__restrict__ im, int nimg, int ndil,
                                                               int nitems, int *queue)
{
  constexpr int K = sq_k_open(R), TW = sq_tw(K, R), FO = open_fo(R);  // K: the opening tile's runs
  static_assert(R <= BATCH_FUSED_RMAX && TW == OPEN_TW, "fused batch geometry");
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ int s_item;
  const int nopen = nitems - ndil;
  for (;;) {
    __syncthreads();  // every thread has read the previous item
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int item = s_item;
    if (item >= nitems) break;
    if (item >= nopen) {
      const int d = item - nopen;
      const BatchImage b = im[batch_find(im, nimg, d)];
      const int tl = d - b.tile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
      batch_tile<R, true>(smem, b.in, b.dil, b, ty * FT, tx * TW);
    } else {
      const int t = item;
      const BatchImage b = im[batch_find<true>(im, nimg, t)];
      const int tl = t - b.otile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
      const int y0 = ty * FO, x0 = tx * OPEN_TW;
      sq_tile_load<K, false, false, R, OPEN_TW, open_pad(R), true>(smem, b.in, b.H, b.W, R, y0 - R, x0);
      __syncthreads();
      sq_open_tile<K, true, false, R, OPEN_TW>(smem, b.open, b.H, b.W, y0, x0);
      __syncthreads();  // LDS is rewritten by the next item
    }
  }
}

typedef void (*batch_kernel_t)(const BatchImage *, int, int);
typedef void (*batch_fused_t)(const BatchImage *, int, int, int, int *);
// fused kernel of radius r in [1, BATCH_FUSED_RMAX] (batch_open.hip)
batch_fused_t batch_fused_for(int r);

// kernel of radius r in [LO, LO + 7] and phase 0 / 1 (batch_r*.hip)
template <int LO> batch_kernel_t batch_kernel_range(int r, int phase);
template <> batch_kernel_t batch_kernel_range<1>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<9>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<17>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<25>(int r, int phase);

}  // namespace pm
