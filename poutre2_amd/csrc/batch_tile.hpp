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

// Radius groups up to this radius run in the one-launch kernel (batch_one_kernel).  The fused
// opening tile is used up to r = 15 here although single images switch to the two-launch
// composition above r = 7 (sq_open.hpp): in a batch the composition's second launch costs
// more than the fused tile's larger halo.
constexpr int BATCH_FUSED_RMAX = 15;

// A radius group of the one-launch kernel (batch_one_kernel): its images are im[img0 ..
// img0 + nimg - 1] of the descriptor table, its items are [item0, item0 + nitems) of the
// launch's work list: first its nopen fused-opening tiles, then its dilation tiles.
struct BatchGroup {
  int r, img0, nimg, item0, nopen, nitems;
};

// Item t of a radius group up to BATCH_FUSED_RMAX: t < nopen is fused-opening tile t
// (sq_open.hpp: erosion and dilation in LDS, no eroded image in memory), otherwise
// dilation tile t - nopen.  im / nimg: the group's images.
template <int R>
__device__ __forceinline__ void batch_fused_item(uint8_t *smem, const BatchImage *__restrict__ im, int nimg, int nopen,
                                                 int t)
{
  constexpr int K = sq_k_open(R), TW = sq_tw(K, R), FO = open_fo(R);  // K: the opening tile's runs
  static_assert(R <= BATCH_FUSED_RMAX && TW == OPEN_TW, "fused batch geometry");
  if (t >= nopen) {
    const int d = t - nopen;
    const BatchImage b = im[batch_find(im, nimg, d)];
    const int tl = d - b.tile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
    batch_tile<R, true>(smem, b.in, b.dil, b, ty * FT, tx * TW);
  } else {
    const BatchImage b = im[batch_find<true>(im, nimg, t)];
    const int tl = t - b.otile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
    const int y0 = ty * FO, x0 = tx * OPEN_TW;
    sq_tile_load<K, false, false, R, OPEN_TW, open_pad(R), true>(smem, b.in, b.H, b.W, R, y0 - R, x0);
    __syncthreads();
    sq_open_tile<K, true, false, R, OPEN_TW>(smem, b.open, b.H, b.W, y0, x0);
    __syncthreads();  // LDS is rewritten by the next item
  }
}

// One persistent launch over every radius group up to BATCH_FUSED_RMAX (batch_open.hip).
// Workgroups take items from a queue: ctr[0] is the next item, ctr[1] counts the workgroups
// that found the queue empty; both are zero at launch, and the last workgroup to finish
// resets them, so the launch can be replayed (graph plans) without clearing them first.
// Groups come in order of decreasing radius (the heavier tiles first, the cheapest last, so
// the launch's tail is short), each group's opening tiles before its dilation tiles; at any
// time the resident workgroups mostly run one radius, i.e. one specialisation of the code.
// Dynamic LDS: the largest tile of the launch's radii (batch_one_lds_bytes).
__global__ __launch_bounds__(FTHREADS) void batch_one_kernel(const BatchImage *__restrict__ im,
                                                             const BatchGroup *__restrict__ grp, int ngrp,
                                                             int nitems, int *ctr);
inline size_t batch_one_lds_bytes(int r)
{
  const size_t a = sq_tile_lds_bytes(r, sq_tw(sq_k_of(r), r)), b = sq_open_lds_bytes(r);
  return a > b ? a : b;
}

typedef void (*batch_kernel_t)(const BatchImage *, int, int);

// kernel of radius r in [LO, LO + 7] and phase 0 / 1 (batch_r*.hip)
template <int LO> batch_kernel_t batch_kernel_range(int r, int phase);
template <> batch_kernel_t batch_kernel_range<1>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<9>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<17>(int r, int phase);
template <> batch_kernel_t batch_kernel_range<25>(int r, int phase);

}  // namespace pm
