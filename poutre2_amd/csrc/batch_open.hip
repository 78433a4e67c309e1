// batch_open.hip -- the one-launch batched dilation + fused opening over every radius group
// up to BATCH_FUSED_RMAX (batch_tile.hpp).
#include "batch_tile.hpp"

namespace pm {

// the next item of the queue, for the whole workgroup
__device__ __forceinline__ int next_item(int *ctr, int *s_item)
{
  __syncthreads();  // every thread has read the previous item
  if (threadIdx.x == 0) *s_item = atomicAdd(ctr, 1);
  __syncthreads();
  return *s_item;
}

// Items of radius group G from `item` on, until the queue hands out an item past the group
// (items come out in increasing order); returns that item.  The loop over one group is a
// loop of its own so that what the compiler keeps live across items is one radius's state.
template <int R>
__device__ __forceinline__ int group_items(uint8_t *smem, const BatchImage *__restrict__ gim, BatchGroup G, int item,
                                        int *ctr, int *s_item)
{
  const int end = G.item0 + G.nitems;
  do {
    batch_fused_item<R>(smem, gim, G.nimg, G.nopen, item - G.item0);
    item = next_item(ctr, s_item);
  } while (item < end);
  return item;
}

__global__ __launch_bounds__(FTHREADS) void batch_one_kernel(const BatchImage *__restrict__ im,
                                                             const BatchGroup *__restrict__ grp, int ngrp,
                                                             int nitems, int *ctr)
{
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ int s_item;
  int item = next_item(ctr, &s_item);
  int g = 0;
  while (item < nitems) {
    // the item's radius group (groups sorted by item0; items only increase)
    while (g + 1 < ngrp && grp[g + 1].item0 <= item) g++;
    const BatchGroup G = grp[g];
    const BatchImage *gim = im + G.img0;
    switch (G.r) {
#define PM_BATCH_CASE(R) \
  case R: item = group_items<R>(smem, gim, G, item, ctr, &s_item); break;
      PM_BATCH_CASE(1) PM_BATCH_CASE(2) PM_BATCH_CASE(3) PM_BATCH_CASE(4) PM_BATCH_CASE(5)
      PM_BATCH_CASE(6) PM_BATCH_CASE(7) PM_BATCH_CASE(8) PM_BATCH_CASE(9) PM_BATCH_CASE(10)
      PM_BATCH_CASE(11) PM_BATCH_CASE(12) PM_BATCH_CASE(13) PM_BATCH_CASE(14) PM_BATCH_CASE(15)
#undef PM_BATCH_CASE
    default: item = nitems; break;  // not reached: every group has a radius in [1, 15]
    }
  }
  static_assert(BATCH_FUSED_RMAX == 15, "one case per radius up to BATCH_FUSED_RMAX");
  // every workgroup is past its last read of ctr[0]: the last one to get here resets the
  // queue for the next launch
  if (threadIdx.x == 0 && atomicAdd(ctr + 1, 1) == (int)gridDim.x - 1) {
    atomicExch(ctr, 0);
    atomicExch(ctr + 1, 0);
  }
}

}  // namespace pm
