// batch.hip -- many uint8 images of uneven sizes and mixed square sizes in one launch.
//
// pm_batch_dilate_open computes, for every image i of a batch,
//   dil[i]  = Dilate(in[i], SESquare2D, r_i)                      (ero_dil_static.cpp:268-305)
//   open[i] = Dilate(Erode(in[i], SESquare2D, r_i), SESquare2D, r_i)
// in ONE cooperative persistent launch.  The work list spans all images: phase 1 holds two
// items per 128x128 tile (its dilation, and its erosion into a scratch image), phase 2 one
// item per tile (the dilation of the eroded tile).  A grid barrier separates the phases; it
// releases the phase-1 writes at device scope so that phase 2 may read eroded tiles written
// by any XCD.  Each item runs the same tile code as the single-image fused kernel
// (sq_u8.hip), specialised by run length K for the item's radius through a switch, so one
// launch serves every radius of the batch.  The reference runs the two Erode/Dilate calls
// per image, one image at a time.
#include <vector>

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

template <bool DIL>
__device__ __forceinline__ void batch_tile(uint8_t *smem, const uint8_t *src, uint8_t *dst, int H, int W, int r, int y0,
                                           int x0)
{
  if (r >= 16) {
    sq_load_region<32, DIL>(smem, src, H, W, r, y0, x0);
    __syncthreads();
    sq_tile_passes<32, DIL, false>(smem, dst, H, W, r, y0, x0);
  } else if (r >= 8) {
    sq_load_region<16, DIL>(smem, src, H, W, r, y0, x0);
    __syncthreads();
    sq_tile_passes<16, DIL, false>(smem, dst, H, W, r, y0, x0);
  } else if (r >= 4) {
    sq_load_region<8, DIL>(smem, src, H, W, r, y0, x0);
    __syncthreads();
    sq_tile_passes<8, DIL, false>(smem, dst, H, W, r, y0, x0);
  } else {
    sq_load_region<4, DIL>(smem, src, H, W, r, y0, x0);
    __syncthreads();
    sq_tile_passes<4, DIL, false>(smem, dst, H, W, r, y0, x0);
  }
  __syncthreads();
}

// Grid-wide barrier for a cooperative launch (all workgroups resident).  Every workgroup
// arrives exactly once, so the wait ends for all of them.
__device__ __forceinline__ void grid_barrier(unsigned *count, unsigned G)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // release this workgroup's writes at device scope
    atomicAdd(count, 1u);
    while (__hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < G) __builtin_amdgcn_s_sleep(4);
    __threadfence();
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void batch_dilopen_kernel(const BatchImage *__restrict__ im, int nimg, int ntiles,
                                                            unsigned *barrier)
{
  extern __shared__ __align__(16) uint8_t smem[];
  const int G = gridDim.x;
  for (int item = blockIdx.x; item < 2 * ntiles; item += G) {
    const int t = item >> 1;
    const BatchImage b = im[batch_find(im, nimg, t)];
    const int tl = t - b.tile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
    if (item & 1) batch_tile<false>(smem, b.in, b.tmp, b.H, b.W, b.r, ty * FT, tx * FT);
    else batch_tile<true>(smem, b.in, b.dil, b.H, b.W, b.r, ty * FT, tx * FT);
  }
  grid_barrier(barrier, (unsigned)G);
  for (int t = blockIdx.x; t < ntiles; t += G) {
    const BatchImage b = im[batch_find(im, nimg, t)];
    const int tl = t - b.tile0, ty = tl / b.tiles_x, tx = tl - ty * b.tiles_x;
    batch_tile<true>(smem, b.tmp, b.open, b.H, b.W, b.r, ty * FT, tx * FT);
  }
}

__global__ void zero_u32_kernel(unsigned *p) { p[0] = 0u; }

}  // namespace pm

using namespace pm;

extern "C" int pm_batch_dilate_open(int n, const pm_image *ins, const int *radii, pm_image *dil_outs,
                                    pm_image *open_outs, void *stream)
{
  clear_error();
  if (n < 0 || (n > 0 && (!ins || !radii || !dil_outs || !open_outs))) PM_FAIL(PM_E_ARG, "batch: bad arguments");
  if (n == 0) return PM_OK;
  hipStream_t s = as_stream(stream);
  std::vector<BatchImage> host(n);
  int64_t tiles = 0, tmp_bytes = 0, maxr = 0;
  for (int i = 0; i < n; i++) {
    PM_TRY(check_unary(&ins[i], &dil_outs[i], "batch dilate"));
    PM_TRY(check_unary(&ins[i], &open_outs[i], "batch open"));
    if (dil_outs[i].data == open_outs[i].data) PM_FAIL(PM_E_ALIAS, "batch: dilation and opening outputs must differ");
    if (ins[i].ptype != PM_PTYPE_U8 || ins[i].rank != 2) PM_FAIL(PM_E_UNSUPPORTED, "batch: uint8 2D images only");
    if (radii[i] < 1 || radii[i] > SQ_TILE_RMAX) PM_FAIL(PM_E_ARG, "batch: radius must be in [1, %d]", SQ_TILE_RMAX);
    const int64_t H = ins[i].shape[0], W = ins[i].shape[1];
    if (H > (1 << 28) || W > (1 << 28)) PM_FAIL(PM_E_SIZE, "batch: image too large");
    BatchImage b;
    b.in = (const uint8_t *)ins[i].data;
    b.dil = (uint8_t *)dil_outs[i].data;
    b.open = (uint8_t *)open_outs[i].data;
    b.tmp = nullptr;
    b.H = (int)H;
    b.W = (int)W;
    b.r = radii[i];
    b.tiles_x = (int)((W + FT - 1) / FT);
    b.tile0 = (int)tiles;
    tiles += (int64_t)b.tiles_x * ((H + FT - 1) / FT);
    if (tiles > INT32_MAX / 4) PM_FAIL(PM_E_SIZE, "batch: too many tiles");
    tmp_bytes += H * W;
    maxr = radii[i] > maxr ? radii[i] : maxr;
    host[i] = b;
  }
  if (tiles == 0) return PM_OK;
  // scratch: eroded images, the descriptor table and the barrier counter
  const size_t desc_off = ((size_t)tmp_bytes + 255) & ~(size_t)255;
  const size_t bar_off = desc_off + (((size_t)n * sizeof(BatchImage) + 255) & ~(size_t)255);
  Scratch sc;
  PM_TRY(sc.alloc(bar_off + 256, s));
  uint8_t *base = (uint8_t *)sc.ptr;
  int64_t off = 0;
  for (int i = 0; i < n; i++) {
    host[i].tmp = base + off;
    off += (int64_t)host[i].H * host[i].W;
  }
  BatchImage *dimg = (BatchImage *)(base + desc_off);
  unsigned *bar = (unsigned *)(base + bar_off);
  PM_HIP(hipMemcpyAsync(dimg, host.data(), (size_t)n * sizeof(BatchImage), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, bar);
  PM_HIP(hipGetLastError());
  const size_t lds = sq_tile_lds_bytes((int)maxr);
  int dev = 0, cus = 0, occ = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_dilopen_kernel, 256, lds));
  if (occ < 1) PM_FAIL(PM_E_HIP, "batch: kernel does not fit on a CU");
  int64_t G = (int64_t)occ * cus;
  if (G > 2 * tiles) G = 2 * tiles;
  int nimg = n, ntiles = (int)tiles;
  void *args[] = {(void *)&dimg, (void *)&nimg, (void *)&ntiles, (void *)&bar};
  PM_HIP(hipLaunchCooperativeKernel((const void *)batch_dilopen_kernel, dim3((unsigned)G), dim3(256), args, lds, s));
  return PM_OK;
}
