// sq_u8.hip -- fused flat dilation / erosion of uint8 images by a (2r+1)^2 square, r <= 32.
//
// This is the kernel behind Dilate/Erode(SESquare2D, iter = r) and DilateX+DilateY for
// uint8 (ero_dil_static.cpp:229-305, ero_dil_line_se_t.hpp:49-280): one launch reads
// every input pixel once from HBM and writes every output once.
//
// One workgroup (256 threads) owns an output tile of 128 rows and TW = 256 (r <= 16) or
// 128 columns.
//  1. The input region (128+2r rows, TW+2*R16 columns, R16 = 16 for r <= 15 and 32 above,
//     so that global loads of aligned rows are 16-byte aligned) is copied to LDS with 128-bit
//     loads (dword-aligned loads realigned in registers when rows are not); pixels
//     outside the image get the neutral value (0 for dilation, 255 for erosion).
//  2. Vertical pass: a thread owns a 4-column quad and a run of K output rows.  It reads
//     K+2r words down the column, splits every word into two u16x2 whose high bytes hold
//     columns 0/2 and 1/3 (one shift)
//     and computes the windowed max with a single-split van Herk: suffix maxima over the
//     first K rows, running prefix over the rest, one v_pk_max_u16 per two pixels.
//     Results are written TRANSPOSED into a second LDS buffer (region column -> row).
//  3. Horizontal pass = the same vertical code over the transposed buffer; the 4x4 byte
//     transpose back to row order is done in registers (v_perm_b32) and the tile is
//     stored with 128-bit (or 32-bit for K < 16) global stores.
// K (run length) is the largest power of two <= min(2r, 32): the suffix covers rows
// [j, K) and the prefix rows [K, j+2r] of output j, which needs K <= 2r + 1.  For r = 1
// a direct 3-tap form with K = 4 is used.
// The transposed buffer T reuses the LDS of the region A (results of the vertical pass
// are held in registers across a barrier), so a workgroup needs max(|A|, |T|) bytes.
// The kernel is persistent (as many workgroups as fit on the chip, each looping over
// tiles).  Tiles are mapped XCD-aware: each XCD walks a contiguous tile range, so
// neighbouring tiles share their halo rows in that XCD's L2.
#include <utility>

#include "pm_common.hpp"
#include "sq_u8.hpp"
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

template <int K, bool DIL, bool ALIGNED, int RC>
__global__ __launch_bounds__(FTHREADS) void sq_u8_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                         int H, int W, int r_rt, int tiles_x, int tiles_per_img,
                                                         int ntiles)
{
  constexpr int TW = sq_tw(K, RC);
  extern __shared__ __align__(16) uint8_t smem[];
  // A tile's region is read with all its loads in flight, then copied to LDS.  The
  // workgroups resident on a CU overlap each other's loads and compute; prefetching the
  // next tile into registers instead was slower (r = 1: 34.9 vs 33.5 us, r = 7: 40.2 vs
  // 36.5 us, r = 31: 55.7 vs 52.1 us on 8192x8192), its registers costing occupancy.
  const TileSched ts = tile_sched(blockIdx.x, gridDim.x, ntiles);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int img = tile / tiles_per_img;
    const int t_in = tile - img * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * FT, x0 = tx * TW;
    const int64_t ibase = (int64_t)img * H * W;
    sq_tile_load<K, DIL, ALIGNED, RC, TW>(smem, in + ibase, H, W, r_rt, y0, x0);
    __syncthreads();
    sq_tile_passes<K, DIL, ALIGNED, RC, TW>(smem, out + ibase, H, W, r_rt, y0, x0);
    __syncthreads();  // A/T are rewritten by the next tile
  }
}

typedef void (*sq_kernel_t)(const uint8_t *, uint8_t *, int, int, int, int, int, int);

// Kernel for radius r (1..32), specialised on r, for 16-byte aligned rows or any rows.
template <bool DIL, int... RS>
static sq_kernel_t sq_pick(int r, bool aligned, std::integer_sequence<int, RS...>)
{
  static constexpr sq_kernel_t spec[] = {sq_u8_kernel<sq_k_of(RS + 1), DIL, true, RS + 1>...};
  static constexpr sq_kernel_t spec_any[] = {sq_u8_kernel<sq_k_of(RS + 1), DIL, false, RS + 1>...};
  return aligned ? spec[r - 1] : spec_any[r - 1];
}

template <bool DIL>
static int sq_u8_once(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int r, hipStream_t s)
{
  const bool aligned = (W % 16) == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0;
  const int TW = sq_tw(sq_k_of(r), r);
  const int tiles_x = (int)((W + TW - 1) / TW), tiles_y = (int)((H + FT - 1) / FT);
  const int64_t per = (int64_t)tiles_x * tiles_y;
  const int64_t ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused square kernel");
  const size_t lds = sq_tile_lds_bytes(r, TW);
  int dev = 0, cus = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const sq_kernel_t k = sq_pick<DIL>(r, aligned, std::make_integer_sequence<int, (int)SQ_U8_RMAX>{});
  int occ = 0;
  PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, FTHREADS, lds));
  int64_t g = (int64_t)(occ > 0 ? occ : 1) * cus;
  if (g >= 8) g &= ~(int64_t)7;
  if (g > ntiles) g = ntiles;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(FTHREADS), lds, s, in, out, (int)H, (int)W, r, tiles_x, (int)per,
                     (int)ntiles);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

bool sq_u8_applicable(const void *, const void *, int64_t H, int64_t W)
{
  return W > 0 && H > 0 && H <= INT32_MAX / 2 && W <= INT32_MAX / 2;
}

template <bool DIL>
int sq_u8(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int64_t r, hipStream_t s)
{
  const int64_t n = B * H * W;
  if (n == 0) return PM_OK;
  if (r <= 0) {
    PM_HIP(hipMemcpyAsync(out, in, (size_t)n, hipMemcpyDeviceToDevice, s));
    return PM_OK;
  }
  const int64_t span = H > W ? H : W;
  if (r > span) r = span;  // a window wider than the image is the whole image (clipped)
  if (r <= SQ_U8_RMAX) return sq_u8_once<DIL>(in, out, B, H, W, (int)r, s);
  // square of radius r = successive squares of radii summing to r (Minkowski sum; clipping
  // at the border commutes with it on a box)
  int nsteps = (int)((r + SQ_U8_RMAX - 1) / SQ_U8_RMAX);
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n, s));
  uint8_t *bufs[2] = {out, (uint8_t *)tmp.ptr};
  int dst = (nsteps % 2 == 1) ? 0 : 1;
  const uint8_t *src = in;
  int64_t left = r;
  for (int i = 0; i < nsteps; i++) {
    const int ri = (int)(left < SQ_U8_RMAX ? left : SQ_U8_RMAX);
    left -= ri;
    PM_TRY(sq_u8_once<DIL>(src, bufs[dst], B, H, W, ri, s));
    src = bufs[dst];
    dst ^= 1;
  }
  return PM_OK;
}

template int sq_u8<true>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, int64_t, hipStream_t);
template int sq_u8<false>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, int64_t, hipStream_t);

}  // namespace pm
