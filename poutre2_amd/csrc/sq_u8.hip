// sq_u8.hip -- fused flat dilation / erosion of uint8 images by a (2r+1)^2 square, r <= 32.
//
// This is the kernel behind Dilate/Erode(SESquare2D, iter = r) and DilateX+DilateY for
// uint8 (ero_dil_static.cpp:229-305, ero_dil_line_se_t.hpp:49-280): one launch reads
// every input pixel once from HBM and writes every output once.
//
// (The 3 x 3 square on images with 16-byte aligned rows runs the register-streamed kernel of
// sq_u8_rs.hip instead.)
// One workgroup (256 threads) owns an output tile of TH = 64 (r <= 3) or 128 rows and TW = 128,
// 192 or 256 columns
// (per device and radius: the table pm_autotune fills, else the rule sq_tw).
//  1. The input region (TH+2r rows, TW+2*R16 columns, R16 = 16 for r <= 15 and 32 above,
//     so that global loads of aligned rows are 16-byte aligned) is copied to LDS with 128-bit
//     loads (at byte addresses when rows are not 16-byte aligned); pixels
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
// K (run length, sq_k_of) is 4 for r <= 3, 16 for r <= 15 and 32 above.  With K <= 2r + 1
// the suffix covers rows [j, K) and the prefix rows [K, j+2r] of output j; for 4 <= r <= 7
// the 16-row run is two blocks of 8 rows split at rows 8 and 16 (needs 2r >= 8).  For r = 1
// a direct 3-tap form with K = 4 is used.
// The transposed buffer T reuses the LDS of the region A (results of the vertical pass
// are held in registers across a barrier), so a workgroup needs max(|A|, |T|) bytes.
// The kernel is persistent (as many workgroups as fit on the chip, each looping over
// tiles).  Tiles are mapped XCD-aware: each XCD walks a contiguous tile range, so
// neighbouring tiles share their halo rows in that XCD's L2.
#include <atomic>
#include <utility>

#include "pm_common.hpp"
#include "sq_u8.hpp"
#include "sq_u8_kernel.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

// ---------------------------------------------------------------- tile-width table
// Tile width per device, alignment class and radius: 0 = not tuned (the built-in rule
// sq_tw), otherwise 128, 192 or 256 as chosen by pm_autotune (or set by pm_tuning_set).
static std::atomic<int> g_tw[PM_MAX_DEVICES][2][SQ_U8_RMAX + 1];

static int tuned_tw(int dev, bool aligned, int r)
{
  const int v = (dev >= 0 && dev < PM_MAX_DEVICES) ? g_tw[dev][aligned ? 1 : 0][r].load(std::memory_order_relaxed) : 0;
  return v ? v : sq_tw(sq_k_of(r), r);
}

static int sq_launch(bool dil, bool aligned, int tw, const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W,
                     int r, int cus, hipStream_t s)
{
  const int th = sq_th(r);
  const int tiles_x = (int)((W + tw - 1) / tw), tiles_y = (int)((H + th - 1) / th);
  const int64_t per = (int64_t)tiles_x * tiles_y;
  const int64_t ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused square kernel");
  const size_t lds = sq_tile_lds_bytes(r, tw, th);
  const sq_kernel_t k = dil ? (aligned ? sq_kernel_for<true, true>(r, tw) : sq_kernel_for<true, false>(r, tw))
                            : (aligned ? sq_kernel_for<false, true>(r, tw) : sq_kernel_for<false, false>(r, tw));
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

template <bool DIL>
static int sq_u8_once(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int r, hipStream_t s)
{
  if (r == 1 && sq_u8_rs_applicable(in, out, H, W)) return sq_u8_rs<DIL>(in, out, B, H, W, s);
  const bool aligned = (W % 16) == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0;
  int dev = 0, cus = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return sq_launch(DIL, aligned, tuned_tw(dev, aligned, r), in, out, B, H, W, r, cus, s);
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

// ---------------------------------------------------------------- autotuner
using namespace pm;

extern "C" int pm_autotune(void *stream)
{ PM_TRACE();
  clear_error();
  hipStream_t s = as_stream(stream);
  int dev = 0, cus = 0;
  PM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= PM_MAX_DEVICES) PM_FAIL(PM_E_HIP, "device %d outside the tuning table", dev);
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // three 8192 x 8192 input/output pairs used in turn (384 MiB, more than the MALL holds,
  // so the timings see HBM as a real workload does); the unaligned class uses 8192 x 8189
  // images at odd offsets.  Min/max work does not depend on the pixel values.
  constexpr int NP = 3;
  const int64_t H = 8192, W = 8192, Wu = 8189;
  const size_t bytes = (size_t)H * W + 64;
  uint8_t *buf[2 * NP] = {};
  int rc = PM_OK;
  for (int i = 0; i < 2 * NP && rc == PM_OK; i++) {
    const hipError_t e = hipMalloc((void **)&buf[i], bytes);
    if (e != hipSuccess) {
      buf[i] = nullptr;
      set_error("autotune: %s", hipGetErrorString(e));
      rc = PM_E_HIP;
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  for (int i = 0; i < NP && rc == PM_OK; i++) rc = fill_bytes(buf[i], 77, bytes, s);
  if (rc == PM_OK && hipEventCreate(&e0) != hipSuccess) rc = PM_E_HIP;
  if (rc == PM_OK && hipEventCreate(&e1) != hipSuccess) rc = PM_E_HIP;
  int turn = 0;
  for (int al = 1; al >= 0 && rc == PM_OK; al--) {
    const int64_t w = al ? W : Wu;
    auto one = [&](int tw, int r) {
      const int p = turn++ % NP;
      return sq_launch(true, al != 0, tw, al ? buf[p] : buf[p] + 1, al ? buf[NP + p] : buf[NP + p] + 3, 1, H, w, r,
                       cus, s);
    };
    for (int r = 1; r <= (int)SQ_U8_RMAX && rc == PM_OK; r++) {
      float best = 1e30f;
      int best_tw = 0;
      for (int tw : {128, 192, 256}) {
        rc = one(tw, r);  // warm-up
        if (rc != PM_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = PM_E_HIP; break; }
        for (int it = 0; it < 3 && rc == PM_OK; it++) rc = one(tw, r);
        if (rc != PM_OK) break;
        float ms = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = PM_E_HIP;
          break;
        }
        if (ms < best) {
          best = ms;
          best_tw = tw;
        }
      }
      if (rc == PM_OK) g_tw[dev][al][r].store(best_tw, std::memory_order_relaxed);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  for (int i = 0; i < 2 * NP; i++)
    if (buf[i]) (void)hipFree(buf[i]);
  if (rc != PM_OK && pm_last_error()[0] == 0) PM_FAIL(PM_E_HIP, "autotune: a timing call failed");
  return rc;
}

extern "C" int pm_tuning_get(int aligned, int r)
{
  clear_error();
  if (r < 1 || r > (int)SQ_U8_RMAX) PM_FAIL(PM_E_ARG, "tuning: radius %d outside [1, %d]", r, (int)SQ_U8_RMAX);
  int dev = 0;
  PM_HIP(hipGetDevice(&dev));
  return tuned_tw(dev, aligned != 0, r);
}

extern "C" int pm_tuning_set(int aligned, int r, int tile_width)
{
  clear_error();
  if (r < 1 || r > (int)SQ_U8_RMAX) PM_FAIL(PM_E_ARG, "tuning: radius %d outside [1, %d]", r, (int)SQ_U8_RMAX);
  if (tile_width != 0 && tile_width != 128 && tile_width != 192 && tile_width != 256)
    PM_FAIL(PM_E_ARG, "tuning: tile width must be 128, 192, 256 or 0 (built-in rule)");
  int dev = 0;
  PM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= PM_MAX_DEVICES) PM_FAIL(PM_E_HIP, "device %d outside the tuning table", dev);
  g_tw[dev][aligned ? 1 : 0][r].store(tile_width, std::memory_order_relaxed);
  return PM_OK;
}
