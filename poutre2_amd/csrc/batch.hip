// batch.hip -- many uint8 images of uneven sizes and mixed square sizes, few launches.
//
// pm_batch_dilate_open computes, for every image i of a batch,
//   dil[i]  = Dilate(in[i], SESquare2D, r_i)                      (ero_dil_static.cpp:268-305)
//   open[i] = Dilate(Erode(in[i], SESquare2D, r_i), SESquare2D, r_i)
// The images are grouped by radius.  Per group, one persistent launch walks a work list
// spanning all its images with two items per tile (the tile's dilation, and its erosion
// into a scratch image), and a second launch dilates the eroded tiles; stream order
// separates the two.  Every item runs the tile code of the single-image fused kernel
// (sq_u8.hip) specialised for the group's radius, with loads and stores for any width.
// The reference runs the two Erode/Dilate calls per image, one image at a time.
#include <vector>

#include "batch_tile.hpp"
#include "pm_common.hpp"

namespace pm {

static batch_kernel_t batch_kernel(int r, int phase)
{
  if (r <= 8) return batch_kernel_range<1>(r, phase);
  if (r <= 16) return batch_kernel_range<9>(r, phase);
  if (r <= 24) return batch_kernel_range<17>(r, phase);
  return batch_kernel_range<25>(r, phase);
}

}  // namespace pm


using namespace pm;

extern "C" int pm_batch_dilate_open(int n, const pm_image *ins, const int *radii, pm_image *dil_outs,
                                    pm_image *open_outs, void *stream)
{
  clear_error();
  if (n < 0 || (n > 0 && (!ins || !radii || !dil_outs || !open_outs))) PM_FAIL(PM_E_ARG, "batch: bad arguments");
  if (n == 0) return PM_OK;
  hipStream_t s = as_stream(stream);
  constexpr int NG = SQ_TILE_RMAX;  // one group per radius 1..32
  std::vector<BatchImage> grp[NG];
  int64_t tiles[NG] = {}, tmp_bytes = 0;
  for (int i = 0; i < n; i++) {
    PM_TRY(check_unary(&ins[i], &dil_outs[i], "batch dilate"));
    PM_TRY(check_unary(&ins[i], &open_outs[i], "batch open"));
    if (dil_outs[i].data == open_outs[i].data) PM_FAIL(PM_E_ALIAS, "batch: dilation and opening outputs must differ");
    if (ins[i].ptype != PM_PTYPE_U8 || ins[i].rank != 2) PM_FAIL(PM_E_UNSUPPORTED, "batch: uint8 2D images only");
    if (radii[i] < 1 || radii[i] > SQ_TILE_RMAX) PM_FAIL(PM_E_ARG, "batch: radius must be in [1, %d]", SQ_TILE_RMAX);
    const int64_t H = ins[i].shape[0], W = ins[i].shape[1];
    if (H > (1 << 28) || W > (1 << 28)) PM_FAIL(PM_E_SIZE, "batch: image too large");
    if (H == 0 || W == 0) continue;
    const int g = radii[i] - 1;
    const int TW = sq_tw(sq_k_of(radii[i]), radii[i]);
    BatchImage b;
    b.in = (const uint8_t *)ins[i].data;
    b.dil = (uint8_t *)dil_outs[i].data;
    b.open = (uint8_t *)open_outs[i].data;
    b.tmp = nullptr;
    b.H = (int)H;
    b.W = (int)W;
    b.r = radii[i];
    b.tiles_x = (int)((W + TW - 1) / TW);
    b.tile0 = (int)tiles[g];
    tiles[g] += (int64_t)b.tiles_x * ((H + FT - 1) / FT);
    if (tiles[g] > INT32_MAX / 4) PM_FAIL(PM_E_SIZE, "batch: too many tiles");
    tmp_bytes += (H * W + 255) & ~(int64_t)255;
    grp[g].push_back(b);
  }
  if (tmp_bytes == 0) return PM_OK;
  // scratch: eroded images (256-byte aligned each), then the descriptor tables
  size_t desc_off = (size_t)tmp_bytes, desc_bytes = 0;
  for (int g = 0; g < NG; g++) desc_bytes += grp[g].size() * sizeof(BatchImage);
  Scratch sc;
  PM_TRY(sc.alloc(desc_off + desc_bytes, s));
  uint8_t *base = (uint8_t *)sc.ptr;
  int64_t off = 0;
  for (int g = 0; g < NG; g++)
    for (auto &b : grp[g]) {
      b.tmp = base + off;
      off += ((int64_t)b.H * b.W + 255) & ~(int64_t)255;
    }
  std::vector<BatchImage> all;
  all.reserve(desc_bytes / sizeof(BatchImage));
  size_t first[NG];
  for (int g = 0; g < NG; g++) {
    first[g] = all.size();
    all.insert(all.end(), grp[g].begin(), grp[g].end());
  }
  BatchImage *dimg = (BatchImage *)(base + desc_off);
  PM_HIP(hipMemcpyAsync(dimg, all.data(), desc_bytes, hipMemcpyHostToDevice, s));
  int dev = 0, cus = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (int phase = 0; phase < 2; phase++)
    for (int g = 0; g < NG; g++) {
      if (grp[g].empty()) continue;
      const int r = g + 1;
      const batch_kernel_t k = batch_kernel(r, phase);
      const size_t lds = sq_tile_lds_bytes(r, sq_tw(sq_k_of(r), r));
      int occ = 0;
      PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, FTHREADS, lds));
      const int64_t items = phase == 0 ? 2 * tiles[g] : tiles[g];
      int64_t G = (int64_t)(occ > 0 ? occ : 1) * cus;
      if (G > items) G = items;
      hipLaunchKernelGGL(k, dim3((unsigned)G), dim3(FTHREADS), lds, s, dimg + first[g], (int)grp[g].size(),
                         (int)items);
      PM_HIP(hipGetLastError());
    }
  return PM_OK;
}
