// batch.hip -- many uint8 images of uneven sizes and mixed square sizes, few launches.
//
// pm_batch_dilate_open computes, for every image i of a batch,
//   dil[i]  = Dilate(in[i], SESquare2D, r_i)                      (ero_dil_static.cpp:268-305)
//   open[i] = Dilate(Erode(in[i], SESquare2D, r_i), SESquare2D, r_i)
// The images are grouped by radius.  All groups up to r = BATCH_FUSED_RMAX (15) run in ONE
// persistent launch (batch_one_kernel) whose work list spans all their images: per group the
// fused opening tiles (sq_open.hpp: erosion and dilation in LDS, nothing in between goes
// through memory), then the dilation tiles.  Larger radii take two launches per group: items
// for the dilation and for the erosion into a scratch image, then the dilation of the eroded
// tiles (stream order separates them).  Every item runs the tile code of the single-image
// fused kernels specialised for its group's radius, with loads and stores for any width and
// alignment.
// The reference runs the two Erode/Dilate calls per image, one image at a time.
#include <cstring>
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


// A batch laid out for launching: images grouped by radius, tile counts, descriptor table
// (host copy) and the scratch bytes the eroded images need.
struct BatchLayout {
  static constexpr int NG = SQ_TILE_RMAX;  // one group per radius 1..32
  std::vector<BatchImage> grp[NG];
  int64_t tiles[NG] = {};
  int64_t otiles[NG] = {};  // fused-opening tiles (groups r <= BATCH_FUSED_RMAX)
  int64_t pixels = 0;
  int64_t tmp_bytes = 0;
  size_t first[NG] = {};
  std::vector<BatchImage> all;
  std::vector<BatchGroup> one;  // groups of the one-launch kernel, in work-list order
  int64_t one_items = 0;
  size_t one_lds = 0;
  BatchGroup *dgrp = nullptr;  // device copy of `one` (scratch)
  int *queue = nullptr;        // the one-launch kernel's queue counters (scratch, 2 ints)
};

static int batch_layout(int n, const pm_image *ins, const int *radii, pm_image *dil_outs, pm_image *open_outs,
                        BatchLayout &L)
{
  if (n < 0 || (n > 0 && (!ins || !radii || !dil_outs || !open_outs))) PM_FAIL(PM_E_ARG, "batch: bad arguments");
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
    b.tile0 = (int)L.tiles[g];
    b.otile0 = 0;
    L.tiles[g] += (int64_t)b.tiles_x * ((H + FT - 1) / FT);
    if (radii[i] <= BATCH_FUSED_RMAX) {
      // fused opening: tiles of open_fo(r) rows, same columns as the dilation tiles
      const int fo = open_fo(radii[i]);
      b.otile0 = (int)L.otiles[g];
      L.otiles[g] += (int64_t)b.tiles_x * ((H + fo - 1) / fo);
    } else {
      L.tmp_bytes += (H * W + 255) & ~(int64_t)255;
    }
    if (L.tiles[g] + L.otiles[g] > INT32_MAX / 4) PM_FAIL(PM_E_SIZE, "batch: too many tiles");
    int64_t fused_items = 0;  // the one-launch kernel's work list spans all groups up to BATCH_FUSED_RMAX
    for (int k = 0; k < BATCH_FUSED_RMAX; k++) fused_items += L.tiles[k] + L.otiles[k];
    if (fused_items > INT32_MAX / 4) PM_FAIL(PM_E_SIZE, "batch: too many tiles");
    L.pixels += H * W;
    L.grp[g].push_back(b);
  }
  return PM_OK;
}

// Scratch layout: eroded images (256-byte aligned each), then the upload area: the
// descriptor table, the group table of the one-launch kernel and its two queue counters
// (each part 256-byte aligned).
static size_t batch_desc_bytes(const BatchLayout &L)
{
  size_t d = 0;
  for (int g = 0; g < BatchLayout::NG; g++) d += L.grp[g].size() * sizeof(BatchImage);
  return (d + 255) & ~(size_t)255;
}
constexpr size_t BATCH_GRP_BYTES = (BatchLayout::NG * sizeof(BatchGroup) + 255) & ~(size_t)255;
static size_t batch_upload_bytes(const BatchLayout &L) { return batch_desc_bytes(L) + BATCH_GRP_BYTES + 2 * sizeof(int); }
static size_t batch_scratch_bytes(const BatchLayout &L) { return (size_t)L.tmp_bytes + batch_upload_bytes(L); }

// point the descriptors at scratch `base` and build the flat host table
static BatchImage *batch_bind(BatchLayout &L, uint8_t *base)
{
  int64_t off = 0;
  for (int g = 0; g < BatchLayout::NG; g++)
    for (auto &b : L.grp[g]) {
      if (b.r <= BATCH_FUSED_RMAX) continue;  // fused opening: no eroded image
      b.tmp = base + off;
      off += ((int64_t)b.H * b.W + 255) & ~(int64_t)255;
    }
  L.all.clear();
  for (int g = 0; g < BatchLayout::NG; g++) {
    L.first[g] = L.all.size();
    L.all.insert(L.all.end(), L.grp[g].begin(), L.grp[g].end());
  }
  // the one-launch work list: groups of decreasing radius (heaviest tiles first)
  L.one.clear();
  L.one_items = 0;
  L.one_lds = 0;
  for (int g = BATCH_FUSED_RMAX - 1; g >= 0; g--) {
    if (L.grp[g].empty()) continue;
    BatchGroup G;
    G.r = g + 1;
    G.img0 = (int)L.first[g];
    G.nimg = (int)L.grp[g].size();
    G.item0 = (int)L.one_items;
    G.nopen = (int)L.otiles[g];
    G.nitems = (int)(L.otiles[g] + L.tiles[g]);
    L.one.push_back(G);
    L.one_items += G.nitems;
    const size_t lds = batch_one_lds_bytes(G.r);
    if (lds > L.one_lds) L.one_lds = lds;
  }
  uint8_t *up = base + L.tmp_bytes;
  L.dgrp = (BatchGroup *)(up + batch_desc_bytes(L));
  L.queue = (int *)(up + batch_desc_bytes(L) + BATCH_GRP_BYTES);
  return (BatchImage *)up;
}

// host image of the upload area (descriptor table, group table, zeroed queue counters)
static std::vector<uint8_t> batch_upload(const BatchLayout &L)
{
  std::vector<uint8_t> h(batch_upload_bytes(L), 0);
  if (!L.all.empty()) memcpy(h.data(), L.all.data(), L.all.size() * sizeof(BatchImage));
  if (!L.one.empty()) memcpy(h.data() + batch_desc_bytes(L), L.one.data(), L.one.size() * sizeof(BatchGroup));
  return h;
}

// Enqueue the batch (kernel launches only, so that the sequence can be captured into a
// graph): the one-launch kernel over every group up to BATCH_FUSED_RMAX, whose queue counters
// must be zero (they are uploaded as zeros and the kernel resets them when it ends), and two
// launches per larger radius.  Launches go to the streams ss[0 .. ns-1] round robin: the
// two-launch groups are independent of the rest, so on several streams they overlap.
static int batch_enqueue(const BatchLayout &L, const BatchImage *dimg, const hipStream_t *ss, int ns)
{
  int dev = 0, cus = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int j = 0;
  if (L.one_items > 0) {
    const hipStream_t s = ss[j++ % ns];
    int occ = 0;
    PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_one_kernel, FTHREADS, L.one_lds));
    int64_t G = (int64_t)(occ > 0 ? occ : 1) * cus;
    if (G > L.one_items) G = L.one_items;
    hipLaunchKernelGGL(batch_one_kernel, dim3((unsigned)G), dim3(FTHREADS), L.one_lds, s, dimg, L.dgrp,
                       (int)L.one.size(), (int)L.one_items, L.queue);
    PM_HIP(hipGetLastError());
  }
  for (int g = BATCH_FUSED_RMAX; g < BatchLayout::NG; g++) {
    if (L.grp[g].empty()) continue;
    const hipStream_t s = ss[j++ % ns];
    const int r = g + 1;
    for (int phase = 0; phase < 2; phase++) {
      const batch_kernel_t k = batch_kernel(r, phase);
      const size_t lds = sq_tile_lds_bytes(r, sq_tw(sq_k_of(r), r));
      int occ = 0;
      PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, FTHREADS, lds));
      const int64_t items = phase == 0 ? 2 * L.tiles[g] : L.tiles[g];
      int64_t G = (int64_t)(occ > 0 ? occ : 1) * cus;
      if (G > items) G = items;
      hipLaunchKernelGGL(k, dim3((unsigned)G), dim3(FTHREADS), lds, s, dimg + L.first[g], (int)L.grp[g].size(),
                         (int)items);
      PM_HIP(hipGetLastError());
    }
  }
  return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" int pm_batch_dilate_open(int n, const pm_image *ins, const int *radii, pm_image *dil_outs,
                                    pm_image *open_outs, void *stream)
{ PM_TRACE();
  clear_error();
  BatchLayout L;
  PM_TRY(batch_layout(n, ins, radii, dil_outs, open_outs, L));
  if (L.pixels == 0) return PM_OK;
  hipStream_t s = as_stream(stream);
  Scratch sc;
  PM_TRY(sc.alloc(batch_scratch_bytes(L), s));
  BatchImage *dimg = batch_bind(L, (uint8_t *)sc.ptr);
  const std::vector<uint8_t> up = batch_upload(L);
  PM_HIP(hipMemcpyAsync(dimg, up.data(), up.size(), hipMemcpyHostToDevice, s));
  return batch_enqueue(L, dimg, &s, 1);
}

// A batch captured once into a graph: its scratch and descriptor table are allocated at
// creation and live with the plan, so a launch is one hipGraphLaunch with no allocation and
// no host-side work per image.
struct pm_batch_plan {
  int device = 0;
  void *scratch = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t pixels = 0;
};

static void plan_free(pm_batch_plan *p)
{
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->scratch) (void)hipFree(p->scratch);
  delete p;
}

extern "C" int pm_batch_plan_create(int n, const pm_image *ins, const int *radii, pm_image *dil_outs,
                                    pm_image *open_outs, pm_batch_plan **plan)
{ PM_TRACE();
  clear_error();
  if (!plan) PM_FAIL(PM_E_ARG, "batch plan: null plan pointer");
  *plan = nullptr;
  BatchLayout L;
  PM_TRY(batch_layout(n, ins, radii, dil_outs, open_outs, L));
  pm_batch_plan *p = new pm_batch_plan();
  {
    const hipError_t ge = hipGetDevice(&p->device);
    if (ge != hipSuccess) {
      plan_free(p);
      PM_FAIL(PM_E_HIP, "batch plan: hipGetDevice failed: %s", hipGetErrorString(ge));
    }
  }
  const size_t bytes = batch_scratch_bytes(L);
  hipError_t e = hipMalloc(&p->scratch, bytes ? bytes : 256);
  if (e != hipSuccess) {
    plan_free(p);
    PM_FAIL(PM_E_HIP, "batch plan: scratch allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  BatchImage *dimg = batch_bind(L, (uint8_t *)p->scratch);
  {
    const std::vector<uint8_t> up = batch_upload(L);
    e = hipMemcpy(dimg, up.data(), up.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      plan_free(p);
      PM_FAIL(PM_E_HIP, "batch plan: descriptor upload failed: %s", hipGetErrorString(e));
    }
  }
  // capture the launches on private streams (default priority): the capture stream forks
  // into NB branches (the one-launch kernel, then one per two-launch group, round robin)
  // that join back before the capture ends, so they are independent graph nodes that may run
  // concurrently
  constexpr int NB = 4;
  hipStream_t st[NB] = {};
  hipEvent_t fork = nullptr, join[NB] = {};
  for (int i = 0; i < NB && e == hipSuccess; i++) e = hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
  for (int i = 1; i < NB && e == hipSuccess; i++) e = hipEventCreateWithFlags(&join[i], hipEventDisableTiming);
  int rc = PM_OK;
  if (e == hipSuccess) e = hipStreamBeginCapture(st[0], hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t ec = hipEventRecord(fork, st[0]);
    for (int i = 1; i < NB && ec == hipSuccess; i++) ec = hipStreamWaitEvent(st[i], fork, 0);
    if (ec == hipSuccess) rc = batch_enqueue(L, dimg, st, NB);
    for (int i = 1; i < NB && ec == hipSuccess; i++) {
      ec = hipEventRecord(join[i], st[i]);
      if (ec == hipSuccess) ec = hipStreamWaitEvent(st[0], join[i], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st[0], &g);
    p->graph = g;
    if (ec != hipSuccess) e = ec;
    else if (rc == PM_OK && e2 != hipSuccess) e = e2;
  }
  if (fork) (void)hipEventDestroy(fork);
  for (int i = 0; i < NB; i++) {
    if (join[i]) (void)hipEventDestroy(join[i]);
    if (st[i]) (void)hipStreamDestroy(st[i]);
  }
  if (rc == PM_OK && e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (rc != PM_OK || e != hipSuccess) {
    plan_free(p);
    if (rc != PM_OK) return rc;
    PM_FAIL(PM_E_HIP, "batch plan: graph capture failed: %s", hipGetErrorString(e));
  }
  for (int g = 0; g < BatchLayout::NG; g++)
    for (const auto &b : L.grp[g]) p->pixels += (int64_t)b.H * b.W;
  *plan = p;
  return PM_OK;
}

extern "C" int pm_batch_plan_launch(pm_batch_plan *plan, void *stream)
{ PM_TRACE();
  clear_error();
  if (!plan || !plan->exec) PM_FAIL(PM_E_ARG, "batch plan: null plan");
  int dev = 0;
  PM_HIP(hipGetDevice(&dev));
  if (dev != plan->device) PM_FAIL(PM_E_ARG, "batch plan: created on device %d, launched on %d", plan->device, dev);
  PM_HIP(hipGraphLaunch(plan->exec, as_stream(stream)));
  return PM_OK;
}

extern "C" int pm_batch_plan_destroy(pm_batch_plan *plan)
{ PM_TRACE();
  clear_error();
  plan_free(plan);
  return PM_OK;
}
