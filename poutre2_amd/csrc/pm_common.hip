// pm_common.hip -- error state, argument checks and scratch allocation shared by all entry points.
#include <atomic>
#include <mutex>

#include "pm_common.hpp"

#include <cstdarg>
#include <cstring>

namespace pm {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

bool same_shape(const pm_image *a, const pm_image *b)
{
  if (a->rank != b->rank) return false;
  for (int i = 0; i < a->rank; i++)
    if (a->shape[i] != b->shape[i]) return false;
  return true;
}

int check_image(const pm_image *im, const char *what)
{
  if (!im) PM_FAIL(PM_E_ARG, "%s: null image descriptor", what);
  if (im->rank < 1 || im->rank > 3) PM_FAIL(PM_E_RANK, "%s: unsupported number of dims %d", what, im->rank);
  if (ptype_size(im->ptype) == 0) PM_FAIL(PM_E_PTYPE, "%s: unsupported PType %d", what, im->ptype);
  for (int i = 0; i < im->rank; i++)
    if (im->shape[i] < 0) PM_FAIL(PM_E_SIZE, "%s: negative extent", what);
  if (!im->data && numel(*im) > 0) PM_FAIL(PM_E_ARG, "%s: null data pointer", what);
  return PM_OK;
}

int check_pair(const pm_image *a, const pm_image *b, const char *what)
{
  if (!a || !b) PM_FAIL(PM_E_ARG, "%s: null image descriptor", what);
  if (!same_shape(a, b)) PM_FAIL(PM_E_SIZE, "%s images have not compatible sizes", what);
  if (a->ptype != b->ptype) PM_FAIL(PM_E_PTYPE, "%s images must have compatible types", what);
  return PM_OK;
}

int check_unary(const pm_image *in, const pm_image *out, const char *what)
{
  clear_error();
  PM_TRY(check_pair(in, out, what));
  if (in->data == out->data && numel(*in) > 0)
    PM_FAIL(PM_E_ALIAS, "%s images input ouput images must be different", what);
  PM_TRY(check_image(in, what));
  return PM_OK;
}

int check_binary(const pm_image *a, const pm_image *b, const pm_image *out, const char *what)
{
  clear_error();
  PM_TRY(check_pair(a, b, what));
  PM_TRY(check_pair(b, out, what));
  if (numel(*out) > 0 && (a->data == out->data || b->data == out->data))
    PM_FAIL(PM_E_ALIAS, "%s output must be != than input images", what);
  PM_TRY(check_image(a, what));
  return PM_OK;
}

// Scratch memory comes from one memory pool per device, owned by the library, whose
// release threshold is unlimited: memory freed by a call stays in the pool for the next
// one instead of being returned to the driver at every synchronisation (which made every
// call with a scratch image pay a fresh mapping).  The pool holds at most the peak scratch
// of the calls made so far on that device, until the process ends.
static hipMemPool_t scratch_pool(int dev, hipError_t &err)
{
  static std::mutex mu;
  static hipMemPool_t pools[PM_MAX_DEVICES] = {};
  err = hipSuccess;
  if (dev < 0 || dev >= PM_MAX_DEVICES) {
    err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!pools[dev]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t p = nullptr;
    err = hipMemPoolCreate(&p, &props);
    if (err != hipSuccess) return nullptr;
    uint64_t keep = UINT64_MAX;
    err = hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep);
    if (err != hipSuccess) {
      (void)hipMemPoolDestroy(p);
      return nullptr;
    }
    pools[dev] = p;
  }
  return pools[dev];
}

static std::atomic<int> g_max_rounds{0}, g_fail_allocs{0};

int round_cap(int own_cap)
{
  const int c = g_max_rounds.load(std::memory_order_relaxed);
  return (c > 0 && c < own_cap) ? c : own_cap;
}

bool take_forced_alloc_failure()
{
  int n = g_fail_allocs.load(std::memory_order_relaxed);
  while (n > 0)
    if (g_fail_allocs.compare_exchange_weak(n, n - 1)) return true;
  return false;
}

int pinned_counters(int **h)
{
  static thread_local int *counters = nullptr;
  if (!counters) PM_HIP(hipHostMalloc((void **)&counters, sizeof(int) * (PM_COUNTERS + 1), hipHostMallocDefault));
  *h = counters;
  return PM_OK;
}

int Scratch::alloc(size_t bytes, hipStream_t s)
{
  stream = s;
  if (bytes == 0) bytes = 256;
  if (take_forced_alloc_failure())
    PM_FAIL(PM_E_HIP, "scratch allocation of %zu bytes failed: out of device memory (forced by the testing hook)",
            bytes);
  int dev = 0;
  PM_HIP(hipGetDevice(&dev));
  hipError_t e = hipSuccess;
  hipMemPool_t pool = scratch_pool(dev, e);
  if (pool) e = hipMallocFromPoolAsync(&ptr, bytes, pool, s);
  if (e != hipSuccess) {
    ptr = nullptr;
    PM_FAIL(PM_E_HIP, "scratch allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  return PM_OK;
}

Scratch::~Scratch()
{
  if (ptr) (void)hipFreeAsync(ptr, stream);
}

// byte fill with 16-byte stores where a 16-byte block of the buffer is whole (a kernel, not
// hipMemsetAsync, so that callers may capture it in a graph)
__global__ void fill_u8_kernel(uint8_t *p, uint8_t v, uint64_t n)
{
  const uint64_t lead = (uint64_t)((16 - ((uintptr_t)p & 15)) & 15);  // bytes before the first aligned block
  const uint64_t nblk = n > lead ? (n - lead) / 16 : 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint32_t w = v * 0x01010101u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nblk; i += stride)
    reinterpret_cast<uint4 *>(p + lead)[i] = make_uint4(w, w, w, w);
  // the unaligned head and the tail
  const uint64_t tail0 = lead + nblk * 16;
  const uint64_t nedge = (lead < n ? lead : n) + (n - (tail0 < n ? tail0 : n));
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nedge; i += stride) {
    const uint64_t h = lead < n ? lead : n;
    p[i < h ? i : tail0 + (i - h)] = v;
  }
}

int fill_bytes(void *p, uint8_t v, size_t n, hipStream_t s)
{
  if (n == 0) return PM_OK;
  hipLaunchKernelGGL(fill_u8_kernel, dim3(grid1d((int64_t)(n / 16 + 1), 256, 4096)), dim3(256), 0, s, (uint8_t *)p, v,
                     (uint64_t)n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

}  // namespace pm

extern "C" const char *pm_last_error(void) { return pm::g_err; }

extern "C" const char *pm_version(void) { return "poutre2_amd 0.1 (gfx950)"; }

namespace pm {
static int set_hook(std::atomic<int> &h, int n, const char *what)
{
  clear_error();
  if (n < 0) PM_FAIL(PM_E_ARG, "%s: negative count", what);
  h.store(n);
  return PM_OK;
}
}  // namespace pm

extern "C" int pm_debug_set_max_rounds(int n) { return pm::set_hook(pm::g_max_rounds, n, "pm_debug_set_max_rounds"); }
extern "C" int pm_debug_fail_allocations(int n)
{ return pm::set_hook(pm::g_fail_allocs, n, "pm_debug_fail_allocations"); }
