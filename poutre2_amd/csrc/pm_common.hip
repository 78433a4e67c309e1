// pm_common.hip -- error state, argument checks and scratch allocation shared by all entry points.
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

static bool same_shape(const pm_image *a, const pm_image *b)
{
  if (a->rank != b->rank) return false;
  for (int i = 0; i < a->rank; i++)
    if (a->shape[i] != b->shape[i]) return false;
  return true;
}

static int check_image(const pm_image *im, const char *what)
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

int Scratch::alloc(size_t bytes, hipStream_t s)
{
  stream = s;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMallocAsync(&ptr, bytes, s);
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

}  // namespace pm

extern "C" const char *pm_last_error(void) { return pm::g_err; }

extern "C" const char *pm_version(void) { return "poutre2_amd 0.1 (gfx950)"; }
