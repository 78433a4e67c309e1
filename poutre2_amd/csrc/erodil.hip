// erodil.hip -- flat erosion / dilation on gfx950.
//
// Kernels
//  * nl_kernel      : max/min over an arbitrary neighbour list (the runtime-SE path,
//                     ero_dil_runtime_nl_se_t.hpp:48-97), out-of-bounds taps skipped.
//  * line_kernel    : max/min over a centred segment of 2r+1 along x or y, computed in an
//                     LDS tile (halo r on both sides) by van Herk / Gil-Werman block
//                     prefix and suffix maxima, three ops per pixel whatever r is: the
//                     reference's van Herk row passes (ero_dil_line_se_t.hpp:49-89) for
//                     every type, and its Y pass (:248-280) without the transposes.
// Host drivers map the reference's Common_NL_SE iterations onto these:
//  SESquare2D^k = (2k+1)^2 rectangle (x pass then y pass), segments^k = lines of radius k,
//  SESquare3D^k = box (x, y, z passes), crosses^k = k neighbour-list passes.
// uint8 fast paths: squares go to the fused kernel of sq_u8.hip; the z passes of boxes and
// z segments and the 3D cross go to the z-streaming kernels of vol_u8.hip.
#include "pm_common.hpp"
#include "erodil.hpp"
#include "sq_u8.hpp"
#include "sq_gen.hpp"
#include "vol_u8.hpp"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace pm {

// ============================================================ neighbour-list kernel
template <typename T, bool DIL>
__global__ __launch_bounds__(256) void nl_kernel(const T *__restrict__ in, T *__restrict__ out, int D, int H, int W,
                                                 int xblocks, NLArgs a)
{
  using L = Lat<T, DIL>;
  const int row = blockIdx.x / xblocks;  // row index over D*H
  const int x = (blockIdx.x - row * xblocks) * 256 + threadIdx.x;
  if (x >= W) return;
  const int z = row / H, y = row - z * H;
  T acc = L::neutral();
  for (int k = 0; k < a.n; k++) {
    const int zz = z + a.dz[k], yy = y + a.dy[k], xx = x + a.dx[k];
    if ((unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W)
      acc = L::op(acc, in[((int64_t)zz * H + yy) * W + xx]);
  }
  out[((int64_t)z * H + y) * W + x] = acc;
}

template <typename T, bool DIL>
int nl_pass(const T *in, T *out, Dims d, const NLArgs &a, hipStream_t s)
{
  if (d.D * d.H * d.W == 0) return PM_OK;
  const int xblocks = (int)((d.W + 255) / 256);
  const int64_t blocks = (int64_t)xblocks * d.D * d.H;
  if (blocks > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the neighbour-list kernel");
  hipLaunchKernelGGL((nl_kernel<T, DIL>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, (int)d.D, (int)d.H,
                     (int)d.W, xblocks, a);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

// ============================================================ line kernel
// Problem view: B batches of an H x W image.  AXIS 0 = window along x, AXIS 1 = along y.
// A block owns S outputs along the axis times C lines across it.  Its LDS region is
// (S + 2r) x C elements, loaded once with the neutral value outside the image.
constexpr int LINE_LDS_BYTES = 32768;  // per ping-pong buffer

template <typename T, bool DIL, int AXIS>
__global__ __launch_bounds__(256) void line_kernel(const T *__restrict__ in, T *__restrict__ out, int H, int W, int r,
                                                   int S, int C, int nb_along, int nb_across)
{
  using L = Lat<T, DIL>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Lr = S + 2 * r;  // region length along the axis
  T *buf0 = reinterpret_cast<T *>(smem);
  T *buf1 = buf0 + (size_t)Lr * C;

  int bid = blockIdx.x;
  const int b_along = bid % nb_along;
  bid /= nb_along;
  const int b_across = bid % nb_across;
  const int batch = bid / nb_across;
  const int64_t base = (int64_t)batch * H * W;

  // region element (i along, c across) lives at buf[i * C + c] for AXIS 1 and at
  // buf[c * Lr + i] for AXIS 0, i.e. always row-major in image order.
  const int n_reg = Lr * C;
  if (AXIS == 0) {
    const int y0 = b_across * C, x0 = b_along * S - r;
    for (int e = threadIdx.x; e < n_reg; e += 256) {
      const int c = e / Lr, i = e - c * Lr;
      const int y = y0 + c, x = x0 + i;
      buf0[e] = (y < H && (unsigned)x < (unsigned)W) ? in[base + (int64_t)y * W + x] : L::neutral();
    }
  } else {
    const int y0 = b_along * S - r, x0 = b_across * C;
    for (int e = threadIdx.x; e < n_reg; e += 256) {
      const int i = e / C, c = e - i * C;
      const int y = y0 + i, x = x0 + c;
      buf0[e] = ((unsigned)y < (unsigned)H && x < W) ? in[base + (int64_t)y * W + x] : L::neutral();
    }
  }
  __syncthreads();

  // van Herk / Gil-Werman: cut the region into blocks of w = 2r+1 taps along the axis; g is
  // the running max from each block's start (into buf1), h the running max from each
  // block's end (in place in buf0); every window [j, j+2r] spans at most two blocks, so
  // out[j] = op(h[j], g[j + 2r]).  Three ops per output whatever r is.
  const int w = 2 * r + 1;
  const int strideA = (AXIS == 0) ? 1 : C;   // step along the axis
  const int strideC = (AXIS == 0) ? Lr : 1;  // step across
  {
    const int nblk = (Lr + w - 1) / w;
    const int ntask = nblk * C;
    for (int e = threadIdx.x; e < ntask; e += 256) {
      const int b = e / C, c = e - b * C;
      const int i0 = b * w, i1 = min(i0 + w, Lr);
      const int oc = c * strideC;
      T acc = L::neutral();
      for (int i = i0; i < i1; i++) {
        const int o = i * strideA + oc;
        acc = L::op(acc, buf0[o]);
        buf1[o] = acc;
      }
      acc = L::neutral();
      for (int i = i1 - 1; i >= i0; i--) {
        const int o = i * strideA + oc;
        acc = L::op(acc, buf0[o]);
        buf0[o] = acc;
      }
    }
  }
  __syncthreads();
  const int sh = (w - 1) * strideA;
  const int n_out = S * C;
  if (AXIS == 0) {
    const int y0 = b_across * C, x0 = b_along * S;
    for (int e = threadIdx.x; e < n_out; e += 256) {
      const int c = e / S, j = e - c * S;
      const int y = y0 + c, x = x0 + j;
      if (y < H && x < W) {
        const int o = j * strideA + c * strideC;
        out[base + (int64_t)y * W + x] = L::op(buf0[o], buf1[o + sh]);
      }
    }
  } else {
    const int y0 = b_along * S, x0 = b_across * C;
    for (int e = threadIdx.x; e < n_out; e += 256) {
      const int j = e / C, c = e - j * C;
      const int y = y0 + j, x = x0 + c;
      if (y < H && x < W) {
        const int o = j * strideA + c * strideC;
        out[base + (int64_t)y * W + x] = L::op(buf0[o], buf1[o + sh]);
      }
    }
  }
}

// Tile geometry of one line pass: returns false if r is too large for one pass.
template <typename T> static bool line_geometry(int axis, int64_t r, int64_t H, int64_t W, int &S, int &C)
{
  const int64_t elems = LINE_LDS_BYTES / (int64_t)sizeof(T);
  if (axis == 0) {
    S = 256;
    if (W < S) S = (int)std::max<int64_t>(W, 1);
    const int64_t rowlen = S + 2 * r;
    if (rowlen > elems) return false;
    int64_t c = elems / rowlen;
    c = std::min<int64_t>(c, 64);
    c = std::min<int64_t>(c, std::max<int64_t>(H, 1));
    C = (int)std::max<int64_t>(c, 1);
  } else {
    C = (int)std::min<int64_t>(256 / (int64_t)sizeof(T), std::max<int64_t>(W, 1));
    S = 32;
    if (H < S) S = (int)std::max<int64_t>(H, 1);
    if ((S + 2 * r) * C > elems) return false;
  }
  return true;
}

template <typename T> static int64_t line_rmax(int axis)
{
  const int64_t elems = LINE_LDS_BYTES / (int64_t)sizeof(T);
  if (axis == 0) return (elems - 256) / 2;
  const int64_t C = 256 / (int64_t)sizeof(T);
  return (elems / C - 32) / 2;
}

template <typename T, bool DIL>
static int line_once(const T *in, T *out, int64_t B, int64_t H, int64_t W, int axis, int64_t r, hipStream_t s)
{
  int S = 0, C = 0;
  if (!line_geometry<T>(axis, r, H, W, S, C)) PM_FAIL(PM_E_ARG, "line pass radius %lld too large", (long long)r);
  const int64_t along = axis == 0 ? W : H, across = axis == 0 ? H : W;
  const int64_t nb_along = (along + S - 1) / S, nb_across = (across + C - 1) / C;
  const int64_t blocks = nb_along * nb_across * B;
  if (blocks > INT32_MAX || H > INT32_MAX || W > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for a line pass");
  const size_t lds = 2 * (size_t)(S + 2 * r) * C * sizeof(T);
  if (axis == 0)
    hipLaunchKernelGGL((line_kernel<T, DIL, 0>), dim3((unsigned)blocks), dim3(256), lds, s, in, out, (int)H, (int)W,
                       (int)r, S, C, (int)nb_along, (int)nb_across);
  else
    hipLaunchKernelGGL((line_kernel<T, DIL, 1>), dim3((unsigned)blocks), dim3(256), lds, s, in, out, (int)H, (int)W,
                       (int)r, S, C, (int)nb_along, (int)nb_across);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <typename T, bool DIL>
int line_pass(const T *in, T *out, int64_t B, int64_t H, int64_t W, int axis, int64_t r, hipStream_t s)
{
  const int64_t n = B * H * W;
  if (n == 0) return PM_OK;
  const int64_t along = axis == 0 ? W : H;
  if (r <= 0) {
    PM_HIP(hipMemcpyAsync(out, in, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, s));
    return PM_OK;
  }
  // a window wider than twice the line is the whole line; clipping keeps the result
  if (r > along) r = along;
  const int64_t rmax = line_rmax<T>(axis);
  if (r <= rmax) return line_once<T, DIL>(in, out, B, H, W, axis, r, s);
  // a segment of radius r is the Minkowski sum of segments of radii r1 + r2 + ... = r, and
  // clipping at the border commutes with that decomposition on an interval
  std::vector<int64_t> steps;
  for (int64_t left = r; left > 0; left -= rmax) steps.push_back(std::min(left, rmax));
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * sizeof(T), s));
  T *bufs[2] = {out, reinterpret_cast<T *>(tmp.ptr)};
  // choose the first destination so that the last step writes `out`
  int dst = (steps.size() % 2 == 1) ? 0 : 1;
  const T *src = in;
  for (size_t i = 0; i < steps.size(); i++) {
    PM_TRY((line_once<T, DIL>(src, bufs[dst], B, H, W, axis, steps[i], s)));
    src = bufs[dst];
    dst ^= 1;
  }
  return PM_OK;
}

// rectangle (2ry+1) x (2rx+1) on B images of H x W: x pass into scratch, y pass into out
template <typename T, bool DIL>
int rect_pass(const T *in, T *out, int64_t B, int64_t H, int64_t W, int64_t ry, int64_t rx, hipStream_t s)
{
  if (ry <= 0) return line_pass<T, DIL>(in, out, B, H, W, 0, rx, s);
  if (rx <= 0) return line_pass<T, DIL>(in, out, B, H, W, 1, ry, s);
  if constexpr (std::is_same<T, uint8_t>::value) {
    if (ry == rx && sq_u8_applicable(in, out, H, W)) return sq_u8<DIL>(in, out, B, H, W, ry, s);
  } else {
    if (ry == rx && sq_gen_applicable(H, W)) return sq_gen<T, DIL>(in, out, B, H, W, ry, s);
  }
  const int64_t n = B * H * W;
  if (n == 0) return PM_OK;
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * sizeof(T), s));
  T *t = reinterpret_cast<T *>(tmp.ptr);
  PM_TRY((line_pass<T, DIL>(in, t, B, H, W, 0, rx, s)));
  PM_TRY((line_pass<T, DIL>(t, out, B, H, W, 1, ry, s)));
  return PM_OK;
}

// window of 2rz+1 planes along z on a D x H x W volume
template <typename T, bool DIL>
static int zline_pass(const T *in, T *out, Dims d, int64_t rz, hipStream_t s)
{
  if constexpr (std::is_same<T, uint8_t>::value) {
    const int64_t k = rz < d.D ? rz : d.D;  // a window longer than the volume is the whole column
    if (k >= 1 && k <= VOL_ZWIN_KMAX && vol_u8_applicable(in, out, d.D, d.H, d.W))
      return zwin_u8<DIL>(in, out, d.D, d.H, d.W, (int)k, s);
  }
  // z pass = y pass on the (D, H*W) view
  return line_pass<T, DIL>(in, out, 1, d.D, d.H * d.W, 1, rz, s);
}

// box (2rz+1) x (2ry+1) x (2rx+1) on a D x H x W volume
template <typename T, bool DIL>
int box_pass(const T *in, T *out, Dims d, int64_t rz, int64_t ry, int64_t rx, hipStream_t s)
{
  if (rz <= 0) return rect_pass<T, DIL>(in, out, d.D, d.H, d.W, ry, rx, s);
  if constexpr (std::is_same<T, uint8_t>::value) {
    // a cube of radius <= VOL_CUBE_RMAX: one launch (it reads planes z-r .. z+r while other
    // workgroups write plane z, so not in place)
    if (rz == ry && ry == rx && rz <= VOL_CUBE_RMAX && vol_u8_applicable(in, out, d.D, d.H, d.W) && in != out)
      return cube_u8<DIL>(in, out, d.D, d.H, d.W, (int)rz, s);
  }
  const int64_t n = d.D * d.H * d.W;
  if (n == 0) return PM_OK;
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * sizeof(T), s));
  T *t = reinterpret_cast<T *>(tmp.ptr);
  PM_TRY((rect_pass<T, DIL>(in, t, d.D, d.H, d.W, ry, rx, s)));
  return zline_pass<T, DIL>(t, out, d, rz, s);
}

// ============================================================ SE tables
int se_rank(int se)
{
  switch (se) {
  case PM_SE_SEGMENT_X_1D: return 1;
  case PM_SE_SQUARE_2D:
  case PM_SE_CROSS_2D:
  case PM_SE_SEGMENT_X_2D:
  case PM_SE_SEGMENT_Y_2D: return 2;
  case PM_SE_SEGMENT_X_3D:
  case PM_SE_SEGMENT_Y_3D:
  case PM_SE_SEGMENT_Z_3D:
  case PM_SE_CROSS_3D:
  case PM_SE_SQUARE_3D: return 3;
  default: return 0;
  }
}

// Neighbour list of a Common_NL_SE as {dz,dy,dx} triples (neighbor_list_static_se_t.hpp:37-432)
int se_nl(int se, NLArgs &a)
{
  a.n = 0;
  auto add = [&](int dz, int dy, int dx) { a.dz[a.n] = dz; a.dy[a.n] = dy; a.dx[a.n] = dx; a.n++; };
  switch (se) {
  case PM_SE_SEGMENT_X_1D:
  case PM_SE_SEGMENT_X_2D:
  case PM_SE_SEGMENT_X_3D: add(0, 0, 0); add(0, 0, -1); add(0, 0, 1); break;
  case PM_SE_SEGMENT_Y_2D:
  case PM_SE_SEGMENT_Y_3D: add(0, 0, 0); add(0, -1, 0); add(0, 1, 0); break;
  case PM_SE_SEGMENT_Z_3D: add(0, 0, 0); add(-1, 0, 0); add(1, 0, 0); break;
  case PM_SE_SQUARE_2D:
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) add(0, dy, dx);
    break;
  case PM_SE_CROSS_2D: add(0, 0, 0); add(0, -1, 0); add(0, 0, -1); add(0, 0, 1); add(0, 1, 0); break;
  case PM_SE_CROSS_3D:
    add(0, 0, 0); add(-1, 0, 0); add(0, -1, 0); add(0, 0, -1); add(0, 0, 1); add(0, 1, 0); add(1, 0, 0);
    break;
  case PM_SE_SQUARE_3D:
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) add(dz, dy, dx);
    break;
  default: PM_FAIL(PM_E_SE, "Unsupported Common_NL_SE %d", se);
  }
  return PM_OK;
}

// k successive passes of one neighbour list, ping-ponging so the last one writes out
template <typename T, bool DIL>
static int nl_iter(const T *in, T *out, Dims d, const NLArgs &a, int k, hipStream_t s)
{
  if (k == 1) return nl_pass<T, DIL>(in, out, d, a, s);
  const int64_t n = d.D * d.H * d.W;
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * sizeof(T), s));
  T *bufs[2] = {out, reinterpret_cast<T *>(tmp.ptr)};
  int dst = (k % 2 == 1) ? 0 : 1;
  const T *src = in;
  for (int i = 0; i < k; i++) {
    PM_TRY((nl_pass<T, DIL>(src, bufs[dst], d, a, s)));
    src = bufs[dst];
    dst ^= 1;
  }
  return PM_OK;
}

// SECross3D^k on a uint8 volume: k passes of the z-streaming cross, ping-ponging so that
// the last one writes out
template <bool DIL> static int cross3_iter(const uint8_t *in, uint8_t *out, Dims d, int k, hipStream_t s)
{
  if (k == 1) return cross3_u8<DIL>(in, out, d.D, d.H, d.W, s);
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)(d.D * d.H * d.W), s));
  uint8_t *bufs[2] = {out, (uint8_t *)tmp.ptr};
  int dst = (k % 2 == 1) ? 0 : 1;
  const uint8_t *src = in;
  for (int i = 0; i < k; i++) {
    PM_TRY(cross3_u8<DIL>(src, bufs[dst], d.D, d.H, d.W, s));
    src = bufs[dst];
    dst ^= 1;
  }
  return PM_OK;
}

// One Common_NL_SE iterated `iter` times (ero_dil_static.cpp:229-305); iter >= 1.
template <typename T, bool DIL>
int static_iter(const T *in, T *out, Dims d, int rank, int se, int iter, hipStream_t s)
{
  const int64_t k = iter;
  switch (se) {
  case PM_SE_SEGMENT_X_1D: return line_pass<T, DIL>(in, out, 1, 1, d.W, 0, k, s);
  case PM_SE_SQUARE_2D: return rect_pass<T, DIL>(in, out, 1, d.H, d.W, k, k, s);
  case PM_SE_SEGMENT_X_2D: return line_pass<T, DIL>(in, out, 1, d.H, d.W, 0, k, s);
  case PM_SE_SEGMENT_Y_2D: return line_pass<T, DIL>(in, out, 1, d.H, d.W, 1, k, s);
  case PM_SE_SEGMENT_X_3D: return line_pass<T, DIL>(in, out, 1, d.D * d.H, d.W, 0, k, s);
  case PM_SE_SEGMENT_Y_3D: return line_pass<T, DIL>(in, out, d.D, d.H, d.W, 1, k, s);
  case PM_SE_SEGMENT_Z_3D: return zline_pass<T, DIL>(in, out, d, k, s);
  case PM_SE_SQUARE_3D: return box_pass<T, DIL>(in, out, d, k, k, k, s);
  case PM_SE_CROSS_2D:
  case PM_SE_CROSS_3D: {
    if constexpr (std::is_same<T, uint8_t>::value) {
      if (se == PM_SE_CROSS_3D && vol_u8_applicable(in, out, d.D, d.H, d.W)) return cross3_iter<DIL>(in, out, d, iter, s);
    }
    NLArgs a;
    PM_TRY(se_nl(se, a));
    return nl_iter<T, DIL>(in, out, d, a, iter, s);
  }
  default: PM_FAIL(PM_E_SE, "Unsupported Common_NL_SE %d", se);
  }
  (void)rank;
}

// explicit instantiations used by other translation units
#define PM_INST(T)                                                                                           \
  template int static_iter<T, true>(const T *, T *, Dims, int, int, int, hipStream_t);                       \
  template int static_iter<T, false>(const T *, T *, Dims, int, int, int, hipStream_t);                      \
  template int rect_pass<T, true>(const T *, T *, int64_t, int64_t, int64_t, int64_t, int64_t, hipStream_t); \
  template int rect_pass<T, false>(const T *, T *, int64_t, int64_t, int64_t, int64_t, int64_t, hipStream_t); \
  template int nl_pass<T, true>(const T *, T *, Dims, const NLArgs &, hipStream_t);                          \
  template int nl_pass<T, false>(const T *, T *, Dims, const NLArgs &, hipStream_t);
PM_INST(uint8_t)
PM_INST(uint16_t)
PM_INST(int32_t)
PM_INST(int64_t)
PM_INST(float)
PM_INST(double)
#undef PM_INST

// ============================================================ entry-point bodies
int check_se_rank(int se, int rank)
{
  const int sr = se_rank(se);
  if (sr == 0) PM_FAIL(PM_E_SE, "Unsupported Common_NL_SE %d", se);
  if (sr != rank) PM_FAIL(PM_E_RANK, "SE and image have not the same rank (%d vs %d)", sr, rank);
  return PM_OK;
}

template <bool DIL> static int erodil_static(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{
  const char *what = DIL ? "Dilate iter" : "Erode iter";
  PM_TRY(check_unary(in, out, what));
  if (iter < 0) PM_FAIL(PM_E_ARG, "%s must be >= 0", what);
  hipStream_t s = as_stream(stream);
  const int64_t n = numel(*in);
  if (iter == 0 || n == 0) {
    if (n) PM_HIP(hipMemcpyAsync(out->data, in->data, (size_t)n * ptype_size(in->ptype), hipMemcpyDeviceToDevice, s));
    return PM_OK;
  }
  PM_TRY(check_se_rank(se, in->rank));
  const Dims d = dims3(*in);
  auto f = [in = in, out = out, d = d, rank = in->rank, se = se, iter = iter, s = s](auto tag) -> int {
      using T = typename decltype(tag)::type;
      return static_iter<T, DIL>((const T *)in->data, (T *)out->data, d, rank, se, iter, s);
    };
  return dispatch_ptype(in->ptype, f);
}

template <bool DIL> static int erodil_line(const pm_image *in, int64_t half, pm_image *out, void *stream, int axis)
{
  const char *what = axis == 0 ? (DIL ? "DilateX" : "ErodeX") : (DIL ? "DilateY" : "ErodeY");
  PM_TRY(check_unary(in, out, what));
  if (axis == 0 && in->rank > 2) PM_FAIL(PM_E_RANK, "%s: 1D and 2D only", what);
  if (axis == 1 && in->rank != 2) PM_FAIL(PM_E_RANK, "%s: 2D only", what);
  const Dims d = dims3(*in);
  auto f = [in = in, out = out, d = d, r = half, axis = axis, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type;
      return line_pass<T, DIL>((const T *)in->data, (T *)out->data, 1, d.H, d.W, axis, r, s);
    };
  return dispatch_ptype(in->ptype, f);
}

template <bool DIL> static int erodil_rect(const pm_image *in, int64_t ry, int64_t rx, pm_image *out, void *stream)
{
  PM_TRY(check_unary(in, out, DIL ? "Dilate rect" : "Erode rect"));
  if (in->rank < 2) PM_FAIL(PM_E_RANK, "rectangle SE needs a 2D or 3D image");
  const Dims d = dims3(*in);
  auto f = [in = in, out = out, d = d, ry = ry, rx = rx, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type;
      return rect_pass<T, DIL>((const T *)in->data, (T *)out->data, d.D, d.H, d.W, ry, rx, s);
    };
  return dispatch_ptype(in->ptype, f);
}

template <bool DIL> static int erodil_nl(const pm_image *in, const int64_t *offs, int n, pm_image *out, void *stream)
{
  PM_TRY(check_unary(in, out, DIL ? "Dilate" : "Erode"));
  if (n <= 0 || n > NL_MAX || !offs) PM_FAIL(PM_E_ARG, "neighbour list must hold 1..%d offsets", NL_MAX);
  NLArgs a;
  a.n = n;
  const int R = in->rank;
  for (int k = 0; k < n; k++) {
    int64_t o[3] = {0, 0, 0};
    for (int j = 0; j < R; j++) o[3 - R + j] = offs[k * R + j];
    for (int j = 0; j < 3; j++)
      if (o[j] < -(1 << 20) || o[j] > (1 << 20)) PM_FAIL(PM_E_ARG, "neighbour offset out of range");
    a.dz[k] = (int)o[0]; a.dy[k] = (int)o[1]; a.dx[k] = (int)o[2];
  }
  const Dims d = dims3(*in);
  auto f = [in = in, out = out, d = d, a = &a, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type; return nl_pass<T, DIL>((const T *)in->data, (T *)out->data, d, *a, s); };
  return dispatch_ptype(in->ptype, f);
}

// Octagon / rhombicuboctahedron (ero_dil_compound_static_se_t.hpp:33-99): the cross applied
// (size - m) times then the square m times, m = round-half-up(size / (1 + sqrt 2)).
template <bool DIL> static int erodil_compound(const pm_image *in, int compound, int size, pm_image *out, void *stream)
{
  const char *what = DIL ? "Dilate compound" : "Erode compound";
  PM_TRY(check_unary(in, out, what));
  if (size < 0) PM_FAIL(PM_E_ARG, "%s: size must be >= 0", what);
  int cross, square;
  if (compound == PM_COMPOUND_OCTAGON) {
    if (in->rank != 2) PM_FAIL(PM_E_RANK, "%s: octagon needs a 2D image", what);
    cross = PM_SE_CROSS_2D; square = PM_SE_SQUARE_2D;
  } else if (compound == PM_COMPOUND_RHOMBICUBOCTAHEDRON) {
    if (in->rank != 3) PM_FAIL(PM_E_RANK, "%s: rhombicuboctahedron needs a 3D image", what);
    // the reference applies SECross2D to the 3D image for size == 1 and throws (:77-79)
    if (size == 1) PM_FAIL(PM_E_UNSUPPORTED, "%s: SE and view have not the same Rank", what);
    cross = PM_SE_CROSS_3D; square = PM_SE_SQUARE_3D;
  } else {
    PM_FAIL(PM_E_SE, "Unsupported Compound_NL_SE %d", compound);
  }
  hipStream_t s = as_stream(stream);
  const int64_t n = numel(*in);
  if (size == 0 || n == 0) {
    if (n) PM_HIP(hipMemcpyAsync(out->data, in->data, (size_t)n * ptype_size(in->ptype), hipMemcpyDeviceToDevice, s));
    return PM_OK;
  }
  const double q = (double)size / (1.0 + 1.4142135623730951);
  const double fl = std::floor(q);
  const int m = (int)((q - fl) < 0.5 ? fl : fl + 1.0);
  const int ncross = size - m;
  const Dims d = dims3(*in);
  auto f = [in = in, out = out, d = d, rank = in->rank, cross = cross, square = square, ncross = ncross, m = m, s = s](auto tag) -> int {
      using T = typename decltype(tag)::type;
      const T *src = (const T *)in->data;
      T *dst = (T *)out->data;
      if (m == 0) return static_iter<T, DIL>(src, dst, d, rank, cross, ncross, s);
      Scratch tmp;
      PM_TRY(tmp.alloc((size_t)(d.D * d.H * d.W) * sizeof(T), s));
      T *t = (T *)tmp.ptr;
      PM_TRY((static_iter<T, DIL>(src, t, d, rank, cross, ncross, s)));
      return static_iter<T, DIL>(t, dst, d, rank, square, m, s);
    };
  return dispatch_ptype(in->ptype, f);
}

}  // namespace pm

using namespace pm;

extern "C" {
int pm_erode_static(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_static<false>(in, se, iter, out, stream); }
int pm_dilate_static(const pm_image *in, int se, int iter, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_static<true>(in, se, iter, out, stream); }
int pm_erode_compound(const pm_image *in, int compound, int size, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_compound<false>(in, compound, size, out, stream); }
int pm_dilate_compound(const pm_image *in, int compound, int size, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_compound<true>(in, compound, size, out, stream); }
int pm_erode_nl(const pm_image *in, const int64_t *offsets, int n, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_nl<false>(in, offsets, n, out, stream); }
int pm_dilate_nl(const pm_image *in, const int64_t *offsets, int n, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_nl<true>(in, offsets, n, out, stream); }
int pm_erode_x(const pm_image *in, int64_t half, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_line<false>(in, half, out, stream, 0); }
int pm_dilate_x(const pm_image *in, int64_t half, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_line<true>(in, half, out, stream, 0); }
int pm_erode_y(const pm_image *in, int64_t half, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_line<false>(in, half, out, stream, 1); }
int pm_dilate_y(const pm_image *in, int64_t half, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_line<true>(in, half, out, stream, 1); }
int pm_erode_rect(const pm_image *in, int64_t ry, int64_t rx, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_rect<false>(in, ry, rx, out, stream); }
int pm_dilate_rect(const pm_image *in, int64_t ry, int64_t rx, pm_image *out, void *stream)
{ PM_TRACE(); return erodil_rect<true>(in, ry, rx, out, stream); }
}
