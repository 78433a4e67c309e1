// label.hip -- connected-component labelling: label_binary, label_flat_zones, label_minima,
// label_maxima (include/poutre/label/label_binary.hpp:35, label_flat_zones.hpp:35,
// label_extrema.hpp:34-39; algorithm of label_t.hpp t_label_helper).
//
// The reference makes one raster pass with a union-find over provisional labels and
// renumbers the roots; its result is fixed by the link graph alone: a pixel is linked to
// each raster-earlier ("upper") neighbour of the SE that satisfies the relation (binary:
// both pixels are foreground and the earlier one is > 0; flat zones: equal values), and each
// component is numbered 1.. by the raster position of its first pixel (the first pixel of a
// component always opens a new provisional label, which stays the component's root because
// the union keeps the smaller root).  On the GPU:
//  1. lab_tile2d / lab_tile: per 32 x 64 tile (8 x 8 x 32 in 3D), the links inside the tile
//                   are merged in LDS (in 2D the links along a row by a ballot over the
//                   row's runs) and every pixel gets the global index of its tile-local root;
//  2. lab_border:   the links that cross tile borders are merged in global memory.  Both use
//                   a lock-free union whose roots only decrease (atomicMin on the larger
//                   root, retried from the value it returns), so each root ends as the
//                   smallest index of its component;
//  3. lab_compress: P[i] = root;
//  4. ranks of the roots in raster order by a block count, a scan of the block counts and a
//     per-block scan (lab_count / lab_scan_blocks / lab_rank);
//  5. lab_output:   out[i] = rank of P[i] (int64, the reference's GINT64 output), 0 for the
//                   background.
// Positive float values at or below the reference's 1e-10 background threshold are not
// linked (the reference reads an unassigned provisional label for them).
#include <cmath>

#include "erodil.hpp"
#include "pm_common.hpp"

namespace pm {

constexpr int LAB_BLOCK = 1024;  // elements per block of the rank scan (256 threads x 4)

struct LabUpper {
  int n;
  int dz[13], dy[13], dx[13];
};

template <typename T> __device__ __forceinline__ bool lab_fg(T v) { return v != T(0); }
template <> __device__ __forceinline__ bool lab_fg<float>(float v) { return fabsf(v) > 1E-10F; }
template <> __device__ __forceinline__ bool lab_fg<double>(double v) { return fabs(v) > 1E-10; }

template <typename T, bool FLAT> __device__ __forceinline__ bool lab_accept(T v) { return FLAT || lab_fg<T>(v); }
// v at the pixel, w at its raster-earlier neighbour
template <typename T, bool FLAT> __device__ __forceinline__ bool lab_link(T v, T w)
{
  if (FLAT) return w == v;
  return lab_fg<T>(v) && lab_fg<T>(w) && w > T(0);
}

// Root of x in global memory, without path compression: while unions run, the parents in
// global memory change by atomics only (plain stores going through another XCD's cache
// could overwrite a link), and halving by atomics measured slower (4096^2 binary 0.65 ->
// 0.74 ms); lab_compress flattens the trees once the unions are done.
__device__ __forceinline__ int lab_find(const int *P, int x)
{
  int y;
  while ((y = P[x]) != x) x = y;
  return x;
}

// Merge the sets of a and b.  Invariant: P[x] <= x and values only decrease, so every find
// terminates and every retry starts from a strictly smaller index.
__device__ __forceinline__ void lab_unite(int *P, int a, int b)
{
  for (;;) {
    a = lab_find(P, a);
    b = lab_find(P, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&P[a], b);  // link the larger root under the smaller one
    if (old == a) return;
    a = old;  // a had been linked meanwhile: keep its new parent connected to b
  }
}

// ---- tiled form: union-find inside a TD x TH x TW tile in LDS, then only the links that
// cross tile borders go through global memory.
// Root of x in LDS with path halving (x's parent moved up to its grandparent as the walk
// goes).  A halving store may race with a union lowering L[x]; unions only link roots and,
// finding x already linked, go on with x's parent, so x keeps an ancestor (a smaller index)
// as parent either way and no link is lost.
__device__ __forceinline__ int lab_find_lds(int *L, int x)
{
  for (;;) {
    const int p = L[x];
    if (p == x) return x;
    const int gp = L[p];
    if (gp == p) return p;
    L[x] = gp;
    x = gp;
  }
}
__device__ __forceinline__ void lab_unite_lds(int *L, int a, int b)
{
  for (;;) {
    a = lab_find_lds(L, a);
    b = lab_find_lds(L, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&L[a], b);
    if (old == a) return;
    a = old;
  }
}

// Tile (blockIdx.x, .y, .z): local labels in LDS (local index = (lz*TH + ly)*TW + lx, which
// orders the tile's pixels as the raster does, so a local root is the component's first
// pixel in the tile), then P[i] = global index of the local root (-1 for background).
template <typename T, bool FLAT, int TD, int TH, int TW>
__global__ __launch_bounds__(256) void lab_tile(const T *__restrict__ in, int *__restrict__ P, int D, int H, int W,
                                                LabUpper u)
{
  constexpr int NT = TD * TH * TW;
  __shared__ int L[NT];
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH, z0 = blockIdx.z * TD;
  for (int e = threadIdx.x; e < NT; e += 256) {
    const int lx = e % TW, ly = (e / TW) % TH, lz = e / (TW * TH);
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    bool a = false;
    if (x < W && y < H && z < D) a = lab_accept<T, FLAT>(in[((int64_t)z * H + y) * W + x]);
    L[e] = a ? e : -1;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NT; e += 256) {
    if (L[e] < 0) continue;  // background or outside the image (a pixel's own entry is only
                             // lowered by unions, never to -1)
    const int lx = e % TW, ly = (e / TW) % TH, lz = e / (TW * TH);
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    const T v = in[((int64_t)z * H + y) * W + x];
    for (int k = 0; k < u.n; k++) {
      const int nx = lx + u.dx[k], ny = ly + u.dy[k], nz = lz + u.dz[k];
      if ((unsigned)nx >= (unsigned)TW || (unsigned)ny >= (unsigned)TH || (unsigned)nz >= (unsigned)TD) continue;
      if (x0 + nx >= W || y0 + ny >= H || z0 + nz >= D) continue;
      const T w = in[((int64_t)(z0 + nz) * H + (y0 + ny)) * W + (x0 + nx)];
      if (lab_link<T, FLAT>(v, w)) lab_unite_lds(L, e, (nz * TH + ny) * TW + nx);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NT; e += 256) {
    const int lx = e % TW, ly = (e / TW) % TH, lz = e / (TW * TH);
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    if (x >= W || y >= H || z >= D) continue;
    int g = -1;
    if (L[e] >= 0) {
      const int r = lab_find_lds(L, e);
      const int rx = r % TW, ry = (r / TW) % TH, rz = r / (TW * TH);
      g = (int)(((int64_t)(z0 + rz) * H + (y0 + ry)) * W + (x0 + rx));
    }
    P[((int64_t)z * H + y) * W + x] = g;
  }
}

// 2D tiles of TH rows x 64 columns, one wave per row at a time (lane = column).  The link to
// the left pixel of the same row is resolved without unions: a ballot marks the lanes that
// start a run (not linked to their left neighbour) and every pixel points at the start of
// its run.  Only the links to the row above (dy = -1) then go through the LDS union-find.
template <typename T, bool FLAT, int TH>
__global__ __launch_bounds__(256) void lab_tile2d(const T *__restrict__ in, int *__restrict__ P, int H, int W,
                                                  LabUpper u, int xlink)
{
  constexpr int TW = 64, NT = TH * TW, RPW = TH / 4;  // rows per wave
  __shared__ int L[NT];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH, x = x0 + lane;
  T v[RPW];
  bool acc[RPW];
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k, y = y0 + ly;
    const bool in_img = x < W && y < H;
    v[k] = in_img ? in[(int64_t)y * W + x] : T(0);
    acc[k] = in_img && lab_accept<T, FLAT>(v[k]);
    bool lk = false;
    if (xlink && acc[k] && lane > 0) lk = lab_link<T, FLAT>(v[k], in[(int64_t)y * W + x - 1]);
    const uint64_t starts = __ballot(!lk);
    const uint64_t m = lane == 63 ? starts : (starts & ((2ull << lane) - 1));
    const int st = 63 - __clzll((long long)m);
    L[ly * TW + lane] = acc[k] ? ly * TW + st : -1;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k, y = y0 + ly;
    if (!acc[k] || ly == 0) continue;
    for (int j = 0; j < u.n; j++) {
      if (u.dz[j] != 0 || u.dy[j] != -1) continue;  // (dz != 0 only for 3D SEs, out of a D = 1 image)
      const int nx = lane + u.dx[j];
      if ((unsigned)nx >= (unsigned)TW || x0 + nx >= W) continue;
      if (lab_link<T, FLAT>(v[k], in[(int64_t)(y - 1) * W + x0 + nx]))
        lab_unite_lds(L, ly * TW + lane, (ly - 1) * TW + nx);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k, y = y0 + ly;
    if (x >= W || y >= H) continue;
    int g = -1;
    if (acc[k]) {
      const int r = lab_find_lds(L, ly * TW + lane);
      g = (int)((int64_t)(y0 + r / TW) * W + x0 + (r % TW));
    }
    P[(int64_t)y * W + x] = g;
  }
}

// The links whose two pixels lie in different tiles, merged in global memory.  After lab_tile
// a pixel's parent is its tile-local root, and it stays so (unions only relink roots), so a
// link can be told by the pair (P[i], P[j]) of tile roots: when the previous lane merged the
// same pair -- consecutive pixels along a tile's top edge mostly do -- the union is skipped.
template <typename T, bool FLAT, int TD, int TH, int TW>
__global__ __launch_bounds__(256) void lab_border(const T *__restrict__ in, int *P, int D, int H, int W, LabUpper u)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t N = (int64_t)D * H * W;
  int x = 0, y = 0, z = 0;
  bool cand = false;
  if (i < N) {
    x = (int)(i % W);
    const int64_t r = i / W;
    y = (int)(r % H);
    z = (int)(r / H);
    const int lx = x % TW, ly = y % TH, lz = z % TD;
    // only pixels on a tile face can have an upper neighbour in another tile (upper offsets
    // have dz <= 0; with dz = -1 they may step +1 in y or x, with dy = -1 +1 in x)
    cand = lx == 0 || lx == TW - 1 || ly == 0 || ly == TH - 1 || (TD > 1 && lz == 0);
  }
  if (!__any(cand)) return;  // the whole wave: no shuffle below loses a participant
  const int ra = cand ? P[i] : -1;
  cand = cand && ra >= 0;
  const T v = cand ? in[i] : T(0);
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < u.n; k++) {
    bool lk = false;
    int rb = -1;
    int64_t j = 0;
    if (cand) {
      const int zz = z + u.dz[k], yy = y + u.dy[k], xx = x + u.dx[k];
      const bool other_tile = (zz / TD != z / TD) || (yy / TH != y / TH) || (xx / TW != x / TW) || zz < 0 || yy < 0 || xx < 0;
      if (other_tile && (unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W) {
        j = ((int64_t)zz * H + yy) * W + xx;
        lk = lab_link<T, FLAT>(v, in[j]);
        if (lk) rb = P[j];
      }
    }
    const int pra = __shfl_up(lk ? ra : -2, 1), prb = __shfl_up(rb, 1);
    const bool dup = lane != 0 && pra == ra && prb == rb;
    if (lk && !dup) lab_unite(P, (int)i, (int)j);
  }
}

__global__ void lab_compress(int *P, int64_t N)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N && P[i] >= 0) P[i] = lab_find(P, (int)i);
}

// roots (P[i] == i) per block of LAB_BLOCK elements
__global__ __launch_bounds__(256) void lab_count(const int *__restrict__ P, int *__restrict__ cnt, int64_t N)
{
  __shared__ int s[256];
  const int64_t base = (int64_t)blockIdx.x * LAB_BLOCK + 4 * threadIdx.x;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) c += (base + k < N && P[base + k] == (int)(base + k)) ? 1 : 0;
  s[threadIdx.x] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) cnt[blockIdx.x] = s[0];
}

// inclusive scan of 256 values in LDS (Hillis-Steele)
__device__ __forceinline__ int lab_scan256(int *s, int v)
{
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int a = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  return s[t];
}

// exclusive offsets of the block counts (one workgroup, chunks of 1024 with a running
// carry) and the total number of labels
__global__ __launch_bounds__(256) void lab_scan_blocks(int *cnt, int nb, int64_t *total)
{
  __shared__ int s[256];
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    int v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int b = c0 + 4 * threadIdx.x + k;
      v[k] = b < nb ? cnt[b] : 0;
      sum += v[k];
    }
    const int incl = lab_scan256(s, sum);
    int run = carry + incl - sum;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int b = c0 + 4 * threadIdx.x + k;
      if (b < nb) cnt[b] = run;
      run += v[k];
    }
    carry += s[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// rank (1-based, raster order) of every root into R
__global__ __launch_bounds__(256) void lab_rank(const int *__restrict__ P, const int *__restrict__ off,
                                                int *__restrict__ R, int64_t N)
{
  __shared__ int s[256];
  const int64_t base = (int64_t)blockIdx.x * LAB_BLOCK + 4 * threadIdx.x;
  bool f[4];
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[k] = base + k < N && P[base + k] == (int)(base + k);
    c += f[k] ? 1 : 0;
  }
  int run = off[blockIdx.x] + lab_scan256(s, c) - c;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (f[k]) R[base + k] = ++run;
}

__global__ void lab_output(const int *__restrict__ P, const int *__restrict__ R, int64_t *__restrict__ out, int64_t N)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N) out[i] = P[i] >= 0 ? (int64_t)R[P[i]] : 0;
}

// upper (raster-earlier) neighbours of a Common_NL_SE (neighbor_list_static_se_t.hpp
// split_coordinates_upper_lower)
static int lab_upper(int se, LabUpper &u)
{
  NLArgs a;
  PM_TRY(se_nl(se, a));
  u.n = 0;
  for (int k = 0; k < a.n; k++) {
    const int sign = a.dz[k] ? a.dz[k] : a.dy[k] ? a.dy[k] : a.dx[k];
    if (sign < 0) {
      u.dz[u.n] = a.dz[k];
      u.dy[u.n] = a.dy[k];
      u.dx[u.n] = a.dx[k];
      u.n++;
    }
  }
  return PM_OK;
}

static int lab_check(const pm_image *in, const pm_image *out, int se, const char *what)
{
  clear_error();
  if (!in || !out) PM_FAIL(PM_E_ARG, "%s: null image descriptor", what);
  if (in->rank == 0) PM_FAIL(PM_E_RANK, "%s Unsupported number of dims:0", what);
  if (in->rank < 1 || in->rank > 3) PM_FAIL(PM_E_RANK, "%s Unsupported number of dims", what);
  if (out->rank != in->rank) PM_FAIL(PM_E_SIZE, "%s incompatible size", what);
  for (int i = 0; i < in->rank; i++) {
    if (in->shape[i] < 0) PM_FAIL(PM_E_SIZE, "%s: negative extent", what);
    if (out->shape[i] != in->shape[i]) PM_FAIL(PM_E_SIZE, "%s incompatible size", what);
  }
  if (ptype_size(in->ptype) == 0) PM_FAIL(PM_E_PTYPE, "%s unsupported PTYPE", what);
  if (out->ptype != PM_PTYPE_I64) PM_FAIL(PM_E_PTYPE, "%s: the output image must be GINT64", what);
  const int64_t n = numel(*in);
  if (n > 0 && (!in->data || !out->data)) PM_FAIL(PM_E_ARG, "%s: null data pointer", what);
  if (n > 0 && in->data == out->data) PM_FAIL(PM_E_ALIAS, "%s output must be != than input images", what);
  if (n >= INT32_MAX) PM_FAIL(PM_E_SIZE, "%s: images of 2^31 pixels or more are not supported", what);
  PM_TRY(check_se_rank(se, in->rank));
  return PM_OK;
}

template <bool FLAT>
static int lab_run(const pm_image *in, int se, pm_image *out, int64_t *count, hipStream_t s, const char *what)
{
  PM_TRY(lab_check(in, out, se, what));
  const int64_t N = numel(*in);
  if (N == 0) {
    if (count) *count = 0;
    return PM_OK;
  }
  LabUpper u;
  PM_TRY(lab_upper(se, u));
  const Dims d = dims3(*in);
  const int64_t nb = (N + LAB_BLOCK - 1) / LAB_BLOCK;
  Scratch sc;
  const size_t pb = ((size_t)N * sizeof(int) + 255) & ~(size_t)255, cb = ((size_t)nb * sizeof(int) + 255) & ~(size_t)255;
  PM_TRY(sc.alloc(2 * pb + cb + 256, s));
  int *P = (int *)sc.ptr;
  int *R = (int *)((char *)sc.ptr + pb);
  int *cnt = (int *)((char *)sc.ptr + 2 * pb);
  int64_t *total = (int64_t *)((char *)sc.ptr + 2 * pb + cb);
  const unsigned g = (unsigned)((N + 255) / 256);
  auto f = [&](auto tag) -> int {
    using T = typename decltype(tag)::type;
    const T *src = (const T *)in->data;
    if (d.D == 1) {  // 1D / 2D: tiles of 32 x 64
      constexpr int TH = 32, TW = 64;
      const dim3 tg((unsigned)((d.W + TW - 1) / TW), (unsigned)((d.H + TH - 1) / TH), 1);
      int xlink = 0;  // is the left neighbour (dy = 0, dx = -1) an upper offset of the SE
      for (int j = 0; j < u.n; j++) xlink |= (u.dz[j] == 0 && u.dy[j] == 0 && u.dx[j] == -1);
      hipLaunchKernelGGL((lab_tile2d<T, FLAT, TH>), tg, dim3(256), 0, s, src, P, (int)d.H, (int)d.W, u, xlink);
      hipLaunchKernelGGL((lab_border<T, FLAT, 1, TH, TW>), dim3(g), dim3(256), 0, s, src, P, 1, (int)d.H, (int)d.W, u);
    } else {  // 3D: tiles of 8 x 8 x 32
      constexpr int TD = 8, TH = 8, TW = 32;
      const dim3 tg((unsigned)((d.W + TW - 1) / TW), (unsigned)((d.H + TH - 1) / TH), (unsigned)((d.D + TD - 1) / TD));
      hipLaunchKernelGGL((lab_tile<T, FLAT, TD, TH, TW>), tg, dim3(256), 0, s, src, P, (int)d.D, (int)d.H, (int)d.W, u);
      hipLaunchKernelGGL((lab_border<T, FLAT, TD, TH, TW>), dim3(g), dim3(256), 0, s, src, P, (int)d.D, (int)d.H,
                         (int)d.W, u);
    }
    PM_HIP(hipGetLastError());
    return PM_OK;
  };
  PM_TRY(dispatch_ptype(in->ptype, f));
  hipLaunchKernelGGL(lab_compress, dim3(g), dim3(256), 0, s, P, N);
  hipLaunchKernelGGL(lab_count, dim3((unsigned)nb), dim3(256), 0, s, P, cnt, N);
  hipLaunchKernelGGL(lab_scan_blocks, dim3(1), dim3(256), 0, s, cnt, (int)nb, total);
  hipLaunchKernelGGL(lab_rank, dim3((unsigned)nb), dim3(256), 0, s, P, cnt, R, N);
  hipLaunchKernelGGL(lab_output, dim3(g), dim3(256), 0, s, P, R, (int64_t *)out->data, N);
  PM_HIP(hipGetLastError());
  if (count) {
    PM_HIP(hipMemcpyAsync(count, total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
  }
  return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" {
int pm_label_binary(const pm_image *in, int se, pm_image *out, int64_t *count, void *stream)
{
  return lab_run<false>(in, se, out, count, as_stream(stream), "label_binary");
}

int pm_label_flat_zones(const pm_image *in, int se, pm_image *out, int64_t *count, void *stream)
{
  return lab_run<true>(in, se, out, count, as_stream(stream), "label_flat_zones");
}

int pm_h_convex(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_h_concave(const pm_image *in, double h, int se, pm_image *out, void *stream);

// label_extrema.cpp: label_binary of h_concave(img, 1) (minima) / h_convex(img, 1) (maxima)
static int lab_extrema(bool minima, const pm_image *in, int se, pm_image *out, void *stream)
{
  const char *what = minima ? "label_minima" : "label_maxima";
  PM_TRY(lab_check(in, out, se, what));
  hipStream_t s = as_stream(stream);
  const int64_t n = numel(*in);
  if (n == 0) return PM_OK;
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * ptype_size(in->ptype), s));
  pm_image t = *in;
  t.data = tmp.ptr;
  PM_TRY(minima ? pm_h_concave(in, 1.0, se, &t, stream) : pm_h_convex(in, 1.0, se, &t, stream));
  return lab_run<false>(&t, se, out, nullptr, s, what);
}

int pm_label_minima(const pm_image *in, int se, pm_image *out, void *stream) { return lab_extrema(true, in, se, out, stream); }
int pm_label_maxima(const pm_image *in, int se, pm_image *out, void *stream) { return lab_extrema(false, in, se, out, stream); }
}
