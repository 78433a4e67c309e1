// sq_u8.hip -- fused flat dilation / erosion of uint8 images by a (2r+1)^2 square, r <= 32.
//
// This is the kernel behind Dilate/Erode(SESquare2D, iter = r) and DilateX+DilateY for
// uint8 (ero_dil_static.cpp:229-305, ero_dil_line_se_t.hpp:49-280): one launch reads
// every input pixel once from HBM and writes every output once.
//
// One workgroup (256 threads) owns a 128x128 output tile.
//  1. The input region (128+2r rows, 128+2*R16 columns, R16 = r rounded up to 16 so that
//     global loads are 16-byte aligned) is copied to LDS with 128-bit loads; pixels
//     outside the image get the neutral value (0 for dilation, 255 for erosion).
//  2. Vertical pass: a thread owns a 4-column quad and a run of K output rows.  It reads
//     K+2r words down the column, unpacks every word into two u16x2 (columns 0/2 and 1/3)
//     and computes the windowed max with a single-split van Herk: suffix maxima over the
//     first K rows, running prefix over the rest, one v_pk_max_u16 per two pixels.
//     Results are written TRANSPOSED into a second LDS buffer (region column -> row).
//  3. Horizontal pass = the same vertical code over the transposed buffer; the 4x4 byte
//     transpose back to row order is done in registers (v_perm_b32) and the tile is
//     stored with 128-bit (or 32-bit for K < 16) global stores.
// K (run length) is the largest power of two <= min(2r, 32): the suffix covers rows
// [j, K) and the prefix rows [K, j+2r] of output j, which needs K <= 2r + 1.  For r = 1
// a direct 3-tap form with K = 4 is used.
// Tiles are mapped XCD-aware: workgroup b runs on XCD b % 8, so consecutive tiles of a
// band go to the same XCD and share their halo rows in that XCD's L2.
#include "pm_common.hpp"
#include "sq_u8.hpp"

namespace pm {

typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

constexpr int FT = 128;       // output tile side
constexpr int FTHREADS = 256;
constexpr int TPAD = 4;       // transposed buffer row padding (bytes)

template <bool DIL> __device__ __forceinline__ uint32_t pk(uint32_t a, uint32_t b)
{
  const u16x2_t x = __builtin_bit_cast(u16x2_t, a), y = __builtin_bit_cast(u16x2_t, b);
  return __builtin_bit_cast(uint32_t, DIL ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
}
// word (c0,c1,c2,c3) -> (c0,c2) and (c1,c3) as u16x2
__device__ __forceinline__ uint32_t even_pairs(uint32_t w) { return __builtin_amdgcn_perm(0u, w, 0x0c020c00u); }
__device__ __forceinline__ uint32_t odd_pairs(uint32_t w) { return __builtin_amdgcn_perm(0u, w, 0x0c030c01u); }

// 4x4 byte transpose: e[j] = (c0_j, c2_j), o[j] = (c1_j, c3_j) for j = 0..3 ->
// w[c] = (c_0, c_1, c_2, c_3) bytes
__device__ __forceinline__ void transpose4(const uint32_t *e, const uint32_t *o, uint32_t *w)
{
  const uint32_t ea = __builtin_amdgcn_perm(e[1], e[0], 0x06020400u);
  const uint32_t eb = __builtin_amdgcn_perm(e[3], e[2], 0x06020400u);
  const uint32_t oa = __builtin_amdgcn_perm(o[1], o[0], 0x06020400u);
  const uint32_t ob = __builtin_amdgcn_perm(o[3], o[2], 0x06020400u);
  w[0] = __builtin_amdgcn_perm(eb, ea, 0x05040100u);
  w[2] = __builtin_amdgcn_perm(eb, ea, 0x07060302u);
  w[1] = __builtin_amdgcn_perm(ob, oa, 0x05040100u);
  w[3] = __builtin_amdgcn_perm(ob, oa, 0x07060302u);
}

__device__ __forceinline__ uint32_t lds_word(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }

// Windowed max/min of 2r+1 rows for K consecutive output rows of one 4-byte column quad.
// base points at the quad in the first input row (output row j covers input rows j..j+2r);
// stride is the row pitch in bytes.
template <int K, bool DIL>
__device__ __forceinline__ void vh_run(const uint8_t *base, int stride, int r, uint32_t (&e)[K], uint32_t (&o)[K])
{
  const uint32_t NEU = DIL ? 0u : 0x00FF00FFu;
  if (K > 4 || 2 * r >= K) {
#pragma unroll
    for (int j = 0; j < K; j++) {
      const uint32_t w = lds_word(base + j * stride);
      e[j] = even_pairs(w);
      o[j] = odd_pairs(w);
    }
#pragma unroll
    for (int j = K - 2; j >= 0; j--) {
      e[j] = pk<DIL>(e[j], e[j + 1]);
      o[j] = pk<DIL>(o[j], o[j + 1]);
    }
    uint32_t pe = NEU, po = NEU;
    const uint8_t *p = base + K * stride;
    for (int t = K; t < 2 * r; t++, p += stride) {
      const uint32_t w = lds_word(p);
      pe = pk<DIL>(pe, even_pairs(w));
      po = pk<DIL>(po, odd_pairs(w));
    }
    p = base + 2 * r * stride;
#pragma unroll
    for (int j = 0; j < K; j++) {
      const uint32_t w = lds_word(p + j * stride);
      pe = pk<DIL>(pe, even_pairs(w));
      po = pk<DIL>(po, odd_pairs(w));
      e[j] = pk<DIL>(e[j], pe);
      o[j] = pk<DIL>(o[j], po);
    }
  } else if constexpr (K == 4) {
    // direct taps for 2r < K (only r = 1 with K = 4 reaches here)
    uint32_t xe[K + 3], xo[K + 3];
#pragma unroll
    for (int t = 0; t < K + 3; t++) {
      if (t < K + 2 * r) {
        const uint32_t w = lds_word(base + t * stride);
        xe[t] = even_pairs(w);
        xo[t] = odd_pairs(w);
      } else {
        xe[t] = NEU;
        xo[t] = NEU;
      }
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
      uint32_t ae = xe[j], ao = xo[j];
#pragma unroll
      for (int d = 1; d <= 3; d++)
        if (d <= 2 * r) {
          ae = pk<DIL>(ae, xe[j + d]);
          ao = pk<DIL>(ao, xo[j + d]);
        }
      e[j] = ae;
      o[j] = ao;
    }
  }
}

__device__ __forceinline__ int xcd_tile(int b, int ntiles)
{
  // workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous
  // range of tiles so that neighbouring tiles share halo rows in one L2
  if ((ntiles & 7) != 0) return b;
  return (b & 7) * (ntiles >> 3) + (b >> 3);
}

template <int K, bool DIL>
__global__ __launch_bounds__(FTHREADS) void sq_u8_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                         int H, int W, int r, int R16, int tiles_x, int tiles_per_img,
                                                         int ntiles)
{
  extern __shared__ __align__(16) uint8_t smem[];
  const int AW = FT + 2 * R16;  // region width in bytes (multiple of 16)
  const int AH = FT + 2 * r;    // region height
  const int TS = FT + TPAD;     // transposed row pitch
  uint8_t *A = smem;
  uint8_t *T = smem + AH * AW;
  const int tid = threadIdx.x;

  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int img = tile / tiles_per_img;
  const int t_in = tile - img * tiles_per_img;
  const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
  const int y0 = ty * FT, x0 = tx * FT;
  const int64_t ibase = (int64_t)img * H * W;

  // ---- 1. region -> LDS (W is a multiple of 16, so a 16-byte chunk is all in or all out)
  {
    const uint32_t nv = DIL ? 0u : 0xFFFFFFFFu;
    const uint4 neutral = make_uint4(nv, nv, nv, nv);
    const int cpr = AW >> 4;
    const int n = AH * cpr;
    for (int c = tid; c < n; c += FTHREADS) {
      const int i = c / cpr, cc = c - i * cpr;
      const int gy = y0 - r + i, gx = x0 - R16 + (cc << 4);
      uint4 v = neutral;
      if ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W)
        v = *reinterpret_cast<const uint4 *>(in + ibase + (int64_t)gy * W + gx);
      *reinterpret_cast<uint4 *>(A + i * AW + (cc << 4)) = v;
    }
  }
  __syncthreads();

  // ---- 2. vertical pass, written transposed into T (T row = region column)
  {
    const int q0 = (R16 - r) >> 2, q1 = (R16 + FT + r + 3) >> 2;
    const int nq = q1 - q0;
    const int ntask = nq * (FT / K);
    for (int task = tid; task < ntask; task += FTHREADS) {
      const int k = task / nq, q = q0 + (task - k * nq);
      uint32_t e[K], o[K];
      vh_run<K, DIL>(A + (k * K) * AW + 4 * q, AW, r, e, o);
#pragma unroll
      for (int g = 0; g < K / 4; g++) {
        uint32_t w[4];
        transpose4(e + 4 * g, o + 4 * g, w);
#pragma unroll
        for (int c = 0; c < 4; c++)
          *reinterpret_cast<uint32_t *>(T + (4 * q + c) * TS + k * K + 4 * g) = w[c];
      }
    }
  }
  __syncthreads();

  // ---- 3. horizontal pass over T, transposed back in registers, stored to HBM
  {
    constexpr int NR = FT / K;
    const int ntask = (FT / 4) * NR;
    for (int task = tid; task < ntask; task += FTHREADS) {
      const int q2 = task / NR, k2 = task - q2 * NR;
      uint32_t e[K], o[K];
      vh_run<K, DIL>(T + (k2 * K + R16 - r) * TS + 4 * q2, TS, r, e, o);
      const int gx = x0 + k2 * K;
      uint32_t rows[4][K / 4];
#pragma unroll
      for (int g = 0; g < K / 4; g++) {
        uint32_t w[4];
        transpose4(e + 4 * g, o + 4 * g, w);
#pragma unroll
        for (int c = 0; c < 4; c++) rows[c][g] = w[c];
      }
      // W is a multiple of 16: a 16-byte (or 4-byte) chunk starting inside the row is whole
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int gy = y0 + 4 * q2 + c;
        if (gy >= H) continue;
        uint8_t *dst = out + ibase + (int64_t)gy * W + gx;
        if (K >= 16) {
#pragma unroll
          for (int hh = 0; hh < K / 16; hh++)
            if (gx + 16 * hh < W)
              *reinterpret_cast<uint4 *>(dst + 16 * hh) =
                  make_uint4(rows[c][4 * hh], rows[c][4 * hh + 1], rows[c][4 * hh + 2], rows[c][4 * hh + 3]);
        } else {
#pragma unroll
          for (int g = 0; g < K / 4; g++)
            if (gx + 4 * g < W) *reinterpret_cast<uint32_t *>(dst + 4 * g) = rows[c][g];
        }
      }
    }
  }
}

template <bool DIL>
static int sq_u8_once(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int r, hipStream_t s)
{
  const int R16 = (r + 15) & ~15;
  const int tiles_x = (int)((W + FT - 1) / FT), tiles_y = (int)((H + FT - 1) / FT);
  const int64_t per = (int64_t)tiles_x * tiles_y;
  const int64_t ntiles = per * B;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "image too large for the fused square kernel");
  const size_t lds = (size_t)(FT + 2 * r) * (FT + 2 * R16) + (size_t)(FT + 2 * R16) * (FT + TPAD);
  const dim3 grid((unsigned)ntiles), block(FTHREADS);
#define PM_SQ_LAUNCH(KK)                                                                                     \
  hipLaunchKernelGGL((sq_u8_kernel<KK, DIL>), grid, block, lds, s, in, out, (int)H, (int)W, r, R16, tiles_x, \
                     (int)per, (int)ntiles)
  if (r >= 16) PM_SQ_LAUNCH(32);
  else if (r >= 8) PM_SQ_LAUNCH(16);
  else if (r >= 4) PM_SQ_LAUNCH(8);
  else PM_SQ_LAUNCH(4);
#undef PM_SQ_LAUNCH
  PM_HIP(hipGetLastError());
  return PM_OK;
}

bool sq_u8_applicable(const void *in, const void *out, int64_t H, int64_t W)
{
  return (W % 16) == 0 && W > 0 && H > 0 && H <= INT32_MAX / 2 && W <= INT32_MAX / 2 &&
         ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0;
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
