// sq_u16.hpp -- fused square erosion / dilation of uint16 images by a (2r+1)^2 square,
// r <= SQ_U16_RMAX, in one launch: Erode/Dilate(img, SESquare2D, r) (ero_dil_static.cpp:229-305,
// the square as a DilateX + DilateY line pair, ero_dil_line_se_t.hpp:49-280).
//
// The uint8 tile kernel (sq_u8_tile.hpp) on 2-byte pixels.  One workgroup (256 threads)
// owns an output tile of TH rows x 128 columns (256-byte row segments), in a persistent
// XCD-aware loop over tiles (tile_sched):
//  1. the region (TH + 2r rows, 128 + 2*PP columns, PP = 8 or 16 pixels so that chunks of
//     aligned rows are 16-byte aligned) is read with all of a thread's 16-byte loads in
//     flight and copied to LDS; pixels outside the image get the neutral value;
//  2. vertical pass: a thread owns a column PAIR (one 32-bit word, two pixels) and a run of
//     K output rows and computes the window with a single-split van Herk run (vh_run_l:
//     suffix over the run, prefix over the rest), one v_pk_max_u16 / v_pk_min_u16 per two
//     pixels.  Results are held in registers across a barrier, transposed 2 x 2 (rows j,
//     j+1 of columns 2q, 2q+1 -> one word per column, v_perm_b32) and written into the
//     transposed buffer T (one row of TH pixels per region column), which reuses the
//     region's LDS;
//  3. horizontal pass = the same run over T (a thread owns a word = a pair of tile rows),
//     transposed back 2 x 2 in registers and stored with 8- or 16-byte stores.
// T is XOR-swizzled: word w of T row R sits at word w ^ ((R >> 1) & (TH/2 - 1)), so that
// the pass-2 writes (lanes on consecutive column pairs, T rows 2 apart) hit distinct banks.
#pragma once
#include "pm_common.hpp"
#include "sq_gen.hpp"
#include "sq_u8_kernel.hpp"

namespace pm {

constexpr int SQ_U16_RMAX = 16;  // largest radius of one launch
constexpr int U16_TW = 128;      // output tile columns (256 bytes)
// tile rows: short tiles for small radii (less LDS and fewer registers, more resident
// workgroups), 128 rows where the halo of a short tile costs more
constexpr int u16_th(int r) { return r <= 3 ? 64 : 128; }
// run length of both passes (K <= 2r, or the direct three-tap form at r = 1)
constexpr int u16_k(int r) { return r >= 8 ? 16 : r >= 4 ? 8 : 4; }
// region columns left and right of the tile: a multiple of 8 pixels (16 bytes), >= r
constexpr int u16_pad(int r) { return r <= 8 ? 8 : 16; }
inline size_t sq_u16_lds_bytes(int r)
{
  const size_t a = (size_t)(u16_th(r) + 2 * r) * 2 * (U16_TW + 2 * u16_pad(r));
  const size_t t = (size_t)(U16_TW + 2 * u16_pad(r)) * 2 * u16_th(r);
  return a > t ? a : t;
}

// byte offset of word w of T row R (rows of TH pixels)
template <int TH> __device__ __forceinline__ int u16_toff(int R, int w)
{
  return R * (2 * TH) + ((w ^ ((R >> 1) & (TH / 2 - 1))) << 2);
}

// 2 x 2 transpose of u16 pairs: a = (x0, x1), b = (y0, y1) -> lo = (x0, y0), hi = (x1, y1)
__device__ __forceinline__ uint32_t pair_lo(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x05040100u); }
__device__ __forceinline__ uint32_t pair_hi(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x07060302u); }

// region rows y0-R .. y0+TH+R-1, columns x0-PP .. x0+TW+PP-1 -> smem (row pitch 2*(TW+2PP)
// bytes).  ALIGNED: W % 8 == 0 and src 16-byte aligned (a chunk of 8 pixels is wholly inside
// or outside the image); otherwise 16-byte loads at 2-byte aligned addresses, and the chunk
// that straddles the right edge is the row's last 16 bytes shifted into place.  Chunks never
// straddle column 0 (x0 and PP are multiples of 8).
template <bool DIL, bool ALIGNED, int R, int TW, int TH, int PP>
__device__ __forceinline__ void u16_load(uint8_t *smem, const uint16_t *__restrict__ src, int H, int W, int y0, int x0,
                                         int tid)
{
  constexpr int cpr = (TW + 2 * PP) / 8, AWB = 2 * (TW + 2 * PP);
  constexpr int nchunk = (TH + 2 * R) * cpr, MAXC = (nchunk + FTHREADS - 1) / FTHREADS;
  const uint32_t nv = DIL ? 0u : 0xFFFFFFFFu;
  const uint8_t *sb = reinterpret_cast<const uint8_t *>(src);
  const int64_t WB = 2 * (int64_t)W;
  const int gx0 = x0 - PP;
  const int ce = ((W - gx0) & 7) ? (W - gx0) >> 3 : -1;  // chunk holding column W, if it straddles
  uint4 v[MAXC];
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    v[i] = make_uint4(nv, nv, nv, nv);
    if (c < nchunk) {
      const int ri = c / cpr, cc = c - ri * cpr;
      const int gy = y0 - R + ri, gx = gx0 + 8 * cc;
      if (ALIGNED) {
        if ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W)
          v[i] = *reinterpret_cast<const uint4 *>(sb + gy * WB + 2 * gx);
      } else {
        if ((unsigned)gy < (unsigned)H && gx >= 0 && gx + 8 <= W) v[i] = load16_any(sb + gy * WB + 2 * gx);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    if (c < nchunk) {
      const int ri = c / cpr, cc = c - ri * cpr;
      if (ALIGNED || cc != ce) *reinterpret_cast<uint4 *>(smem + ri * AWB + 16 * cc) = v[i];
    }
  }
  if (!ALIGNED && ce >= 0 && ce < cpr) {
    for (int ri = tid; ri < TH + 2 * R; ri += FTHREADS) {
      const int gy = y0 - R + ri;
      uint4 e = make_uint4(nv, nv, nv, nv);
      if ((unsigned)gy < (unsigned)H) e = load_edge_chunk(sb + gy * WB, (int)WB, 2 * (gx0 + 8 * ce), DIL ? 0 : 255);
      *reinterpret_cast<uint4 *>(smem + ri * AWB + 16 * ce) = e;
    }
  }
}

// K pixels of one output row at column gx (p = the row's K/2 pixel pairs)
template <int K, bool ALIGNED>
__device__ __forceinline__ void u16_store_run(uint16_t *__restrict__ d, const uint32_t (&p)[K / 2], int gx, int W,
                                              bool whole)
{
  uint8_t *b = reinterpret_cast<uint8_t *>(d);
  if (whole || gx + K <= W) {
    if constexpr (K == 4) {
      if (ALIGNED) *reinterpret_cast<uint2 *>(b) = make_uint2(p[0], p[1]);
      else {
        store4_any(b, p[0]);
        store4_any(b + 4, p[1]);
      }
    } else {
#pragma unroll
      for (int h = 0; h < K / 8; h++) {
        if (ALIGNED) *reinterpret_cast<uint4 *>(b + 16 * h) = make_uint4(p[4 * h], p[4 * h + 1], p[4 * h + 2], p[4 * h + 3]);
        else store16_any(b + 16 * h, p[4 * h], p[4 * h + 1], p[4 * h + 2], p[4 * h + 3]);
      }
    }
    return;
  }
  // the run crosses the right edge: pixel stores, clipped
#pragma unroll
  for (int j = 0; j < K; j++)
    if (gx + j < W) d[j] = (uint16_t)(p[j >> 1] >> (16 * (j & 1)));
}

// Passes 2 and 3 of one tile whose region is in smem.
template <bool DIL, bool ALIGNED, int R, int TW, int TH, int PP>
__device__ __forceinline__ void u16_passes(uint8_t *smem, uint16_t *__restrict__ dst, int H, int W, int y0, int x0,
                                           int tid)
{
  using L = LatPk16<DIL>;
  constexpr int K = u16_k(R), AWB = 2 * (TW + 2 * PP);
  // ---- 2. vertical pass: column pair q (region pixel columns 2q, 2q+1), output rows
  // k*K .. k*K+K-1; the pairs cover region columns [PP - R, PP + TW + R)
  {
    constexpr int q0 = (PP - R) >> 1, q1 = (PP + TW + R + 1) >> 1, nq = q1 - q0;
    constexpr int ntask = nq * (TH / K), MAXT = (ntask + FTHREADS - 1) / FTHREADS;
    uint32_t res[MAXT][K];
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
        const uint8_t *base = smem + (k * K) * AWB + 4 * q;
        uint32_t o[K];
        vh_run_l<L, uint32_t, K, R>([&](int t) { return lds_word(base + t * AWB); }, o);
#pragma unroll
        for (int j = 0; j < K; j += 2) {
          res[sl][j] = pair_lo(o[j], o[j + 1]);      // column 2q, rows j, j+1
          res[sl][j + 1] = pair_hi(o[j], o[j + 1]);  // column 2q+1
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
#pragma unroll
        for (int j = 0; j < K / 2; j++) {
          const int w = (k * K) / 2 + j;
          *reinterpret_cast<uint32_t *>(smem + u16_toff<TH>(2 * q, w)) = res[sl][2 * j];
          *reinterpret_cast<uint32_t *>(smem + u16_toff<TH>(2 * q + 1, w)) = res[sl][2 * j + 1];
        }
      }
    }
  }
  __syncthreads();
  // ---- 3. horizontal pass: tile rows 2*q2, 2*q2+1 (word q2 of T), output columns
  // k2*K .. k2*K+K-1, whose windows start at T row PP - R + k2*K
  {
    constexpr int NR = TW / K, ntask = (TH / 2) * NR, C0 = PP - R;
    const bool full = (y0 + TH <= H) && (x0 + TW <= W);
    for (int task = tid; task < ntask; task += FTHREADS) {
      const int q2 = task / NR, k2 = task - q2 * NR;
      // T row C0 + k2*K + t has swizzle key (k2*K/2 + ((C0 + t) >> 1)) & (TH/2 - 1): K is even,
      // so the per-row part is known at compile time
      const int m = k2 * (K / 2);
      const uint8_t *Tb = smem + (C0 + k2 * K) * (2 * TH);
      uint32_t o[K];
      vh_run_l<L, uint32_t, K, R>(
          [&](int t) {
            const int u = (C0 + t) >> 1;
            return lds_word(Tb + t * (2 * TH) + ((q2 ^ ((m + u) & (TH / 2 - 1))) << 2));
          },
          o);
      uint32_t a[K / 2], b[K / 2];  // rows 2*q2 and 2*q2+1, pixel pairs
#pragma unroll
      for (int g = 0; g < K / 2; g++) {
        a[g] = pair_lo(o[2 * g], o[2 * g + 1]);
        b[g] = pair_hi(o[2 * g], o[2 * g + 1]);
      }
      const int gy = y0 + 2 * q2, gx = x0 + k2 * K;
      uint16_t *d = dst + (int64_t)gy * W + gx;
      if (full) {
        u16_store_run<K, ALIGNED>(d, a, gx, W, true);
        u16_store_run<K, ALIGNED>(d + W, b, gx, W, true);
      } else if (gx < W) {
        if (gy < H) u16_store_run<K, ALIGNED>(d, a, gx, W, false);
        if (gy + 1 < H) u16_store_run<K, ALIGNED>(d + W, b, gx, W, false);
      }
    }
  }
}

// Thread-index-derived offsets recomputed per tile (tile_tid<true>) rather than hoisted out of
// the tile loop: only where the hoisted offsets (168-193 VGPRs from r = 9 on) leave 2 resident
// workgroups per CU instead of 3.  4096x4096 dilation, us, hoisted vs per tile: r = 1 17.2 /
// 17.7, r = 7 17.5 / 18.0, r = 9 18.2 / 20.0, r = 11 20.8 / 19.7, r = 14 21.4 / 20.5,
// r = 16 18.6 / 20.2.
constexpr bool u16_opq(int r) { return r >= 11 && r <= 14; }
template <bool DIL, bool ALIGNED, int R, bool OPQ = u16_opq(R)>
__global__ __launch_bounds__(FTHREADS) void sq_u16_kernel(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                                          int H, int W, int tiles_x, int tiles_per_img, int ntiles)
{
  extern __shared__ __align__(16) uint8_t smem[];
  constexpr int TH = u16_th(R), PP = u16_pad(R), TW = U16_TW;
  static_assert(R >= 1 && R <= PP && R <= SQ_U16_RMAX, "radius within the pad");
  const TileSched ts = tile_sched(blockIdx.x, gridDim.x, ntiles);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int tid = tile_tid<OPQ>();
    const int img = tile / tiles_per_img;
    const int t_in = tile - img * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * TH, x0 = tx * TW;
    const int64_t ibase = (int64_t)img * H * W;
    u16_load<DIL, ALIGNED, R, TW, TH, PP>(smem, in + ibase, H, W, y0, x0, tid);
    __syncthreads();
    u16_passes<DIL, ALIGNED, R, TW, TH, PP>(smem, out + ibase, H, W, y0, x0, tid);
    __syncthreads();  // the next tile's region overwrites T
  }
}

typedef void (*u16_kernel_t)(const uint16_t *, uint16_t *, int, int, int, int, int);
// kernel of radius r (1 .. SQ_U16_RMAX), instantiated in sq_u16_k*.hip
template <bool DIL, bool ALIGNED> u16_kernel_t sq_u16_kernel_for(int r);
template <> u16_kernel_t sq_u16_kernel_for<true, true>(int r);
template <> u16_kernel_t sq_u16_kernel_for<true, false>(int r);
template <> u16_kernel_t sq_u16_kernel_for<false, true>(int r);
template <> u16_kernel_t sq_u16_kernel_for<false, false>(int r);

}  // namespace pm
