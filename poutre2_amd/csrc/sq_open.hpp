// sq_open.hpp -- fused opening / closing of uint8 images by a (2r+1)^2 square (sq_open*.hip).
//
// Opening = Dilate(Erode(in, SESquare2D, r), SESquare2D, r) and closing the dual
// (the compositions of ero_dil_static.cpp:229-305 that pypoutre users chain by hand), in one
// launch: every tile reads its input region once, runs both operators in LDS and writes
// its output once, so the eroded (dilated) image never goes through HBM.
//
// Tile geometry (output tile FO rows x TW columns at (y0, x0), FO = (128 - 2r) rounded down
// to a multiple of the run length K):
//   region   rows y0-2r .. y0+128-1      (128 + 2r rows), columns x0-PAD .. x0+TW+PAD-1,
//            PAD = 2r rounded up to 16; pixels outside the image hold the first operator's
//            neutral value
//   band     rows y0-r .. y0-r+127 (128 rows), columns x0-r .. x0+TW+r-1: the first
//            operator's result, with pixels outside the image replaced by the second
//            operator's neutral value (the reference pads both passes of the composition)
//   output   rows y0 .. y0+FO-1, columns x0 .. x0+TW-1
// Four window passes, each the single-split van Herk run of sq_u8_tile.hpp (vh_run):
//   P1 vertical, first operator:  region (natural layout)  -> T1 (transposed, swizzled)
//   P2 horizontal, first operator: T1 -> T2 (transposed, not re-transposed), band mask
//   P3 horizontal, second operator: T2 -> N (natural layout, transposed back in registers)
//   P4 vertical, second operator:  N -> HBM (coalesced rows: lanes on consecutive quads)
// Every pass holds its results in registers across a barrier and overwrites the storage it
// read, so one buffer of (128 + 2r) x (TW + 2*PAD) bytes serves all four.
#pragma once
#include "pm_common.hpp"
#include "sq_u8_kernel.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

// Largest radius of the fused form.  Beyond it the two-launch composition is faster on
// 8192x8192: its intermediate image stays in the 256 MiB MALL between the launches, while the
// fused tile's halo (2r rows per 128-2r output rows) and its four passes grow with r
// (us, fused vs composition: r = 1 47.6 vs 58.1, r = 4 50.6 vs 59.7, r = 7 58.7 vs 64.1,
// r = 8 60.7 vs 58.5, r = 15 76.5 vs 64.6).
constexpr int64_t SQ_U8_OPEN_RMAX = 7;
constexpr int OPEN_TW = 256;

constexpr int open_pad(int r) { return 2 * r <= 16 ? 16 : ((2 * r + 15) / 16) * 16; }
constexpr int open_fo(int r) { return (FT - 2 * r) & ~(sq_k_open(r) - 1); }
inline size_t sq_open_lds_bytes(int r) { return (size_t)(FT + 2 * r) * (OPEN_TW + 2 * open_pad(r)); }

// N: natural-layout band (128 rows, row pitch TW).  Word w of row R is stored at
// w ^ ((R >> 2) & (K/4 - 1)), so that P3's stores (lanes over runs k2 and row groups q2) hit
// distinct banks; P4's loads (lanes over consecutive words of one row) stay conflict-free.
template <int K, int TW> __device__ __forceinline__ int n_off(int R, int w)
{
  return R * TW + ((w ^ ((R >> 2) & (K / 4 - 1))) << 2);
}

// the word (c0, c1, c2, c3) from its high-byte pairs e = (c0, c2), o = (c1, c3)
__device__ __forceinline__ uint32_t join_hi(uint32_t e, uint32_t o) { return __builtin_amdgcn_perm(o, e, 0x07030501u); }

template <int K, bool OPEN, bool ALIGNED, int RC, int TW>
__device__ __forceinline__ void sq_open_tile(uint8_t *smem, uint8_t *__restrict__ dst, int H, int W, int y0, int x0)
{
  constexpr bool D1 = !OPEN, D2 = OPEN;  // an opening erodes first
  constexpr int r = RC, PAD = open_pad(RC), AW = TW + 2 * PAD, FO = open_fo(RC);
  static_assert(PAD >= 2 * r && FO >= K && FT % K == 0 && TW % K == 0, "tile geometry");
  const int tid = tile_tid<true>();  // 205 -> 78..114 VGPRs (see tile_tid)
  uint8_t *L = smem;

  // ---- P1: vertical first operator over region rows 0 .. 127+2r -> band rows 0..127 for
  // region columns [PAD-2r, PAD+TW+2r), written transposed (T1 row = region column)
  {
    constexpr int q0 = (PAD - 2 * r) >> 2, q1 = (PAD + TW + 2 * r + 3) >> 2, nq = q1 - q0;
    constexpr int ntask = nq * (FT / K), MAXT = (ntask + FTHREADS - 1) / FTHREADS;
    static_assert(4 * q1 <= AW, "P1 quads inside the region");
    uint32_t res[MAXT][K];
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
        uint32_t e[K], o[K];
        const uint8_t *base = L + (k * K) * AW + 4 * q;
        vh_run<K, D1, RC>([&](int t) { return lds_word(base + t * AW); }, r, e, o);
#pragma unroll
        for (int g = 0; g < K / 4; g++) transpose4(e + 4 * g, o + 4 * g, &res[sl][4 * g]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
#pragma unroll
        for (int g = 0; g < K / 4; g++)
#pragma unroll
          for (int c = 0; c < 4; c++)
            *reinterpret_cast<uint32_t *>(L + t_off(4 * q + c, (k * K >> 2) + g)) = res[sl][4 * g + c];
      }
    }
  }
  __syncthreads();

  // ---- P2: horizontal first operator over T1 -> band columns j = 0 .. TW+2r-1 (T1 rows
  // PAD-2r+j .. PAD+j), written into T2 row j without transposing; band pixels outside
  // the image become the second operator's neutral value.  The last run may compute lines
  // past TW+2r-1 from T1 rows that P1 left unwritten (still inside the region's storage);
  // P3 never reads those lines.
  {
    constexpr int NOUT = TW + 2 * r, NR = (NOUT + K - 1) / K, ntask = (FT / 4) * NR;
    constexpr int MAXT = (ntask + FTHREADS - 1) / FTHREADS;
    static_assert(PAD + NR * K <= AW && NR * K <= AW, "P2 rows inside the region");
    const uint32_t NEU2 = D2 ? 0u : 0xFFFFFFFFu;
    const bool inner = y0 - r >= 0 && y0 - r + FT <= H && x0 - r >= 0 && x0 + TW + r <= W;
    uint32_t res[MAXT][K];
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int q2 = task / NR, k2 = task - q2 * NR;
        uint32_t e[K], o[K];
        sq_tline<K, D1, RC>(L, r, PAD - 2 * r, q2, k2, e, o);
        uint32_t keep = 0xFFFFFFFFu;  // band rows 4*q2 .. 4*q2+3 inside the image
        if (!inner) {
          const int gy = y0 - r + 4 * q2;
#pragma unroll
          for (int b = 0; b < 4; b++)
            if ((unsigned)(gy + b) >= (unsigned)H) keep &= ~(0xFFu << (8 * b));
        }
#pragma unroll
        for (int t = 0; t < K; t++) {
          uint32_t w = join_hi(e[t], o[t]);
          if (!inner) {
            const int gx = x0 - r + k2 * K + t;
            const uint32_t m = (unsigned)gx < (unsigned)W ? keep : 0u;
            w = (w & m) | (NEU2 & ~m);
          }
          res[sl][t] = w;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int q2 = task / NR, k2 = task - q2 * NR;
#pragma unroll
        for (int t = 0; t < K; t++) *reinterpret_cast<uint32_t *>(L + t_off(k2 * K + t, q2)) = res[sl][t];
      }
    }
  }
  __syncthreads();

  // ---- P3: horizontal second operator over T2 -> tile columns 0 .. TW-1 of the 128 band
  // rows, transposed back and written into N (natural layout)
  {
    constexpr int NR = TW / K, ntask = (FT / 4) * NR, MAXT = (ntask + FTHREADS - 1) / FTHREADS;
    uint32_t res[MAXT][4][K / 4];
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int q2 = task / NR, k2 = task - q2 * NR;
        sq_pass3_rows<K, D2, RC>(L, r, 0, q2, k2, res[sl]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int q2 = task / NR, k2 = task - q2 * NR;
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
          for (int g = 0; g < K / 4; g++)
            *reinterpret_cast<uint32_t *>(L + n_off<K, TW>(4 * q2 + c, k2 * (K / 4) + g)) = res[sl][c][g];
      }
    }
  }
  __syncthreads();

  // ---- P4: vertical second operator over N -> output rows 0 .. FO-1 (band rows j .. j+2r),
  // stored with lanes on consecutive 4-byte quads of one row
  {
    constexpr int NQ = TW / 4, NRUN = FO / K, ntask = NQ * NRUN;
    const bool full = y0 + FO <= H && x0 + TW <= W;
    for (int task = tid; task < ntask; task += FTHREADS) {
      const int k = task / NQ, w = task - k * NQ;
      uint32_t e[K], o[K];
      // rows k*K + t: the swizzle key is (t >> 2) & (K/4 - 1) (k*K is a multiple of K)
      const uint8_t *nb = L + (k * K) * TW;
      vh_run<K, D2, RC>(
          [&](int t) { return lds_word(nb + t * TW + ((w ^ ((t >> 2) & (K / 4 - 1))) << 2)); }, r, e, o);
      const int gx = x0 + 4 * w;
      uint8_t *d0 = dst + (int64_t)(y0 + k * K) * W + gx;
#pragma unroll
      for (int t = 0; t < K; t++) {
        const uint32_t v = join_hi(e[t], o[t]);
        uint8_t *d = d0 + (int64_t)t * W;
        if (ALIGNED) {
          if (full || (y0 + k * K + t < H && gx < W)) *reinterpret_cast<uint32_t *>(d) = v;
        } else if (full || (y0 + k * K + t < H && gx + 4 <= W)) {
          store4_any(d, v);  // any alignment: one dword store at the byte address
        } else if (y0 + k * K + t < H) {
#pragma unroll
          for (int bb = 0; bb < 4; bb++)
            if (gx + bb < W) d[bb] = (uint8_t)(v >> (8 * bb));
        }
      }
    }
  }
}

template <int K, bool OPEN, bool ALIGNED, int RC, int TW>
__global__ __launch_bounds__(FTHREADS) void sq_open_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                           int H, int W, int tiles_x, int tiles_per_img, int ntiles)
{
  extern __shared__ __align__(16) uint8_t smem[];
  constexpr int FO = open_fo(RC);
  const TileSched ts = tile_sched(blockIdx.x, gridDim.x, ntiles);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int img = tile / tiles_per_img;
    const int t_in = tile - img * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * FO, x0 = tx * TW;
    const int64_t ibase = (int64_t)img * H * W;
    // region rows y0-2r .. y0+127: the load's rows around a band starting at y0-r
    sq_tile_load<K, !OPEN, ALIGNED, RC, TW, open_pad(RC)>(smem, in + ibase, H, W, RC, y0 - RC, x0);
    __syncthreads();
    sq_open_tile<K, OPEN, ALIGNED, RC, TW>(smem, out + ibase, H, W, y0, x0);
    __syncthreads();
  }
}

typedef void (*open_kernel_t)(const uint8_t *, uint8_t *, int, int, int, int, int);

// kernel of radius r (1 .. SQ_U8_OPEN_RMAX) for aligned rows (sq_open_a.hip) or rows of any
// alignment (sq_open_u.hip)
template <bool OPEN, bool ALIGNED> open_kernel_t sq_open_kernel_for(int r);
template <> open_kernel_t sq_open_kernel_for<true, true>(int r);
template <> open_kernel_t sq_open_kernel_for<false, true>(int r);
template <> open_kernel_t sq_open_kernel_for<true, false>(int r);
template <> open_kernel_t sq_open_kernel_for<false, false>(int r);

// Opening (OPEN) or closing of B images of H x W by the (2r+1)^2 square in one launch
int sq_u8_open_close(bool open, const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int64_t r,
                     hipStream_t s);

}  // namespace pm
