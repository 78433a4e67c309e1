// sq_u8_tile.hpp -- device code of one 128-row output tile of the fused uint8 square
// erosion / dilation (used by the single-image kernel of sq_u8.hip and by the batched
// kernel of batch.hip).  See sq_u8.hip for the algorithm.
#pragma once
#include "pm_common.hpp"

namespace pm {

typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

constexpr int FT = 128;       // default output tile height (and default width)
constexpr int FTHREADS = 256;
// Output tile height of the single-image square kernel for radius r.  Small radii read
// little halo, and 64-row tiles need half the LDS and fewer registers (6 resident
// workgroups per CU instead of 3); from r = 4 on the halo of the shorter tile costs more
// than the occupancy gains.  Measured on 8192x8192 dilation (us, 128 vs 64 rows):
// r = 1 33.6 / 31.3, r = 2 35.7 / 31.6, r = 3 36.9 / 32.4, r = 4 34.6 / 34.2,
// r = 5 35.3 / 35.4, r = 7 36.4 / 36.3, r = 8 33.5 / 36.7.
constexpr int sq_th(int r) { return r <= 3 ? 64 : FT; }
// Transposed buffer T of a tile TH rows high: one TH-byte row per region column, 32-bit
// word w of row R stored at word (w ^ ((R >> 2) & (TH/4 - 1))).  Pass 1 writes rows
// 4q..4q+3 from lanes of consecutive q and pass 2 reads rows k2*K+t at word q2; with the
// swizzle both land on distinct banks (a plain pitch maps them onto 8 banks).
template <int TH = FT> __device__ __forceinline__ int t_off(int R, int w)
{
  return R * TH + ((w ^ ((R >> 2) & (TH / 4 - 1))) << 2);
}

template <bool DIL> __device__ __forceinline__ uint32_t pk(uint32_t a, uint32_t b)
{
  const u16x2_t x = __builtin_bit_cast(u16x2_t, a), y = __builtin_bit_cast(u16x2_t, b);
  return __builtin_bit_cast(uint32_t, DIL ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
}
// The four bytes (c0, c1, c2, c3) of a word are processed as two u16x2 values whose HIGH
// bytes carry the pixels: "odd" = the word itself (c1, c3 in the high bytes) and "even" =
// the word shifted left by 8 (c0, c2 in the high bytes).  The max / min of two u16 values
// has the max / min of their high bytes as its high byte whatever the low bytes hold, so
// one v_pk_max_u16 / v_pk_min_u16 handles two pixels and unpacking costs one shift.
__device__ __forceinline__ uint32_t even_hi(uint32_t w) { return w << 8; }
__device__ __forceinline__ uint32_t odd_hi(uint32_t w) { return w; }

// 4x4 byte transpose of high-byte pairs: e[j] = (c0_j, c2_j), o[j] = (c1_j, c3_j) in bytes
// 1 and 3, j = 0..3 -> w[c] = (c_0, c_1, c_2, c_3) bytes
__device__ __forceinline__ void transpose4(const uint32_t *e, const uint32_t *o, uint32_t *w)
{
  const uint32_t ea = __builtin_amdgcn_perm(e[1], e[0], 0x07030501u);
  const uint32_t eb = __builtin_amdgcn_perm(e[3], e[2], 0x07030501u);
  const uint32_t oa = __builtin_amdgcn_perm(o[1], o[0], 0x07030501u);
  const uint32_t ob = __builtin_amdgcn_perm(o[3], o[2], 0x07030501u);
  w[0] = __builtin_amdgcn_perm(eb, ea, 0x05040100u);
  w[2] = __builtin_amdgcn_perm(eb, ea, 0x07060302u);
  w[1] = __builtin_amdgcn_perm(ob, oa, 0x05040100u);
  w[3] = __builtin_amdgcn_perm(ob, oa, 0x07060302u);
}

// The thread index, optionally (OPQ) as a value the compiler cannot prove loop-invariant.
// In a persistent tile loop every LDS and region offset derived from the thread index alone
// is otherwise hoisted out of the loop and kept live across all passes of every tile: fewer
// instructions per tile, more VGPRs.  Which is faster depends on whether the registers limit
// occupancy (LDS often does first).  Measured on 8192x8192 (us, hoisted vs opaque):
// dilation r = 7 / 31: 38.5 / 47.7 vs 40.8 / 52.1 (hoisted kept); fused opening r = 7: 75.4
// vs 59.7 (opaque passes, 205 -> 86 VGPRs); fused gradient r = 15: 81.0 vs 67.3 (opaque load
// and passes), r = 7: 59.5 vs 63.1.
template <bool OPQ> __device__ __forceinline__ int tile_tid()
{
  int t = threadIdx.x;
  if constexpr (OPQ) __asm__ volatile("" : "+v"(t));
  return t;
}

__device__ __forceinline__ uint32_t lds_word(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }

// load16_any / store16_any / store4_any (unaligned 16- and 4-byte accesses): pm_common.hpp

// Shift the 16 bytes of (v0..v3) by b bytes towards higher (LEFT = true: out byte j =
// in byte j - b) or lower (out byte j = in byte j + b) positions, 0 <= b <= 15, filling
// with the byte value of fill.  b may differ between lanes.
template <bool LEFT>
__device__ __forceinline__ uint4 shift16(const uint4 &v, int b, uint32_t fill)
{
  const uint32_t in[4] = {v.x, v.y, v.z, v.w};
  const int q = b >> 2, rem = b & 3;
  // word m of the input shifted by q whole words (fill outside)
  auto wsh = [&](int m) {
    const int i = LEFT ? m - q : m + q;
    uint32_t x = fill;
#pragma unroll
    for (int t = 0; t < 4; t++) x = (i == t) ? in[t] : x;
    return x;
  };
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (LEFT) {
      const uint32_t hi = wsh(k), lo = wsh(k - 1);  // out = high rem bytes of lo, then low 4-rem of hi
      o[k] = rem ? __builtin_amdgcn_alignbyte(hi, lo, 4 - rem) : hi;
    } else {
      const uint32_t lo = wsh(k), hi = wsh(k + 1);  // out = bytes rem.. of lo, then low rem of hi
      o[k] = rem ? __builtin_amdgcn_alignbyte(hi, lo, rem) : lo;
    }
  }
  return make_uint4(o[0], o[1], o[2], o[3]);
}

// The 16-byte region chunk starting at image column gx < W < gx + 16 of row `row` (the
// chunk that straddles the right edge; tile columns are multiples of 16 bytes apart from
// column 0, so no chunk straddles the left edge): the row's last 16 bytes shifted into
// place with neutral fill; rows narrower than 16 bytes are read bytewise.
__device__ __forceinline__ uint4 load_edge_chunk(const uint8_t *__restrict__ row, int W, int gx, uint8_t nb)
{
  const uint32_t nv = nb * 0x01010101u;
  if (W >= 16) return shift16<false>(load16_any(row + (W - 16)), gx + 16 - W, nv);
  uint32_t wd[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t acc = 0;
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
      const int x = gx + 4 * k + bb;
      const uint32_t byte = ((unsigned)x < (unsigned)W) ? row[x] : nb;
      acc |= byte << (8 * bb);
    }
    wd[k] = acc;
  }
  return make_uint4(wd[0], wd[1], wd[2], wd[3]);
}

// Radius r and run length K (output rows per vertical-pass task) of the fused square
// kernels: 4 for r <= 3, 16 for 4 <= r <= 15 (for r < 8 as two blocks of 8 rows sharing
// their loads, see vh_run), 32 above; the region is padded by R16 = 16 columns on each side
// for K <= 16 (r <= 15) and by 32 for K = 32 (16 <= r <= 32).  Runs of 16 rather than 8
// rows for r = 4..7 read 8 + 4r rows per 16 outputs instead of 16 + 4r and halve the tasks:
// 8192^2 dilation, us, 8-row vs 16-row runs: r = 4 34.9 / 32.4, r = 5 35.7 / 32.9,
// r = 6 36.5 / 33.1, r = 7 36.7 / 33.9.  Every kernel instantiates K with a compile-time
// radius RC, which the two-block form needs.
constexpr int sq_k_of(int r) { return r >= 16 ? 32 : r >= 4 ? 16 : 4; }
// Run length of the four-pass fused opening / closing and of the fused gradient: 8 for
// r = 4..7, whose results they hold in more registers (8192^2, us, 8-row vs 16-row runs:
// opening r = 4 54.3-55.8 / 57.1-57.8, r = 7 60.8-63.5 / 61.6-64.8; gradient r = 7
// 58.6-60.7 / 60.4-63.0).
constexpr int sq_k_open(int r) { return r >= 16 ? 32 : r >= 8 ? 16 : r >= 4 ? 8 : 4; }
template <int K> constexpr int sq_r16() { return K == 32 ? 32 : 16; }
// Tile width (output columns): 256 up to r = 16, 128 above (RC = 0: run-time radius of
// run length K).  Wider tiles make longer contiguous row segments for HBM; above r = 16
// the registers and LDS of the wider region cost more than they save (measured on
// 8192x8192: r = 1 37.9 -> 34.9 us, r = 15 42.5 -> 41.4 us, r = 31 55.5 -> 67.4 us).
constexpr int sq_tw(int K, int RC) { return (RC > 0 ? RC <= 16 : K < 32) ? 256 : 128; }
// largest radius served by run length K
constexpr int sq_rmax_of(int K) { return K == 4 ? 3 : K == 8 ? 7 : K == 16 ? 15 : 32; }
// vertical-pass tasks (column quad x run of K rows) per thread for tiles TW wide: the
// quads cover columns [R16 - r, R16 + TW + r) of the region
constexpr int sq_maxt(int K, int TW, int TH = FT)
{
  const int r = sq_rmax_of(K), R16 = K == 32 ? 32 : 16;
  const int nq = ((R16 + TW + r + 3) >> 2) - ((R16 - r) >> 2);  // largest at the largest radius
  return (nq * (TH / K) + FTHREADS - 1) / FTHREADS;
}
// 16-byte region chunks per thread (region of TH + 2r rows, TW + 2*R16 columns)
constexpr int sq_maxc(int K, int TW, int TH = FT)
{
  const int r = sq_rmax_of(K), R16 = K == 32 ? 32 : 16;
  return ((TH + 2 * r) * ((TW + 2 * R16) >> 4) + FTHREADS - 1) / FTHREADS;
}

// Windowed max/min of 2r+1 rows for K consecutive output rows of one 4-byte column quad.
// row(t) returns the quad's 32-bit word in input row t (output row j covers rows j..j+2r).
// RC > 0: the radius is the compile-time constant RC (r is ignored), so every row offset
// is an immediate and the prefix loop has its exact length.
template <int K, bool DIL, int RC, typename LD>
__device__ __forceinline__ void vh_run(LD &&row, int r_rt, uint32_t (&e)[K], uint32_t (&o)[K])
{
  const uint32_t NEU = DIL ? 0u : 0xFFFFFFFFu;
  const int r = RC > 0 ? RC : r_rt;
  if constexpr (RC > 0 && K == 16 && 2 * RC < K) {
    // 16-row run for 4 <= r <= 7 as two blocks of 8 sharing their loads: outputs 0..7 take
    // the suffix over rows j..7 and the prefix over rows 8..j+2r, outputs 8..15 the suffix
    // over rows j..15 and the prefix over rows 16..j+2r (both splits need 2r >= 8)
    static_assert(2 * RC >= 8, "two-block runs need 2r >= 8");
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const uint32_t w = row(j);
      e[j] = even_hi(w);
      o[j] = odd_hi(w);
    }
    uint32_t pe = NEU, po = NEU;
#pragma unroll
    for (int t = 8; t < 2 * RC; t++) {  // prefix rows every output of block 1 shares
      pe = pk<DIL>(pe, e[t]);
      po = pk<DIL>(po, o[t]);
    }
#pragma unroll
    for (int j = 6; j >= 0; j--) {
      e[j] = pk<DIL>(e[j], e[j + 1]);
      o[j] = pk<DIL>(o[j], o[j + 1]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int t = 2 * RC + j;  // rows 8..15 are still the raw rows here
      uint32_t we, wo;
      if (t < 16) {
        we = e[t];
        wo = o[t];
      } else {
        const uint32_t w = row(t);
        we = even_hi(w);
        wo = odd_hi(w);
      }
      pe = pk<DIL>(pe, we);
      po = pk<DIL>(po, wo);
      e[j] = pk<DIL>(e[j], pe);
      o[j] = pk<DIL>(o[j], po);
    }
#pragma unroll
    for (int j = 14; j >= 8; j--) {
      e[j] = pk<DIL>(e[j], e[j + 1]);
      o[j] = pk<DIL>(o[j], o[j + 1]);
    }
    pe = NEU;
    po = NEU;
#pragma unroll
    for (int t = 16; t < 8 + 2 * RC; t++) {  // prefix rows every output of block 2 shares
      const uint32_t w = row(t);
      pe = pk<DIL>(pe, even_hi(w));
      po = pk<DIL>(po, odd_hi(w));
    }
#pragma unroll
    for (int j = 8; j < 16; j++) {
      const uint32_t w = row(2 * RC + j);
      pe = pk<DIL>(pe, even_hi(w));
      po = pk<DIL>(po, odd_hi(w));
      e[j] = pk<DIL>(e[j], pe);
      o[j] = pk<DIL>(o[j], po);
    }
  } else if (K > 4 || 2 * r >= K) {
#pragma unroll
    for (int j = 0; j < K; j++) {
      const uint32_t w = row(j);
      e[j] = even_hi(w);
      o[j] = odd_hi(w);
    }
#pragma unroll
    for (int j = K - 2; j >= 0; j--) {
      e[j] = pk<DIL>(e[j], e[j + 1]);
      o[j] = pk<DIL>(o[j], o[j + 1]);
    }
    // prefix over rows [K, 2r): at most K rows (2r <= 2K for every radius mapped to K).
    // The loop is unrolled over K rows and predicated, so that its LDS reads issue back to
    // back instead of one per dependent iteration; rows past 2r-1 are read (they lie inside
    // the region since K <= 2r and K <= R16 + r) and ignored.
    uint32_t pe = NEU, po = NEU;
    if constexpr (RC > 0) {
#pragma unroll
      for (int t = 0; t < 2 * RC - K; t++) {
        const uint32_t w = row(K + t);
        pe = pk<DIL>(pe, even_hi(w));
        po = pk<DIL>(po, odd_hi(w));
      }
    } else {
#pragma unroll
      for (int t = 0; t < K; t++) {
        const uint32_t w = row(K + t);
        const bool use = K + t < 2 * r;
        pe = use ? pk<DIL>(pe, even_hi(w)) : pe;
        po = use ? pk<DIL>(po, odd_hi(w)) : po;
      }
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
      const uint32_t w = row(2 * r + j);
      pe = pk<DIL>(pe, even_hi(w));
      po = pk<DIL>(po, odd_hi(w));
      e[j] = pk<DIL>(e[j], pe);
      o[j] = pk<DIL>(o[j], po);
    }
  } else if constexpr (K == 4) {
    // r = 1 (2r < K): direct three-tap form over six rows
    uint32_t xe[6], xo[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
      const uint32_t w = row(t);
      xe[t] = even_hi(w);
      xo[t] = odd_hi(w);
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
      e[j] = pk<DIL>(pk<DIL>(xe[j], xe[j + 1]), xe[j + 2]);
      o[j] = pk<DIL>(pk<DIL>(xo[j], xo[j + 1]), xo[j + 2]);
    }
  }
}

// Horizontal pass of one task (rows 4*q2 .. 4*q2+3 of the tile, output columns
// k2*K .. k2*K+K-1) over the transposed buffer T, whose row c0 holds output column 0's
// window start: rows[c][g] = output word g of tile row 4*q2 + c.
// Windowed max/min along the rows of the transposed buffer T for the task (word q2, output
// rows k2*K .. k2*K+K-1 of T's line order, whose row c0 holds line 0's window start): the
// high-byte pairs e / o of vh_run.
template <int K, bool DIL, int RC, int TH = FT>
__device__ __forceinline__ void sq_tline(const uint8_t *T, int r, int c0, int q2, int k2, uint32_t (&e)[K],
                                         uint32_t (&o)[K])
{
  // T row R0 + t, word q2: the swizzle key (R >> 2) & 31 is (m + u) & 31 with
  // u = ((R0 & 3) + t) >> 2 known at compile time when r is, so the rows of one key
  // group share a base address and differ by immediate offsets.
  const int R0 = k2 * K + c0;
  const int m = k2 * (K / 4) + (c0 >> 2), s0 = c0 & 3;
  const uint8_t *Tb = T + R0 * TH;
  vh_run<K, DIL, RC>(
      [&](int t) {
        const int u = (s0 + t) >> 2;
        return lds_word((Tb + ((q2 ^ ((m + u) & (TH / 4 - 1))) << 2)) + t * TH);
      },
      r, e, o);
}

template <int K, bool DIL, int RC, int TH = FT>
__device__ __forceinline__ void sq_pass3_rows(const uint8_t *T, int r, int c0, int q2, int k2, uint32_t (&rows)[4][K / 4])
{
  uint32_t e[K], o[K];
  sq_tline<K, DIL, RC, TH>(T, r, c0, q2, k2, e, o);
#pragma unroll
  for (int g = 0; g < K / 4; g++) {
    uint32_t w[4];
    transpose4(e + 4 * g, o + 4 * g, w);
#pragma unroll
    for (int c = 0; c < 4; c++) rows[c][g] = w[c];
  }
}

// Store rows[c] (K bytes of tile row 4*q2 + c at output column k2*K) to the image.
// ALIGNED: W % 16 == 0 and dst 16-byte aligned; otherwise the same stores at byte addresses.
template <int K, bool ALIGNED, int TW>
__device__ __forceinline__ void sq_store_rows(const uint32_t (&rows)[4][K / 4], uint8_t *__restrict__ dst_img, int H,
                                              int W, int y0, int x0, int q2, int k2, bool full)
{
  const int gx = x0 + k2 * K;
  uint8_t *d0 = dst_img + (int64_t)(y0 + 4 * q2) * W + gx;
  auto run = [&](uint8_t *dst, int c) {
    if (K >= 16) {
#pragma unroll
      for (int hh = 0; hh < K / 16; hh++) {
        if (ALIGNED)
          *reinterpret_cast<uint4 *>(dst + 16 * hh) =
              make_uint4(rows[c][4 * hh], rows[c][4 * hh + 1], rows[c][4 * hh + 2], rows[c][4 * hh + 3]);
        else
          store16_any(dst + 16 * hh, rows[c][4 * hh], rows[c][4 * hh + 1], rows[c][4 * hh + 2], rows[c][4 * hh + 3]);
      }
    } else {
#pragma unroll
      for (int g = 0; g < K / 4; g++) {
        if (ALIGNED) reinterpret_cast<uint32_t *>(dst)[g] = rows[c][g];
        else store4_any(dst + 4 * g, rows[c][g]);
      }
    }
  };
  if (full) {
#pragma unroll
    for (int c = 0; c < 4; c++) run(d0 + (int64_t)c * W, c);
    return;
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int gy = y0 + 4 * q2 + c;
    if (gy >= H) continue;
    uint8_t *dst = d0 + (int64_t)c * W;
    if (ALIGNED) {
      // W is a multiple of 16: a 16-byte (or 4-byte) chunk starting inside the row is whole
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
    } else if (gx + K <= W) {
      run(dst, c);
    } else {
      // the run crosses the right edge: byte stores, clipped
#pragma unroll
      for (int g = 0; g < K / 4; g++)
#pragma unroll
        for (int bb = 0; bb < 4; bb++)
          if (gx + 4 * g + bb < W) dst[4 * g + bb] = (uint8_t)(rows[c][g] >> (8 * bb));
    }
  }
}

// Passes 2 and 3 of one tile of TH rows and TW columns whose input region (rows
// y0-r .. y0+TH+r-1, columns x0-R16 .. x0+TW+R16-1 of the image) is in smem (row pitch
// AW = TW + 2*R16).  dst is the output image (row pitch W).  ALIGNED: W % 16 == 0 and dst
// 16-byte aligned.  RC > 0: the radius is the compile-time constant RC (r_rt is ignored).
// OPQ: thread-index-derived offsets recomputed per tile (see tile_tid).
template <int K, bool DIL, bool ALIGNED, int RC = 0, int TW = FT, int TH = FT, bool OPQ = false>
__device__ __forceinline__ void sq_tile_passes(uint8_t *smem, uint8_t *__restrict__ dst_img, int H, int W, int r_rt,
                                               int y0, int x0)
{
  constexpr int R16 = sq_r16<K>();
  constexpr int AW = TW + 2 * R16;
  const int r = RC > 0 ? RC : r_rt;
  uint8_t *A = smem;
  uint8_t *T = smem;
  const int tid = tile_tid<OPQ>();
  // ---- 2. vertical pass.  Every thread keeps the (transposed) results of its at most
  // MAXT tasks in registers; after a barrier they are written into T, which reuses the
  // storage of A.  Aliasing the two buffers halves the LDS footprint and doubles the
  // number of resident workgroups.
  {
    constexpr int MAXT = sq_maxt(K, TW, TH);
    const int q0 = (R16 - r) >> 2, q1 = (R16 + TW + r + 3) >> 2;
    const int nq = q1 - q0;
    const int ntask = nq * (TH / K);
    uint32_t res[MAXT][K];
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
        uint32_t e[K], o[K];
        const uint8_t *base = A + (k * K) * AW + 4 * q;
        vh_run<K, DIL, RC>([&](int t) { return lds_word(base + t * AW); }, r, e, o);
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
            *reinterpret_cast<uint32_t *>(T + t_off<TH>(4 * q + c, (k * K >> 2) + g)) = res[sl][4 * g + c];
      }
    }
  }
  __syncthreads();

  // ---- 3. horizontal pass over T, transposed back in registers, stored to HBM
  {
    constexpr int NR = TW / K;
    constexpr int ntask = (TH / 4) * NR;
    // tile wholly inside the image: no bounds checks on the stores
    const bool full = (y0 + TH <= H) && (x0 + TW <= W);
    for (int task = tid; task < ntask; task += FTHREADS) {
      const int q2 = task / NR, k2 = task - q2 * NR;
      uint32_t rows[4][K / 4];
      sq_pass3_rows<K, DIL, RC, TH>(T, r, R16 - r, q2, k2, rows);
      sq_store_rows<K, ALIGNED, TW>(rows, dst_img, H, W, y0, x0, q2, k2, full);
    }
  }
}

// Copy the input region of the tile at (y0, x0) -- rows y0-r .. y0+TH+r-1, columns
// x0-R16 .. x0+TW+R16-1 -- into smem (row pitch TW + 2*R16) with all of a thread's
// 16-byte loads in flight before the LDS writes.  Pixels outside the image get the neutral
// value.  ALIGNED: W % 16 == 0 and src 16-byte aligned (a chunk is wholly inside or
// outside the image); otherwise load16_any, and the chunk that straddles the right edge
// of the image comes from load_edge_chunk.
// PAD (a multiple of 16, default R16) replaces R16 as the column pad; TH is the tile height.
template <int K, bool DIL, bool ALIGNED, int RC = 0, int TW = FT, int PAD = sq_r16<K>(), bool OPQ = false,
          int TH = FT>
__device__ __forceinline__ void sq_tile_load(uint8_t *smem, const uint8_t *__restrict__ src, int H, int W, int r_rt,
                                             int y0, int x0)
{
  constexpr int MAXC = PAD == sq_r16<K>() ? sq_maxc(K, TW, TH)
                                          : ((TH + 2 * (RC > 0 ? RC : sq_rmax_of(K))) * ((TW + 2 * PAD) >> 4) +
                                             FTHREADS - 1) / FTHREADS;
  constexpr int R16 = PAD;
  constexpr int AW = TW + 2 * R16, cpr = AW >> 4;
  const int r = RC > 0 ? RC : r_rt;
  const int nchunk = (TH + 2 * r) * cpr;
  const uint32_t nv = DIL ? 0u : 0xFFFFFFFFu;
  const int gx0 = x0 - R16;  // image column of region column 0 (a multiple of 16)
  const int tid = tile_tid<OPQ>();
  // any width: the chunk of each region row that straddles the right edge (column W) is
  // skipped by the main loop and filled by the edge loop below
  const int ce = ((W - gx0) & 15) ? (W - gx0) >> 4 : -1;
  uint4 v[MAXC];
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    v[i] = make_uint4(nv, nv, nv, nv);
    if (c < nchunk) {
      const int ri = c / cpr, cc = c - ri * cpr;
      const int gy = y0 - r + ri, gx = gx0 + (cc << 4);
      if (ALIGNED) {
        if ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W)
          v[i] = *reinterpret_cast<const uint4 *>(src + (int64_t)gy * W + gx);
      } else {
        if ((unsigned)gy < (unsigned)H && gx >= 0 && gx + 16 <= W) v[i] = load16_any(src + (int64_t)gy * W + gx);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    if (c < nchunk) {
      const int ri = c / cpr, cc = c - ri * cpr;
      if (ALIGNED || cc != ce) *reinterpret_cast<uint4 *>(smem + ri * AW + (cc << 4)) = v[i];
    }
  }
  if (!ALIGNED && ce >= 0 && ce < cpr) {
    for (int ri = tid; ri < TH + 2 * r; ri += FTHREADS) {
      const int gy = y0 - r + ri;
      uint4 e = make_uint4(nv, nv, nv, nv);
      if ((unsigned)gy < (unsigned)H) e = load_edge_chunk(src + (int64_t)gy * W, W, gx0 + (ce << 4), DIL ? 0 : 255);
      *reinterpret_cast<uint4 *>(smem + ri * AW + (ce << 4)) = e;
    }
  }
}


// ---------------------------------------------------------------- fused gradient
// Gradient = dilation - erosion by the same square.  One region serves both: pixels outside
// the image are replaced by the nearest image pixel (rows and columns clamped).  A box
// window that reaches outside the image always contains the clamped-to pixel itself, so
// max and min over the clamped region equal max and min over the window's image part, i.e.
// the reference's neutral padding for both operators at once.

// value of the 16-byte region chunk at column gx of image row `row` with columns clamped
// to [0, W)
template <bool ALIGNED>
__device__ __forceinline__ uint4 load_chunk_clamped(const uint8_t *__restrict__ row, int W, int gx)
{
  if (gx + 16 <= 0) {
    const uint32_t b = row[0] * 0x01010101u;
    return make_uint4(b, b, b, b);
  }
  if (gx >= W) {
    const uint32_t b = row[W - 1] * 0x01010101u;
    return make_uint4(b, b, b, b);
  }
  if (gx >= 0 && gx + 16 <= W) return ALIGNED ? *reinterpret_cast<const uint4 *>(row + gx) : load16_any(row + gx);
  // straddles the right edge (chunks never straddle column 0)
  const uint8_t last = row[W - 1];
  if (W >= 16) return shift16<false>(load16_any(row + (W - 16)), gx + 16 - W, last * 0x01010101u);
  uint32_t wd[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t acc = 0;
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
      const int x = gx + 4 * k + bb;
      const uint32_t byte = x < 0 ? row[0] : x < W ? row[x] : last;
      acc |= byte << (8 * bb);
    }
    wd[k] = acc;
  }
  return make_uint4(wd[0], wd[1], wd[2], wd[3]);
}

// region of the tile at (y0, x0) with clamped rows and columns (see above)
template <int K, bool ALIGNED, int RC, int TW, bool OPQ = false>
__device__ __forceinline__ void sq_tile_load_clamped(uint8_t *smem, const uint8_t *__restrict__ src, int H, int W,
                                                     int r_rt, int y0, int x0)
{
  constexpr int MAXC = sq_maxc(K, TW);
  constexpr int R16 = sq_r16<K>();
  constexpr int AW = TW + 2 * R16, cpr = AW >> 4;
  const int r = RC > 0 ? RC : r_rt;
  const int nchunk = (FT + 2 * r) * cpr;
  const int gx0 = x0 - R16;
  const int tid = tile_tid<OPQ>();
  uint4 v[MAXC];
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    v[i] = make_uint4(0u, 0u, 0u, 0u);
    if (c < nchunk) {
      const int ri = c / cpr, cc = c - ri * cpr;
      const int gy = min(max(y0 - r + ri, 0), H - 1);
      v[i] = load_chunk_clamped<ALIGNED>(src + (int64_t)gy * W, W, gx0 + (cc << 4));
    }
  }
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    if (c < nchunk) {
      const int ri = c / cpr, cc = c - ri * cpr;
      *reinterpret_cast<uint4 *>(smem + ri * AW + (cc << 4)) = v[i];
    }
  }
}

// LDS bytes of the gradient tile: region / max-transposed buffer, then the min-transposed one
inline size_t sq_grad_lds_bytes(int r, int TW)
{
  const int R16 = sq_k_of(r) == 32 ? 32 : 16;
  const size_t a = (size_t)(FT + 2 * r) * (TW + 2 * R16), t = (size_t)(TW + 2 * R16) * FT;
  return (a > t ? a : t) + t;
}

// Passes 2 and 3 of the gradient: the vertical pass computes max and min of each task from
// the same LDS words, the two transposed results go to T_max (reusing the region's
// storage) and T_min (after it), and the horizontal pass subtracts min from max (bytewise,
// with no borrow since max >= min in every byte).
template <int K, bool ALIGNED, int RC, int TW, bool OPQ = false>
__device__ __forceinline__ void sq_tile_grad_passes(uint8_t *smem, uint8_t *__restrict__ dst_img, int H, int W, int r_rt,
                                                    int y0, int x0)
{
  constexpr int R16 = sq_r16<K>();
  constexpr int AW = TW + 2 * R16;
  const int r = RC > 0 ? RC : r_rt;
  uint8_t *A = smem;
  uint8_t *TM = smem;
  const size_t a_bytes = (size_t)(FT + 2 * r) * AW, t_bytes = (size_t)AW * FT;
  uint8_t *Tm = smem + (a_bytes > t_bytes ? a_bytes : t_bytes);
  const int tid = tile_tid<OPQ>();
  {
    constexpr int MAXT = sq_maxt(K, TW);
    const int q0 = (R16 - r) >> 2, q1 = (R16 + TW + r + 3) >> 2;
    const int nq = q1 - q0;
    const int ntask = nq * (FT / K);
    uint32_t rmax[MAXT][K], rmin[MAXT][K];
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
        uint32_t e[K], o[K];
        const uint8_t *base = A + (k * K) * AW + 4 * q;
        vh_run<K, true, RC>([&](int t) { return lds_word(base + t * AW); }, r, e, o);
#pragma unroll
        for (int g = 0; g < K / 4; g++) transpose4(e + 4 * g, o + 4 * g, &rmax[sl][4 * g]);
        vh_run<K, false, RC>([&](int t) { return lds_word(base + t * AW); }, r, e, o);
#pragma unroll
        for (int g = 0; g < K / 4; g++) transpose4(e + 4 * g, o + 4 * g, &rmin[sl][4 * g]);
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
          for (int c = 0; c < 4; c++) {
            const int off = t_off(4 * q + c, (k * K >> 2) + g);
            *reinterpret_cast<uint32_t *>(TM + off) = rmax[sl][4 * g + c];
            *reinterpret_cast<uint32_t *>(Tm + off) = rmin[sl][4 * g + c];
          }
      }
    }
  }
  __syncthreads();
  {
    constexpr int NR = TW / K;
    constexpr int ntask = (FT / 4) * NR;
    const bool full = (y0 + FT <= H) && (x0 + TW <= W);
    for (int task = tid; task < ntask; task += FTHREADS) {
      const int q2 = task / NR, k2 = task - q2 * NR;
      uint32_t rows[4][K / 4], lo[4][K / 4];
      sq_pass3_rows<K, true, RC>(TM, r, R16 - r, q2, k2, rows);
      sq_pass3_rows<K, false, RC>(Tm, r, R16 - r, q2, k2, lo);
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int g = 0; g < K / 4; g++) rows[c][g] -= lo[c][g];
      sq_store_rows<K, ALIGNED, TW>(rows, dst_img, H, W, y0, x0, q2, k2, full);
    }
  }
}

// Passes 2 and 3 of the gradient with ONE transposed buffer: the vertical pass's max and min
// results are held in registers; the max results go to T (reusing the region's storage), the
// horizontal max pass keeps its rows in registers while the min results replace them in T,
// and the horizontal min pass subtracts and stores.  Half the LDS of the two-buffer form, for
// the long runs (K = 32) whose per-thread task counts keep the held rows few.
template <int K, bool ALIGNED, int RC, int TW, bool OPQ = false>
__device__ __forceinline__ void sq_tile_grad_passes1(uint8_t *smem, uint8_t *__restrict__ dst_img, int H, int W,
                                                     int r_rt, int y0, int x0)
{
  constexpr int R16 = sq_r16<K>();
  constexpr int AW = TW + 2 * R16;
  const int r = RC > 0 ? RC : r_rt;
  uint8_t *A = smem;
  uint8_t *T = smem;
  const int tid = tile_tid<OPQ>();
  constexpr int MAXT = sq_maxt(K, TW);
  const int q0 = (R16 - r) >> 2, q1 = (R16 + TW + r + 3) >> 2;
  const int nq = q1 - q0;
  const int ntask = nq * (FT / K);
  uint32_t rmax[MAXT][K], rmin[MAXT][K];
#pragma unroll
  for (int sl = 0; sl < MAXT; sl++) {
    const int task = tid + sl * FTHREADS;
    if (task < ntask) {
      const int k = task / nq, q = q0 + (task - k * nq);
      uint32_t e[K], o[K];
      const uint8_t *base = A + (k * K) * AW + 4 * q;
      vh_run<K, true, RC>([&](int t) { return lds_word(base + t * AW); }, r, e, o);
#pragma unroll
      for (int g = 0; g < K / 4; g++) transpose4(e + 4 * g, o + 4 * g, &rmax[sl][4 * g]);
      vh_run<K, false, RC>([&](int t) { return lds_word(base + t * AW); }, r, e, o);
#pragma unroll
      for (int g = 0; g < K / 4; g++) transpose4(e + 4 * g, o + 4 * g, &rmin[sl][4 * g]);
    }
  }
  // the transposed results of one operator into T
  auto put = [&](const uint32_t (&res)[MAXT][K]) {
#pragma unroll
    for (int sl = 0; sl < MAXT; sl++) {
      const int task = tid + sl * FTHREADS;
      if (task < ntask) {
        const int k = task / nq, q = q0 + (task - k * nq);
#pragma unroll
        for (int g = 0; g < K / 4; g++)
#pragma unroll
          for (int c = 0; c < 4; c++)
            *reinterpret_cast<uint32_t *>(T + t_off(4 * q + c, (k * K >> 2) + g)) = res[sl][4 * g + c];
      }
    }
  };
  constexpr int NR = TW / K;
  constexpr int ntask3 = (FT / 4) * NR, MAXH = (ntask3 + FTHREADS - 1) / FTHREADS;
  __syncthreads();
  put(rmax);
  __syncthreads();
  uint32_t hi[MAXH][4][K / 4];
#pragma unroll
  for (int i = 0; i < MAXH; i++) {
    const int task = tid + i * FTHREADS;
    if (task < ntask3) {
      const int q2 = task / NR, k2 = task - q2 * NR;
      sq_pass3_rows<K, true, RC>(T, r, R16 - r, q2, k2, hi[i]);
    }
  }
  __syncthreads();
  put(rmin);
  __syncthreads();
  const bool full = (y0 + FT <= H) && (x0 + TW <= W);
#pragma unroll
  for (int i = 0; i < MAXH; i++) {
    const int task = tid + i * FTHREADS;
    if (task < ntask3) {
      const int q2 = task / NR, k2 = task - q2 * NR;
      uint32_t lo[4][K / 4];
      sq_pass3_rows<K, false, RC>(T, r, R16 - r, q2, k2, lo);
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int g = 0; g < K / 4; g++) hi[i][c][g] -= lo[c][g];
      sq_store_rows<K, ALIGNED, TW>(hi[i], dst_img, H, W, y0, x0, q2, k2, full);
    }
  }
}

constexpr int SQ_TILE_RMAX = 32;
// LDS bytes of one workgroup for radius r, tile width TW and height TH (region A and
// transposed buffer T alias)
inline size_t sq_tile_lds_bytes(int r, int TW = FT, int TH = FT)
{
  const int R16 = sq_k_of(r) == 32 ? 32 : 16;
  const size_t a = (size_t)(TH + 2 * r) * (TW + 2 * R16), t = (size_t)(TW + 2 * R16) * TH;
  return a > t ? a : t;
}

}  // namespace pm
