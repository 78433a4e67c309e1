// maxplus.hip -- differentiable float32 dilation / erosion (flat and max-plus) on gfx950.
//
// No reference counterpart: this extends the flat Dilate/Erode of ero_dil.hpp:32-47 to a
// batch of float32 images with a structuring FUNCTION w (K x K, K odd) and a backward pass.
//   dilation: y[q] = max_t ( x[q + t] + w[t] )        erosion: y[q] = min_t ( x[q + t] - w[t] )
//   flat:     w == 0 on the K x K square (the SESquare2D^((K-1)/2) operator of the reference)
// t = (dy, dx) in [-R, R]^2, R = (K-1)/2; taps falling outside the image are skipped (the
// reference's neutral padding).  The selected tap of q is the FIRST maximiser (minimiser) in
// raster order of t (dy outer, dx inner), i.e. torch.max's first-occurrence rule over an
// F.unfold window.  Backward (gather form, deterministic: no atomics, every sum in a fixed order):
//   gx[p] = sum over taps t of [sel(p - t) == t] * gy[p - t]
//   gw[t] = +-sum over q of [sel(q) == t] * gy[q]      (+ for dilation, - for erosion)
// gw is reduced per workgroup (re-selecting backward) or per wave (stored-tap backward) into a
// partials buffer, then summed over those in a fixed order (gw_reduce: many partials first in
// contiguous ranges, one workgroup each, then per tap).
// Memory per pixel: forward reads x, writes y (8 B); the re-selecting backward (pm_maxplus_bwd,
// LDS tiles, taps in raster order) reads x and gy and writes gx (12 B).  The _sel forms store
// the selected tap of every output in the forward (9 B) and read it in the backward instead
// of x (9 B, no re-selection); that backward sums gx over tap rows, then tap columns.  Flat
// windows are separable (row extremum, then column extremum) and, like the stored-tap
// backward, run one wave per strip with DPP lane shifts: 2K work per pixel instead of K^2.
#include <cmath>

#include "maxplus.hpp"
#include "pm_common.hpp"

namespace pm {

// LDS tiles of the max-plus forward and of the re-selecting backward: MTW columns; MTH_F rows
// in the forward, MTH_B in the backward, whose four LDS arrays (inputs with a 2R halo,
// gradients, selected taps) limit residency (measured on [64,32,256,256] 3x3: forward
// 16 / 32 / 64 rows 0.36 / 0.32 / 0.31 ms, backward best at 16-32 rows, 64 rows 1.5-1.6x
// slower)
constexpr int MTW = 64, MTH_F = 64, MTH_B = 32, MTHREADS = 256;

template <bool DIL> __device__ __forceinline__ float pad_value() { return DIL ? -INFINITY : INFINITY; }

// load a (rows x cols) float region whose origin is (y0, x0) in image b into LDS, padding with pad
__device__ __forceinline__ void load_region(float *dst, int rows, int cols, const float *__restrict__ src, int H, int W,
                                            int y0, int x0, float pad)
{
  const int n = rows * cols;
  for (int e = threadIdx.x; e < n; e += MTHREADS) {
    const int i = e / cols, j = e - i * cols;
    const int gy = y0 + i, gx = x0 + j;
    dst[e] = ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) ? src[(int64_t)gy * W + gx] : pad;
  }
}

// load_region for a region whose first column x0 - R lies R <= 4 floats left of a 16-byte
// aligned image column x0 (rows 16-byte aligned: W % 4 == 0 and src aligned, x0 % 4 == 0):
// the region rows are read as aligned float4 chunks starting at image column x0 - 4, one
// 16-byte load per chunk inside the image (element loads with padding for chunks crossing
// an image edge), every element landing in LDS column (image column) - (x0 - R) when that
// lies in [0, cols).  A quarter of the loads and index computations of load_region.
template <int R>
__device__ __forceinline__ void load_region4(float *dst, int rows, int cols, const float *__restrict__ src, int H,
                                             int W, int y0, int x0, float pad)
{
  static_assert(R >= 0 && R <= 4, "the region starts at most one chunk left of x0");
  const int nk = (cols - R + 3) / 4 + 1;  // chunk k: image columns x0 - 4 + 4k .. x0 - 1 + 4k
  const int n = rows * nk;
  for (int e = threadIdx.x; e < n; e += MTHREADS) {
    const int i = e / nk, k = e - i * nk;
    const int gy = y0 + i, g0 = x0 - 4 + 4 * k;
    float v[4];
    if ((unsigned)gy < (unsigned)H && g0 >= 0 && g0 + 4 <= W) {
      const float4 f = *reinterpret_cast<const float4 *>(src + (int64_t)gy * W + g0);
      v[0] = f.x;
      v[1] = f.y;
      v[2] = f.z;
      v[3] = f.w;
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int gx = g0 + u;
        v[u] = ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) ? src[(int64_t)gy * W + gx] : pad;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int c = 4 * k - 4 + R + u;
      if (c >= 0 && c < cols) dst[i * cols + c] = v[u];
    }
  }
}

// selected tap (raster index dy*K + dx) and value of output q whose window starts at row
// pointer p (region pitch pitch); out-of-image taps hold +-inf and are never selected
template <int K, bool DIL, bool FLAT>
__device__ __forceinline__ int select_tap(const float *p, int pitch, const float *w, float &best)
{
  int sel = 0;
  best = pad_value<DIL>();
#pragma unroll
  for (int dy = 0; dy < K; dy++)
#pragma unroll
    for (int dx = 0; dx < K; dx++) {
      const float xv = p[dy * pitch + dx];
      const float v = FLAT ? xv : (DIL ? xv + w[dy * K + dx] : xv - w[dy * K + dx]);
      const bool better = DIL ? (v > best) : (v < best);
      if (better) {
        best = v;
        sel = dy * K + dx;
      }
    }
  return sel;
}

// Max-plus forward over LDS tiles.  A thread computes four horizontally adjacent outputs from
// row segments read with 16-byte LDS loads (NV float4 per tap row instead of 4K scalar reads).
// Without SEL the window reduces to an IEEE max / min chain (v_max3 / v_min3): it ignores NaN
// like a strict comparison, and differs from one only in which zero it returns when +0 and -0
// tie for the extremum, so an output whose extremum is zero is recomputed with the strict
// comparison.
// SEL (a call autograd records) stores the index of the FIRST maximiser in raster order, which
// a strict comparison of every tap in turn finds at 4 VALU operations per tap.  Instead each
// tap row gives its extremum by the max / min chain, and the first row whose extremum
// strictly beats the running best is the first row that holds the window's extremum (zeros of
// either sign compare equal, NaN never wins); that row's taps are then recomputed from LDS and
// the first one equal to the best is the first maximiser, whose own value (sign of zero
// included) is the output.  When no row beats the neutral value (every tap +-inf towards the
// neutral side, or NaN) the strict rule keeps tap 0 and the neutral value, and so does this.
// 5x5 on [64,32,256,256]: ~100 -> ~60 VALU operations per output.
template <int K, bool DIL, bool FLAT, bool SEL>
__global__ __launch_bounds__(MTHREADS) void mp_fwd_kernel(const float *__restrict__ x, const float *__restrict__ wg,
                                                          float *__restrict__ y, uint8_t *__restrict__ selo, int H,
                                                          int W, int tiles_x, int tiles_per_img, int al)
{
  constexpr int R = (K - 1) / 2, RR = MTH_F + K - 1;
  constexpr int RCP = (MTW + K - 1 + 3) & ~3;  // region pitch, a multiple of 4 floats
  constexpr int NV = (4 + K - 1 + 3) / 4;      // float4 per row segment of 4 outputs
  static_assert(MTW % 4 == 0 && MTW - 4 + 4 * NV <= RCP, "row segments stay inside the region row");
  __shared__ __align__(16) float X[RR * RCP];
  __shared__ float WL[SEL ? K * K : 1];  // w for the per-lane row of the SEL search
  const int img = blockIdx.x / tiles_per_img, t = blockIdx.x - img * tiles_per_img;
  const int ty = t / tiles_x, tx = t - ty * tiles_x;
  const int y0 = ty * MTH_F, x0 = tx * MTW;
  const int64_t ib = (int64_t)img * H * W;
  if (al) load_region4<R>(X, RR, RCP, x + ib, H, W, y0 - R, x0, pad_value<DIL>());  // 16-byte aligned rows
  else load_region(X, RR, RCP, x + ib, H, W, y0 - R, x0 - R, pad_value<DIL>());
  float wv[K * K];  // uniform: scalar loads
#pragma unroll
  for (int k = 0; k < K * K; k++) wv[k] = FLAT ? 0.f : wg[k];
  if (SEL && threadIdx.x < K * K) WL[threadIdx.x] = FLAT ? 0.f : wg[threadIdx.x];
  __syncthreads();
  auto tap = [&](float xv, int k) { return FLAT ? xv : (DIL ? xv + wv[k] : xv - wv[k]); };
  for (int e = threadIdx.x; e < MTH_F * (MTW / 4); e += MTHREADS) {
    const int i = e / (MTW / 4), j = (e - i * (MTW / 4)) * 4;
    const int gy = y0 + i, gx = x0 + j;
    if (gy >= H || gx >= W) continue;
    const float *base = &X[i * RCP + j];
    float best[4];
    int sel[4];  // SEL: first the row of the first maximiser (-1: none beats the neutral value)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      best[q] = pad_value<DIL>();
      sel[q] = SEL ? -1 : 0;
    }
#pragma unroll
    for (int dy = 0; dy < K; dy++) {
      float r[4 * NV];
#pragma unroll
      for (int u = 0; u < NV; u++) {
        const float4 f = *reinterpret_cast<const float4 *>(base + dy * RCP + 4 * u);
        r[4 * u] = f.x;
        r[4 * u + 1] = f.y;
        r[4 * u + 2] = f.z;
        r[4 * u + 3] = f.w;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        float m = tap(r[q], dy * K);
#pragma unroll
        for (int dx = 1; dx < K; dx++) m = DIL ? fmaxf(m, tap(r[q + dx], dy * K + dx)) : fminf(m, tap(r[q + dx], dy * K + dx));
        if constexpr (SEL) {
          const bool better = DIL ? (m > best[q]) : (m < best[q]);
          best[q] = better ? m : best[q];
          sel[q] = better ? dy : sel[q];
        } else {
          best[q] = DIL ? fmaxf(best[q], m) : fminf(best[q], m);
        }
      }
    }
    if constexpr (SEL) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = sel[q] < 0 ? 0 : sel[q];
        const float *rp = base + row * RCP + q;
        const float *wr = WL + row * K;
        float yv = best[q];
        int sd = 0;
#pragma unroll
        for (int dx = K - 1; dx >= 0; dx--) {
          const float v = FLAT ? rp[dx] : (DIL ? rp[dx] + wr[dx] : rp[dx] - wr[dx]);
          const bool eq = v == best[q];
          sd = eq ? dx : sd;
          yv = eq ? v : yv;
        }
        const bool none = sel[q] < 0;
        best[q] = none ? pad_value<DIL>() : yv;
        sel[q] = none ? 0 : row * K + sd;
      }
    }
    if constexpr (!SEL) {
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (best[q] == 0.0f) {  // +0 / -0 tie: the strict comparison decides the sign
          float b = pad_value<DIL>();
#pragma unroll
          for (int dy = 0; dy < K; dy++)
#pragma unroll
            for (int dx = 0; dx < K; dx++) {
              const float v = tap(base[dy * RCP + q + dx], dy * K + dx);
              b = (DIL ? (v > b) : (v < b)) ? v : b;
            }
          best[q] = b;
        }
    }
    float *yo = y + ib + (int64_t)gy * W + gx;
    uint8_t *so = SEL ? selo + ib + (int64_t)gy * W + gx : nullptr;
    if (gx + 4 <= W && ((uintptr_t)yo & 15) == 0) {
      *reinterpret_cast<float4 *>(yo) = make_float4(best[0], best[1], best[2], best[3]);
      if (SEL) {
        const uint32_t packed = (uint32_t)sel[0] | ((uint32_t)sel[1] << 8) | ((uint32_t)sel[2] << 16) | ((uint32_t)sel[3] << 24);
        if (((uintptr_t)so & 3) == 0) *reinterpret_cast<uint32_t *>(so) = packed;
        else store4_any(so, packed);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (gx + q < W) {
          yo[q] = best[q];
          if (SEL) so[q] = (uint8_t)sel[q];
        }
    }
  }
}

// Flat forward and flat backward run one wave per strip of WNR rows and 64 - 2R output columns,
// lane l holding input column (first output column) - R + l: a row's K taps are lanes l..l+2R,
// reached by DPP shifts (no LDS, no barrier); the vertical window is a ring of K rows in
// registers.  Lanes 64 - 2R.. only feed their left neighbours.
// WNR: strip rows (the max-plus backward with gw takes 2 * WNR: fewer halo rows per output
// row and fewer per-strip gw partials; measured on [64,32,256,256] 5x5 with the partials in
// LDS bins, backward alone: 32 / 64 / 128 rows 381 / 351-358 / 362 us, while 64-row strips
// slowed the flat 3x3 backward, 0.65 -> 0.71 ms).  WCH: rows whose loads a wave issues
// together (16 gained nothing over 8; 16 in the backward alone: 3x3 slightly faster, 5x5
// and 7x7 slower).
constexpr int WNR = 32, WCH = 8;

// lane i receives lane i + 1's value (lane 63 receives 0; no output lane reads it)
__device__ __forceinline__ float from_right(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ int from_right(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }

struct WaveStrip {
  int img, r0, ox, c, cc, lane;
  bool cin;
};

// the strip of wave w: image, first row, this lane's output column ox and input column c
template <int R, int NR>
__device__ __forceinline__ bool wave_strip(int nwaves, int strips_x, int strips_per_img, int W, WaveStrip &ws)
{
  const int wave = blockIdx.x * (MTHREADS / 64) + (threadIdx.x >> 6);
  if (wave >= nwaves) return false;
  ws.lane = threadIdx.x & 63;
  ws.img = wave / strips_per_img;
  const int t = wave - ws.img * strips_per_img;
  const int sy = t / strips_x, sx = t - sy * strips_x;
  ws.r0 = sy * NR;
  ws.ox = sx * (64 - 2 * R) + ws.lane;
  ws.c = ws.ox - R;
  ws.cin = (unsigned)ws.c < (unsigned)W;
  ws.cc = ws.c < 0 ? 0 : (ws.c >= W ? W - 1 : ws.c);
  return true;
}

// Flat forward, separable: each row of the window gives its extremum over the K taps and the
// first arg dx; the column pass takes the first row dy whose value beats the running best.
// That row holds the window's extremum and no earlier row does, so dy * K + dx is the first
// maximiser in raster order (the tap select_tap finds) and y has the same bits.  2K compares
// per output instead of K^2.
template <int K, bool DIL, bool SEL>
__global__ __launch_bounds__(MTHREADS) void mp_fwd_flat_kernel(const float *__restrict__ x, float *__restrict__ y,
                                                               uint8_t *__restrict__ selo, int H, int W, int strips_x,
                                                               int strips_per_img, int nwaves)
{
  constexpr int R = (K - 1) / 2;
  WaveStrip ws;
  if (!wave_strip<R, WNR>(nwaves, strips_x, strips_per_img, W, ws)) return;  // whole waves only
  const int64_t ib = (int64_t)ws.img * H * W;
  const float *xi = x + ib;
  const bool out_lane = ws.lane < 64 - 2 * R && ws.ox < W;
  auto load = [&](int r) {
    const int rr = r < 0 ? 0 : (r >= H ? H - 1 : r);
    const float v = xi[(int64_t)rr * W + ws.cc];
    return (ws.cin && (unsigned)r < (unsigned)H) ? v : pad_value<DIL>();
  };
  float m[K];
  int a[K];
  auto row_pass = [&](float v, float &mo, int &ao) {
    float best = pad_value<DIL>();
    int arg = 0;
#pragma unroll
    for (int dx = 0; dx < K; dx++) {
      if (dx) v = from_right(v);
      const bool better = DIL ? (v > best) : (v < best);
      best = better ? v : best;
      arg = better ? dx : arg;
    }
    mo = best;
    ao = arg;
  };
  {
    float v[K - 1];  // all loads issued before the first use
#pragma unroll
    for (int k = 0; k < K - 1; k++) v[k] = load(ws.r0 - R + k);
#pragma unroll
    for (int k = 0; k < K - 1; k++) row_pass(v[k], m[k], a[k]);
  }
  for (int i0 = 0; i0 < WNR; i0 += WCH) {
    if (ws.r0 + i0 >= H) break;  // uniform over the wave
    float v[WCH];  // WCH rows in flight per wave
#pragma unroll
    for (int u = 0; u < WCH; u++) v[u] = load(ws.r0 + i0 + u + R);
#pragma unroll
    for (int u = 0; u < WCH; u++) {
      const int r = ws.r0 + i0 + u;
      if (r >= H) break;
      row_pass(v[u], m[K - 1], a[K - 1]);
      float best = pad_value<DIL>();
      int sel = 0;
#pragma unroll
      for (int dy = 0; dy < K; dy++) {
        const bool better = DIL ? (m[dy] > best) : (m[dy] < best);
        best = better ? m[dy] : best;
        if (SEL) sel = better ? dy * K + a[dy] : sel;
      }
      if (out_lane) {
        y[ib + (int64_t)r * W + ws.ox] = best;
        if (SEL) selo[ib + (int64_t)r * W + ws.ox] = (uint8_t)sel;
      }
#pragma unroll
      for (int k = 0; k < K - 1; k++) {
        m[k] = m[k + 1];
        a[k] = a[k + 1];
      }
    }
  }
}

template <int K, bool DIL, bool FLAT>
__global__ __launch_bounds__(MTHREADS) void mp_bwd_kernel(const float *__restrict__ x, const float *__restrict__ gy,
                                                          const float *__restrict__ wg, float *__restrict__ gx,
                                                          float *__restrict__ gw_part, int H, int W, int tiles_x,
                                                          int tiles_per_img)
{
  constexpr int R = (K - 1) / 2;
  constexpr int QC = MTW + K - 1, QR = MTH_B + K - 1;            // outputs q whose window reaches the tile
  constexpr int XC = MTW + 2 * (K - 1), XR = MTH_B + 2 * (K - 1);  // inputs those q read
  __shared__ float X[XR * XC];
  __shared__ float G[QR * QC];
  __shared__ unsigned char S[QR * QC];
  __shared__ float w[K * K];
  __shared__ float red[MTHREADS];
  const int img = blockIdx.x / tiles_per_img, t = blockIdx.x - img * tiles_per_img;
  const int ty = t / tiles_x, tx = t - ty * tiles_x;
  const int y0 = ty * MTH_B, x0 = tx * MTW;
  const int64_t ib = (int64_t)img * H * W;
  if (!FLAT && threadIdx.x < K * K) w[threadIdx.x] = wg[threadIdx.x];
  load_region(X, XR, XC, x + ib, H, W, y0 - 2 * R, x0 - 2 * R, pad_value<DIL>());
  // gy of out-of-image q is 0: they receive nothing and contribute nothing
  load_region(G, QR, QC, gy + ib, H, W, y0 - R, x0 - R, 0.0f);
  __syncthreads();
  // selected tap of every q in the tile extended by R (q = (y0 - R + i, x0 - R + j))
  for (int e = threadIdx.x; e < QR * QC; e += MTHREADS) {
    const int i = e / QC, j = e - i * QC;
    float best;
    S[e] = (unsigned char)select_tap<K, DIL, FLAT>(&X[i * XC + j], XC, w, best);
  }
  __syncthreads();
  // gather: p = (y0 + i, x0 + j) receives gy[q] from q = p - t when sel(q) == t
  for (int e = threadIdx.x; e < MTH_B * MTW; e += MTHREADS) {
    const int i = e / MTW, j = e - i * MTW;
    const int py = y0 + i, px = x0 + j;
    float acc = 0.0f;
#pragma unroll
    for (int dy = 0; dy < K; dy++)
#pragma unroll
      for (int dx = 0; dx < K; dx++) {
        // q = p - (dy - R, dx - R); in q-region coordinates (i + R - (dy - R), j + R - (dx - R))
        const int qi = i + 2 * R - dy, qj = j + 2 * R - dx;
        const int qe = qi * QC + qj;
        if (S[qe] == dy * K + dx) acc += G[qe];
      }
    if (py < H && px < W) gx[ib + (int64_t)py * W + px] = acc;
  }
  if (FLAT) return;
  // structuring-function gradient: per-thread partial per tap over this tile's q, then a
  // fixed-order tree over the workgroup
  float part[K * K];
#pragma unroll
  for (int k = 0; k < K * K; k++) part[k] = 0.0f;
  for (int e = threadIdx.x; e < MTH_B * MTW; e += MTHREADS) {
    const int i = e / MTW, j = e - i * MTW;
    const int qe = (i + R) * QC + (j + R);
    const int s = S[qe];
    const float g = G[qe];
#pragma unroll
    for (int k = 0; k < K * K; k++) part[k] += (s == k) ? g : 0.0f;
  }
  float *dst = gw_part + (int64_t)blockIdx.x * K * K;
  for (int k = 0; k < K * K; k++) {
    red[threadIdx.x] = part[k];
    __syncthreads();
    for (int sft = MTHREADS / 2; sft > 0; sft >>= 1) {
      if (threadIdx.x < sft) red[threadIdx.x] += red[threadIdx.x + sft];
      __syncthreads();
    }
    if (threadIdx.x == 0) dst[k] = DIL ? red[0] : -red[0];
    __syncthreads();
  }
}

// Backward from stored taps, separable.  For input p = (r, c + dx - R):
//   gx[p] = sum_dx sum_dy [sel(q) == (dy, dx)] gy[q],  q = (r - dy + R, c)
// Stage A sums, for input row r and the output column c of this lane, gy over the q of column
// c whose tap row lands on r (GR), and keeps their tap column (AD).  When every such q agrees
// on dx (always so for taps of the flat forward: a q's dx is the first arg of row r at column
// c), stage B gathers gx[p] = sum_dx [AD(c) == dx] GR(c) over the lanes c = p + R - dx.  A row
// where some lane mixes tap columns (taps of a structuring function) keeps one sum per tap
// column instead (K^2 work).  Sums run dy then dx in ascending order in both (deterministic,
// the same bits where both apply).
// GW: also the structuring-function gradient gw[t] = +-sum_q [sel(q) == t] gy[q].  Lanes
// R .. R + WOUT - 1 own the q of their column in the strip's rows (every q once over the
// grid); each keeps K^2 partials, a fixed-order butterfly sums them over the wave, and the
// wave's K^2 sums go to gw_part[wave] for gw_reduce_kernel.  For K <= 7 a lane's partials
// live in LDS (one private bin per tap and lane, plus a bin that takes the q a lane does not
// own), and each q adds its gy to its tap's bin by an LDS read, add and write: the same
// additions in the same order as K^2 compare-and-add steps in registers, which cost 3K^2
// VALU operations per pixel (5x5 on [64,32,256,256]: 573 -> 366 us; the LDS float atomic
// ds_add_f32 in place of the read-add-write ran at 915 us).
template <int K, bool GW, bool DIL>
__global__ __launch_bounds__(MTHREADS) void mp_bwd_flat_kernel(const uint8_t *__restrict__ sel,
                                                               const float *__restrict__ gy, float *__restrict__ gx,
                                                               float *__restrict__ gw_part, int H, int W,
                                                               int strips_x, int strips_per_img, int nwaves)
{
  constexpr int R = (K - 1) / 2, NR = GW ? 2 * WNR : WNR;
  constexpr bool GWL = GW && K <= 7;  // partials in LDS bins
  __shared__ float gwb[GWL ? (MTHREADS / 64) * (K * K + 1) * 64 : 1];
  WaveStrip ws;
  if (!wave_strip<R, NR>(nwaves, strips_x, strips_per_img, W, ws)) return;  // whole waves only
  const int64_t ib = (int64_t)ws.img * H * W;
  const bool out_lane = ws.lane < 64 - 2 * R && ws.ox < W;
  // ring k <-> output row r - R + k (q rows of the window of input row r); outside: no tap, gy 0
  int sv[K];
  float gv[K];
  float pw[GW ? K * K : 1];  // with GWL only the wave sums, after the rows
  const bool own = ws.lane >= R && ws.lane < R + 64 - 2 * R;
  // this lane's bin of tap t: bins[t * 64]; t = K^2 takes the q the lane does not own
  float *bins = gwb + (GWL ? (threadIdx.x >> 6) * (K * K + 1) * 64 + ws.lane : 0);
  if (GWL) {
#pragma unroll
    for (int t = 0; t <= K * K; t++) bins[t * 64] = 0.0f;
  } else if (GW) {
#pragma unroll
    for (int t = 0; t < K * K; t++) pw[t] = 0.0f;
  }
  auto load = [&](int q, int &so, float &go) {
    const int qq = q < 0 ? 0 : (q >= H ? H - 1 : q);
    const int64_t e = ib + (int64_t)qq * W + ws.cc;
    const int s8 = sel[e];
    const float g = gy[e];
    const bool in = ws.cin && (unsigned)q < (unsigned)H;
    so = in ? s8 : 255;
    go = in ? g : 0.0f;
  };
#pragma unroll
  for (int k = 0; k < K - 1; k++) load(ws.r0 - R + k, sv[k], gv[k]);
  for (int i0 = 0; i0 < NR; i0 += WCH) {
    if (ws.r0 + i0 >= H) break;  // uniform over the wave
    int cs[WCH];
    float cg[WCH];
#pragma unroll
    for (int u = 0; u < WCH; u++) load(ws.r0 + i0 + u + R, cs[u], cg[u]);
#pragma unroll
    for (int u = 0; u < WCH; u++) {
      const int r = ws.r0 + i0 + u;
      if (r >= H) break;
      sv[K - 1] = cs[u];
      gv[K - 1] = cg[u];
      float GR = 0.0f;
      int ad = -1, bad = 0;
#pragma unroll
      for (int dy = 0; dy < K; dy++) {
        // q = (r + R - dy, c) is ring entry 2R - dy; its tap column when its tap row is dy
        const unsigned d = (unsigned)(sv[2 * R - dy] - dy * K);
        const bool hit = d < (unsigned)K;
        GR += hit ? gv[2 * R - dy] : 0.0f;
        bad |= (int)(hit & (ad >= 0) & (ad != (int)d));
        ad = hit ? (int)d : ad;
      }
      // stage B: lane l + k holds output column ox + k - R, whose tap column 2R - k lands on
      // ox.  With part(dx) the lane's sum for tap column dx, gx = sum over dx ascending of
      // part(dx) from lane l + 2R - dx, accumulated Horner-wise (h = h from lane l + 1 +
      // part(dx), one DPP add per step): lane-wise the same additions in the same order as
      // shifting every part(dx) into place and summing (no part(dx) is -0, so the leading
      // 0 + part(0) of that sum changes nothing).
      float acc;
      if (!__any(bad)) {
        acc = (ad == 0) ? GR : 0.0f;
#pragma unroll
        for (int dx = 1; dx < K; dx++) acc = from_right(acc) + ((ad == dx) ? GR : 0.0f);
      } else {
        float part[K];
#pragma unroll
        for (int dx = 0; dx < K; dx++) part[dx] = 0.0f;
#pragma unroll
        for (int dy = 0; dy < K; dy++) {
          const int d = sv[2 * R - dy] - dy * K;
#pragma unroll
          for (int dx = 0; dx < K; dx++) part[dx] += (d == dx) ? gv[2 * R - dy] : 0.0f;
        }
        acc = part[0];
#pragma unroll
        for (int dx = 1; dx < K; dx++) acc = from_right(acc) + part[dx];
      }
      if (out_lane) gx[ib + (int64_t)r * W + ws.ox] = acc;
      if (GWL) {
        // q = (r, c) is ring entry R; outside the image its tap is 255 and matches nothing
        const int sq = sv[R];
        const int t = (own && (unsigned)sq < (unsigned)(K * K)) ? sq : K * K;
        bins[t * 64] += gv[R];
      } else if (GW) {
        const int sq = sv[R];
        const float g = own ? gv[R] : 0.0f;
#pragma unroll
        for (int t = 0; t < K * K; t++) pw[t] += (sq == t) ? g : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < K - 1; k++) {
        sv[k] = sv[k + 1];
        gv[k] = gv[k + 1];
      }
    }
  }
  if (GW) {
    const int wave = blockIdx.x * (MTHREADS / 64) + (threadIdx.x >> 6);
#pragma unroll
    for (int t = 0; t < K * K; t++) {
      float v = GWL ? bins[t * 64] : pw[t];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      pw[t] = v;
    }
    if (ws.lane == 0) {
#pragma unroll
      for (int t = 0; t < K * K; t++) gw_part[(int64_t)wave * K * K + t] = DIL ? pw[t] : -pw[t];
    }
  }
}

// First stage of the gw reduction over many partials: workgroup j sums the partials of the
// contiguous range of blocks [j*n/G, (j+1)*n/G) for every tap into part2[j][k].  Thread t
// owns tap t % KK and every (256 / KK)-th block of the range, so consecutive threads read
// consecutive floats; the per-thread sums are added in thread order.  Fixed order
// throughout: the same bits on every run.
__global__ __launch_bounds__(256) void gw_reduce1_kernel(const float *__restrict__ part, int nblocks, int KK,
                                                         float *__restrict__ part2)
{
  __shared__ float red[256];
  const int j = blockIdx.x, G = gridDim.x;
  const int b0 = (int)((int64_t)j * nblocks / G), b1 = (int)((int64_t)(j + 1) * nblocks / G);
  const int per = 256 / KK, t = threadIdx.x, k = t % KK, g = t / KK;
  float acc = 0.0f;
  if (g < per)
    for (int b = b0 + g; b < b1; b += per) acc += part[(int64_t)b * KK + k];
  red[t] = acc;
  __syncthreads();
  if (t < KK) {
    float s = 0.0f;
    for (int u = 0; u < per; u++) s += red[u * KK + t];
    part2[(int64_t)j * KK + t] = s;
  }
}

// partials per first-stage workgroup; above GW_DIRECT partials the reduction runs in two stages
constexpr int GW_STAGE1 = 256, GW_DIRECT = 2048;

// gw[k] = sum over blocks b (in order) of part[b][k]: one workgroup per tap, fixed-order tree
__global__ __launch_bounds__(256) void gw_reduce_kernel(const float *__restrict__ part, int nblocks, int KK,
                                                        float *__restrict__ gw)
{
  __shared__ float red[256];
  const int k = blockIdx.x;
  float acc = 0.0f;
  for (int b = threadIdx.x; b < nblocks; b += 256) acc += part[(int64_t)b * KK + k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) gw[k] = red[0];
}

// floats of scratch the gw reduction of n partials of KK taps needs (partials + first stage)
static size_t gw_scratch_floats(int64_t n, int KK)
{
  return (size_t)n * KK + (n > GW_DIRECT ? (size_t)GW_STAGE1 * KK : 0);
}

// gw from n per-block partials at part (scratch of gw_scratch_floats(n, KK) floats)
static int gw_reduce(float *part, int n, int KK, float *gw, hipStream_t s)
{
  const float *src = part;
  int m = n;
  if (n > GW_DIRECT) {
    float *part2 = part + (size_t)n * KK;
    hipLaunchKernelGGL(gw_reduce1_kernel, dim3(GW_STAGE1), dim3(256), 0, s, part, n, KK, part2);
    src = part2;
    m = GW_STAGE1;
  }
  hipLaunchKernelGGL(gw_reduce_kernel, dim3(KK), dim3(256), 0, s, src, m, KK, gw);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

static int mp_check(int64_t B, int64_t H, int64_t W, int K, const void *a, const void *b)
{
  clear_error();
  if (B < 0 || H < 0 || W < 0) PM_FAIL(PM_E_SIZE, "maxplus: negative extent");
  if (K != 3 && K != 5 && K != 7 && K != 9) PM_FAIL(PM_E_ARG, "maxplus: K must be 3, 5, 7 or 9 (got %d)", K);
  if (B * H * W > 0 && (!a || !b)) PM_FAIL(PM_E_ARG, "maxplus: null pointer");
  if (H > INT32_MAX || W > INT32_MAX) PM_FAIL(PM_E_SIZE, "maxplus: image too large");
  return PM_OK;
}

static int64_t mp_blocks(int64_t B, int64_t H, int64_t W, int th, int &tiles_x, int &per)
{
  tiles_x = (int)((W + MTW - 1) / MTW);
  per = tiles_x * (int)((H + th - 1) / th);
  return B * per;
}

// strips of the wave kernels: nr rows x (64 - (K - 1)) output columns, one wave each, four
// waves per workgroup; returns the workgroup count (0: nothing to do, -1: too many)
static int64_t mp_strips(int64_t B, int64_t H, int64_t W, int K, int nr, int &strips_x, int &per, int &nwaves)
{
  const int64_t wout = 64 - (K - 1);
  const int64_t sx = (W + wout - 1) / wout, sy = (H + nr - 1) / nr, n = B * sx * sy;
  if (n == 0) return 0;
  if (n > INT32_MAX - 3) return -1;
  strips_x = (int)sx;
  per = (int)(sx * sy);
  nwaves = (int)n;
  return (n + MTHREADS / 64 - 1) / (MTHREADS / 64);
}

}  // namespace pm

using namespace pm;

#define PM_MP_DISPATCH(MACRO)                                                                                \
  switch (K) {                                                                                               \
  case 3: MACRO(3); break;                                                                                   \
  case 5: MACRO(5); break;                                                                                   \
  case 7: MACRO(7); break;                                                                                   \
  default: MACRO(9); break;                                                                                  \
  }

static int maxplus_fwd(const float *x, int64_t B, int64_t H, int64_t W, const float *w, int K, int dilate, float *y,
                       uint8_t *selo, void *stream)
{
  PM_TRY(mp_check(B, H, W, K, x, y));
  if (x == y && B * H * W > 0) PM_FAIL(PM_E_ALIAS, "maxplus: output must be != input");
  int tiles_x, per, strips_x = 0, sper = 0, nwaves = 0;
  const int64_t nb = mp_blocks(B, H, W, MTH_F, tiles_x, per);
  if (nb == 0) return PM_OK;
  if (nb > INT32_MAX) PM_FAIL(PM_E_SIZE, "maxplus: batch too large");
  hipStream_t s = as_stream(stream);
  const bool flat = (w == nullptr);
  const int al = (W % 4 == 0 && ((uintptr_t)x & 15) == 0) ? 1 : 0;  // rows of 16-byte aligned float4
  const int64_t nbs = mp_strips(B, H, W, K, WNR, strips_x, sper, nwaves);
  if (flat && nbs < 0) PM_FAIL(PM_E_SIZE, "maxplus: batch too large");
#define L(KK, SELV)                                                                                          \
  do {                                                                                                       \
    if (dilate) {                                                                                            \
      if (flat && !al) hipLaunchKernelGGL((mp_fwd_flat_kernel<KK, true, SELV>), dim3((unsigned)nbs), dim3(MTHREADS), 0, s, x, y, selo, (int)H, (int)W, strips_x, sper, nwaves); \
      else if (flat) hipLaunchKernelGGL((mp_fwd_kernel<KK, true, true, SELV>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, w, y, selo, (int)H, (int)W, tiles_x, per, al); \
      else hipLaunchKernelGGL((mp_fwd_kernel<KK, true, false, SELV>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, w, y, selo, (int)H, (int)W, tiles_x, per, al); \
    } else {                                                                                                 \
      if (flat && !al) hipLaunchKernelGGL((mp_fwd_flat_kernel<KK, false, SELV>), dim3((unsigned)nbs), dim3(MTHREADS), 0, s, x, y, selo, (int)H, (int)W, strips_x, sper, nwaves); \
      else if (flat) hipLaunchKernelGGL((mp_fwd_kernel<KK, false, true, SELV>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, w, y, selo, (int)H, (int)W, tiles_x, per, al); \
      else hipLaunchKernelGGL((mp_fwd_kernel<KK, false, false, SELV>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, w, y, selo, (int)H, (int)W, tiles_x, per, al); \
    }                                                                                                        \
  } while (0)
#define LS(KK) L(KK, true)
#define LN(KK) L(KK, false)
  if (selo) {
    PM_MP_DISPATCH(LS)
  } else {
    PM_MP_DISPATCH(LN)
  }
#undef LN
#undef LS
#undef L
  PM_HIP(hipGetLastError());
  return PM_OK;
}

extern "C" int pm_maxplus_fwd(const float *x, int64_t B, int64_t H, int64_t W, const float *w, int K, int dilate,
                              float *y, void *stream)
{ PM_TRACE();
  return maxplus_fwd(x, B, H, W, w, K, dilate, y, nullptr, stream);
}

extern "C" int pm_maxplus_fwd_sel(const float *x, int64_t B, int64_t H, int64_t W, const float *w, int K, int dilate,
                                  float *y, uint8_t *sel, void *stream)
{ PM_TRACE();
  if (!sel && B * H * W > 0) PM_FAIL(PM_E_ARG, "maxplus: null tap buffer");
  return maxplus_fwd(x, B, H, W, w, K, dilate, y, sel, stream);
}

extern "C" int pm_maxplus_bwd_sel(const uint8_t *sel, const float *gy, int64_t B, int64_t H, int64_t W, int K,
                                  int dilate, int with_w, float *gx, float *gw, void *stream)
{ PM_TRACE();
  PM_TRY(mp_check(B, H, W, K, gy, gx));
  if (!sel && B * H * W > 0) PM_FAIL(PM_E_ARG, "maxplus: null tap buffer");
  if (gx == gy && B * H * W > 0) PM_FAIL(PM_E_ALIAS, "maxplus: gx must be a distinct buffer");
  if (with_w && !gw) PM_FAIL(PM_E_ARG, "maxplus: gw buffer required for a structuring function");
  hipStream_t s = as_stream(stream);
  int strips_x = 0, sper = 0, nwaves = 0;
  const int64_t nbs = mp_strips(B, H, W, K, with_w ? 2 * WNR : WNR, strips_x, sper, nwaves);
  if (nbs < 0) PM_FAIL(PM_E_SIZE, "maxplus: batch too large");
  if (nbs == 0) {
    if (with_w) return fill_bytes(gw, 0, sizeof(float) * K * K, s);
    return PM_OK;
  }
  Scratch part;
  if (with_w) PM_TRY(part.alloc(sizeof(float) * gw_scratch_floats(nwaves, K * K), s));
  float *pp = (float *)part.ptr;
#define L(KK)                                                                                                \
  do {                                                                                                       \
    if (!with_w) hipLaunchKernelGGL((mp_bwd_flat_kernel<KK, false, true>), dim3((unsigned)nbs), dim3(MTHREADS), 0, s, sel, gy, gx, pp, (int)H, (int)W, strips_x, sper, nwaves); \
    else if (dilate) hipLaunchKernelGGL((mp_bwd_flat_kernel<KK, true, true>), dim3((unsigned)nbs), dim3(MTHREADS), 0, s, sel, gy, gx, pp, (int)H, (int)W, strips_x, sper, nwaves); \
    else hipLaunchKernelGGL((mp_bwd_flat_kernel<KK, true, false>), dim3((unsigned)nbs), dim3(MTHREADS), 0, s, sel, gy, gx, pp, (int)H, (int)W, strips_x, sper, nwaves); \
  } while (0)
  PM_MP_DISPATCH(L)
#undef L
  PM_HIP(hipGetLastError());
  if (with_w) return gw_reduce(pp, nwaves, K * K, gw, s);
  return PM_OK;
}

extern "C" int pm_maxplus_bwd(const float *x, const float *gy, int64_t B, int64_t H, int64_t W, const float *w, int K,
                              int dilate, float *gx, float *gw, void *stream)
{ PM_TRACE();
  PM_TRY(mp_check(B, H, W, K, x, gx));
  if (!gy && B * H * W > 0) PM_FAIL(PM_E_ARG, "maxplus: null gradient");
  if ((gx == x || gx == gy) && B * H * W > 0) PM_FAIL(PM_E_ALIAS, "maxplus: gx must be a distinct buffer");
  const bool flat = (w == nullptr);
  if (!flat && !gw) PM_FAIL(PM_E_ARG, "maxplus: gw buffer required for a structuring function");
  int tiles_x, per;
  const int64_t nb = mp_blocks(B, H, W, MTH_B, tiles_x, per);
  hipStream_t s = as_stream(stream);
  if (nb == 0) {
    if (!flat) return fill_bytes(gw, 0, sizeof(float) * K * K, s);
    return PM_OK;
  }
  if (nb > INT32_MAX) PM_FAIL(PM_E_SIZE, "maxplus: batch too large");
  Scratch part;
  if (!flat) PM_TRY(part.alloc(sizeof(float) * gw_scratch_floats(nb, K * K), s));
  float *pp = (float *)part.ptr;
#define L(KK)                                                                                                \
  do {                                                                                                       \
    if (dilate) {                                                                                            \
      if (flat) hipLaunchKernelGGL((mp_bwd_kernel<KK, true, true>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, gy, w, gx, pp, (int)H, (int)W, tiles_x, per); \
      else hipLaunchKernelGGL((mp_bwd_kernel<KK, true, false>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, gy, w, gx, pp, (int)H, (int)W, tiles_x, per); \
    } else {                                                                                                 \
      if (flat) hipLaunchKernelGGL((mp_bwd_kernel<KK, false, true>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, gy, w, gx, pp, (int)H, (int)W, tiles_x, per); \
      else hipLaunchKernelGGL((mp_bwd_kernel<KK, false, false>), dim3((unsigned)nb), dim3(MTHREADS), 0, s, x, gy, w, gx, pp, (int)H, (int)W, tiles_x, per); \
    }                                                                                                        \
  } while (0)
  PM_MP_DISPATCH(L)
#undef L
  PM_HIP(hipGetLastError());
  if (!flat) return gw_reduce(pp, (int)nb, K * K, gw, s);
  return PM_OK;
}
