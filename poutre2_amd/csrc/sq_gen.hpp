// sq_gen.hpp -- fused square erosion / dilation for int32, int64, float32 and float64
// images: Erode/Dilate(img, SESquare2D, r) (ero_dil_static.cpp:229-305, the (2r+1)^2 square
// as a 2D line pass pair, ero_dil_line_se_t.hpp:49-280) in one launch per radius up to
// SQ_GEN_RMAX (larger radii chain launches: a square of side 2a+1 followed by one of side
// 2b+1 is the square of side 2(a+b)+1, also at the image borders).  uint16 images have a
// kernel of their own (sq_u16.hpp), which uses the run helpers below on pixel pairs.
//
// One workgroup (256 threads) owns an output tile of 64 rows x TWE columns (gen_twe), in a
// persistent loop over tiles (XCD-aware order, sq_u8_kernel.hpp):
//  1. the region (64+2r rows x TWE+2r columns) is read into LDS with an odd row pitch, all of
//     a thread's loads in flight; pixels outside the image hold the operator's neutral value.
//     Rows of 16-byte aligned chunks (al: W a multiple of 16 / sizeof(T), 16-byte aligned
//     buffer) are read as aligned 16-byte chunks, one load per chunk inside the image;
//     otherwise element by element, consecutive lanes on consecutive columns (4096^2 side 15
//     with chunks: float32 43.1 -> 40.9 us, int32 40.4 -> 38.1 us; side 3 unchanged);
//  2. horizontal pass: a lane owns one region row and a run of K output columns (lanes on
//     consecutive rows, so the odd pitch puts them on distinct banks) and computes the
//     window with a single-split van Herk run (suffix over the run, prefix over the rest:
//     about three operations per pixel whatever r is); results are held in registers across
//     a barrier and written over the region (pitch gen_pitch2: TWE + 1);
//  3. vertical pass: a lane owns one column and a run of K output rows (lanes on
//     consecutive columns: conflict-free reads, and each store instruction writes one
//     contiguous 256-byte row segment of the tile).
#pragma once
#include "pm_common.hpp"
#include "sq_u8_kernel.hpp"

namespace pm {

constexpr int SQ_GEN_RMAX = 16;  // largest radius of one launch
constexpr int GEN_FT = 64;       // output tile rows

// tile columns: 256-byte row segments
template <typename T> constexpr int gen_twe() { return 256 / (int)sizeof(T); }
constexpr int gen_k(int r) { return r >= 8 ? 16 : r >= 4 ? 8 : 4; }
// region pitch and the horizontal pass's output pitch: odd (elements)
template <typename T> constexpr int gen_pitch(int r) { return gen_twe<T>() + 2 * r + 1; }
template <typename T> constexpr int gen_pitch2() { return gen_twe<T>() + 1; }
template <typename T> inline size_t gen_lds_bytes(int r)
{
  const size_t a = (size_t)(GEN_FT + 2 * r) * gen_pitch<T>(r), b = (size_t)(GEN_FT + 2 * r) * gen_pitch2<T>();
  return (a > b ? a : b) * sizeof(T);
}

// the lattice of a dilation / erosion on two uint16 pixels packed in one 32-bit word
template <bool DIL> struct LatPk16 {
  typedef unsigned short v2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ uint32_t op(uint32_t a, uint32_t b)
  {
    const v2 x = __builtin_bit_cast(v2, a), y = __builtin_bit_cast(v2, b);
    return __builtin_bit_cast(uint32_t, DIL ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
  }
  static __device__ __forceinline__ uint32_t neutral() { return DIL ? 0u : 0xFFFFFFFFu; }
};

// out[j] = op over row(j .. j+2R), j < K, in the lattice L
template <typename L, typename T, int K, int R, typename LD> __device__ __forceinline__ void vh_run_l(LD &&row, T (&out)[K])
{
  if constexpr (2 * R >= K) {
#pragma unroll
    for (int j = 0; j < K; j++) out[j] = row(j);
#pragma unroll
    for (int j = K - 2; j >= 0; j--) out[j] = L::op(out[j], out[j + 1]);  // suffix over [j, K)
    T p = L::neutral();
#pragma unroll
    for (int t = K; t < 2 * R; t++) p = L::op(p, row(t));  // prefix over [K, 2R)
#pragma unroll
    for (int j = 0; j < K; j++) {
      p = L::op(p, row(2 * R + j));  // [K, 2R + j]
      out[j] = L::op(out[j], p);
    }
  } else {
    static_assert(R == 1 && K == 4, "direct form for r = 1");
    T x[K + 2];
#pragma unroll
    for (int t = 0; t < K + 2; t++) x[t] = row(t);
#pragma unroll
    for (int j = 0; j < K; j++) out[j] = L::op(L::op(x[j], x[j + 1]), x[j + 2]);
  }
}
template <typename T, int K, bool DIL, int R, typename LD> __device__ __forceinline__ void vh_run_t(LD &&row, T (&out)[K])
{
  vh_run_l<Lat<T, DIL>, T, K, R>(row, out);
}

template <typename T, bool DIL, int R>
__global__ __launch_bounds__(256) void sq_gen_kernel(const T *__restrict__ in, T *__restrict__ out, int H, int W,
                                                     int tiles_x, int tiles_per_img, int ntiles, int al)
{
  using L = Lat<T, DIL>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *S = reinterpret_cast<T *>(smem_raw);
  constexpr int TWE = gen_twe<T>(), K = gen_k(R), RH = GEN_FT + 2 * R, RW = TWE + 2 * R;
  constexpr int P = gen_pitch<T>(R), P2 = gen_pitch2<T>();
  static_assert(TWE % K == 0 && GEN_FT % K == 0, "runs tile the output");
  constexpr int NEL = RH * RW, NL = (NEL + 255) / 256;
  constexpr int NRH = TWE / K, HT = RH * NRH, NH = (HT + 255) / 256;
  constexpr int NRV = GEN_FT / K, VT = TWE * NRV;
  const TileSched ts = tile_sched(blockIdx.x, gridDim.x, ntiles);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int tid = tile_tid<true>();
    const int img = tile / tiles_per_img;
    const int t_in = tile - img * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * GEN_FT, x0 = tx * TWE;
    const T *src = in + (int64_t)img * H * W;
    T *dst = out + (int64_t)img * H * W;
    // ---- 1. region rows y0-R .., columns x0-R ..
    if (al) {
      // aligned 16-byte chunks from image column x0 - CA (x0 and CA are multiples of EPC);
      // element u of chunk k lands in region column EPC*k + u + R - CA when inside [0, RW)
      constexpr int EPC = 16 / (int)sizeof(T), CA = (R + EPC - 1) / EPC * EPC;
      constexpr int NK = (RW + CA - R + EPC - 1) / EPC, NCH = RH * NK, NC = (NCH + 255) / 256;
      static_assert(TWE % EPC == 0, "tile columns start on chunk boundaries");
      T v[NC][EPC];
#pragma unroll
      for (int i = 0; i < NC; i++) {
        const int e = tid + i * 256;
#pragma unroll
        for (int u = 0; u < EPC; u++) v[i][u] = L::neutral();
        if (e < NCH) {
          const int row = e / NK, k = e - row * NK;
          const int gy = y0 - R + row, g0 = x0 - CA + EPC * k;
          if ((unsigned)gy < (unsigned)H) {
            const T *p = src + (int64_t)gy * W + g0;
            if (g0 >= 0 && g0 + EPC <= W) {
              const uint4 c = *reinterpret_cast<const uint4 *>(p);
              __builtin_memcpy(v[i], &c, 16);
            } else {
#pragma unroll
              for (int u = 0; u < EPC; u++)
                if ((unsigned)(g0 + u) < (unsigned)W) v[i][u] = p[u];
            }
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NC; i++) {
        const int e = tid + i * 256;
        if (e < NCH) {
          const int row = e / NK, k = e - row * NK;
#pragma unroll
          for (int u = 0; u < EPC; u++) {
            const int c = EPC * k + u + R - CA;
            if (c >= 0 && c < RW) S[row * P + c] = v[i][u];
          }
        }
      }
    } else {
      T v[NL];
#pragma unroll
      for (int i = 0; i < NL; i++) {
        const int e = tid + i * 256;
        v[i] = L::neutral();
        if (e < NEL) {
          const int row = e / RW, col = e - row * RW;
          const int gy = y0 - R + row, gx = x0 - R + col;
          if ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) v[i] = src[(int64_t)gy * W + gx];
        }
      }
#pragma unroll
      for (int i = 0; i < NL; i++) {
        const int e = tid + i * 256;
        if (e < NEL) {
          const int row = e / RW, col = e - row * RW;
          S[row * P + col] = v[i];
        }
      }
    }
    __syncthreads();
    // ---- 2. horizontal pass: region row `row`, output columns run*K .. run*K+K-1
    {
      T res[NH][K];
#pragma unroll
      for (int i = 0; i < NH; i++) {
        const int task = tid + i * 256;
        if (task < HT) {
          const int run = task / RH, row = task - run * RH;
          const T *b = S + row * P + run * K;
          vh_run_t<T, K, DIL, R>([&](int t) { return b[t]; }, res[i]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NH; i++) {
        const int task = tid + i * 256;
        if (task < HT) {
          const int run = task / RH, row = task - run * RH;
#pragma unroll
          for (int j = 0; j < K; j++) S[row * P2 + run * K + j] = res[i][j];
        }
      }
    }
    __syncthreads();
    // ---- 3. vertical pass: column `col`, output rows run*K .. run*K+K-1, stored
    const bool full = y0 + GEN_FT <= H && x0 + TWE <= W;
    for (int task = tid; task < VT; task += 256) {
      const int run = task / TWE, col = task - run * TWE;
      T o[K];
      const T *b = S + (run * K) * P2 + col;
      vh_run_t<T, K, DIL, R>([&](int t) { return b[t * P2]; }, o);
      const int gx = x0 + col, gy0 = y0 + run * K;
      T *d = dst + (int64_t)gy0 * W + gx;
      if (full) {
#pragma unroll
        for (int j = 0; j < K; j++) d[(int64_t)j * W] = o[j];
      } else if (gx < W) {
#pragma unroll
        for (int j = 0; j < K; j++)
          if (gy0 + j < H) d[(int64_t)j * W] = o[j];
      }
    }
    __syncthreads();  // the next tile's region overwrites S
  }
}

template <typename T, bool DIL> using gen_kernel_t = void (*)(const T *, T *, int, int, int, int, int, int);

// kernel of radius r (1 .. SQ_GEN_RMAX), instantiated per type in sq_gen_<type>.hip
template <typename T, bool DIL> gen_kernel_t<T, DIL> sq_gen_kernel_for(int r);
#define PM_GEN_DECL(T)                                                                                       \
  template <> gen_kernel_t<T, true> sq_gen_kernel_for<T, true>(int r);                                       \
  template <> gen_kernel_t<T, false> sq_gen_kernel_for<T, false>(int r);
PM_GEN_DECL(int32_t)
PM_GEN_DECL(int64_t)
PM_GEN_DECL(float)
PM_GEN_DECL(double)
#undef PM_GEN_DECL

// True if the fused generic kernel can run on these sizes.
inline bool sq_gen_applicable(int64_t H, int64_t W) { return H > 0 && W > 0 && H <= INT32_MAX / 2 && W <= INT32_MAX / 2; }
// (2r+1)^2 square on B images of H x W, any r >= 1 (radii above SQ_GEN_RMAX chain launches
// through a scratch image)
template <typename T, bool DIL> int sq_gen(const T *in, T *out, int64_t B, int64_t H, int64_t W, int64_t r, hipStream_t s);

}  // namespace pm
