// vol_u8.hip -- uint8 volume kernels that stream along z (D x H x W, rows of W % 16 == 0
// bytes, 16-byte aligned buffers).
//
//  * zwin_u8_kernel<DIL,K>: out(z) = max/min of in(z-K .. z+K) (planes outside the volume
//    are skipped): the z pass of SESquare3D^K (a (2K+1)^3 box = 2D square per plane, then
//    this) and SESegmentZ3D^K (ero_dil_static_se_t.hpp:214-265, 761-839).
//  * cross3_u8_kernel<DIL>: one pass of the 6-connected cross SECross3D (7 taps,
//    ero_dil_static_se_t.hpp:841-915); SECross3D^k is k passes.
// A thread owns one 16-byte chunk of a plane (4 words, consecutive lanes on consecutive
// chunks, so every load and store is a coalesced 128-bit access) and walks VOL_ZB planes.
// The cross keeps the chunk's planes z-1, z, z+1 in registers; the y +- 1 chunks and the two
// dwords left and right of the chunk come through L1/L2 (they are other lanes' own chunks).
// Bytes are combined in the high-byte pair form of sq_u8_tile.hpp: the word itself carries
// bytes 1 and 3 in the high bytes of its u16 halves, the word shifted left by 8 carries
// bytes 0 and 2, and one v_pk_max_u16 / v_pk_min_u16 handles two pixels.
#include <utility>

#include "pm_common.hpp"
#include "sq_u8_tile.hpp"
#include "vol_u8.hpp"

namespace pm {

// max / min accumulator over words, bytewise
template <bool DIL> struct Acc4 {
  uint32_t o, e;
  __device__ __forceinline__ Acc4() : o(DIL ? 0u : 0xFFFFFFFFu), e(DIL ? 0u : 0xFFFFFFFFu) {}
  __device__ __forceinline__ void add(uint32_t w)
  {
    o = pk<DIL>(o, odd_hi(w));
    e = pk<DIL>(e, even_hi(w));
  }
  // bytes 0 and 2 from e's high bytes, 1 and 3 from o's
  __device__ __forceinline__ uint32_t get() const { return __builtin_amdgcn_perm(o, e, 0x07030501u); }
};

__device__ __forceinline__ uint32_t word(const uint4 &v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }

// planes per thread: 2 measured best on 512^3 (cross 46.9 us vs 54.5 / 53.6 / 69.9 us for
// 4 / 8 / 16 planes; z window r = 1: 38.4 vs 44.8 / 47.8 / 61.6 us): short z runs keep many
// more rows in flight at once than long ones
constexpr int VOL_ZB = 2;

template <bool DIL, int K>
__global__ __launch_bounds__(256) void zwin_u8_kernel(const uint4 *__restrict__ in, uint4 *__restrict__ out, int D,
                                                      int cpp)
{
  const int c = blockIdx.x * 256 + threadIdx.x;  // chunk index within a plane
  if (c >= cpp) return;
  const int z0 = blockIdx.y * VOL_ZB;
  const int z1 = min(D, z0 + VOL_ZB);
  for (int z = z0; z < z1; z++) {
    uint4 v[2 * K + 1];
#pragma unroll
    for (int t = 0; t <= 2 * K; t++) {
      // planes outside the volume read plane z (valid) and are replaced by the neutral value
      const int zz = z - K + t;
      const bool ok = (unsigned)zz < (unsigned)D;
      const uint4 x = in[(int64_t)(ok ? zz : z) * cpp + c];
      v[t] = ok ? x : (DIL ? make_uint4(0u, 0u, 0u, 0u) : make_uint4(~0u, ~0u, ~0u, ~0u));
    }
    Acc4<DIL> a[4];
#pragma unroll
    for (int t = 0; t <= 2 * K; t++) {
      a[0].add(v[t].x);
      a[1].add(v[t].y);
      a[2].add(v[t].z);
      a[3].add(v[t].w);
    }
    out[(int64_t)z * cpp + c] = make_uint4(a[0].get(), a[1].get(), a[2].get(), a[3].get());
  }
}

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

template <bool DIL>
__global__ __launch_bounds__(256) void cross3_u8_kernel(const uint4 *__restrict__ in4, uint4 *__restrict__ out, int D,
                                                        int H, int W16)
{
  const int cpp = H * W16;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cpp) return;
  const int y = c / W16, cx = c - y * W16;
  const int z0 = blockIdx.y * VOL_ZB;
  const int z1 = min(D, z0 + VOL_ZB);
  const uint32_t n32 = DIL ? 0u : ~0u;
  const u32x4_t NEU = {n32, n32, n32, n32};
  const u32x4_t *in = reinterpret_cast<const u32x4_t *>(in4);
  const uint32_t *in32 = reinterpret_cast<const uint32_t *>(in4);
  // Every load reads a valid address (a neighbour outside the volume reads the chunk itself)
  // and the neutral value is selected afterwards, so all loads are unconditional 128-bit
  // (or 32-bit) loads.
  const int64_t b0 = (int64_t)z0 * cpp + c;
  u32x4_t zm = in[z0 > 0 ? b0 - cpp : b0];
  if (z0 == 0) zm = NEU;
  u32x4_t cur = in[b0];
  const bool has_ym = y > 0, has_yp = y + 1 < H, has_l = cx > 0, has_r = cx + 1 < W16;
  const int64_t dym = has_ym ? -W16 : 0, dyp = has_yp ? W16 : 0;
  for (int z = z0; z < z1; z++) {
    const int64_t base = (int64_t)z * cpp + c;
    const bool has_zp = z + 1 < D;
    u32x4_t zp = in[base + (has_zp ? cpp : 0)];
    u32x4_t ym = in[base + dym];
    u32x4_t yp = in[base + dyp];
    uint32_t lw = in32[base * 4 + (has_l ? -1 : 0)];  // the dword left of the chunk
    uint32_t rw = in32[base * 4 + (has_r ? 4 : 0)];   // the dword right of it
    if (!has_zp) zp = NEU;
    if (!has_ym) ym = NEU;
    if (!has_yp) yp = NEU;
    if (!has_l) lw = n32;
    if (!has_r) rw = n32;
    uint32_t res[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint32_t w = cur[i];
      const uint32_t wl = i == 0 ? lw : cur[i - 1];
      const uint32_t wr = i == 3 ? rw : cur[i + 1];
      Acc4<DIL> a;
      a.add(w);
      a.add(zm[i]);
      a.add(zp[i]);
      a.add(ym[i]);
      a.add(yp[i]);
      a.add(__builtin_amdgcn_alignbyte(w, wl, 3));  // byte j <- byte j-1
      a.add(__builtin_amdgcn_alignbyte(wr, w, 1));  // byte j <- byte j+1
      res[i] = a.get();
    }
    out[base] = make_uint4(res[0], res[1], res[2], res[3]);
    zm = cur;
    cur = zp;
  }
}

bool vol_u8_applicable(const void *in, const void *out, int64_t D, int64_t H, int64_t W)
{
  return D > 0 && H > 0 && W > 0 && W % 16 == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0 &&
         H * (W / 16) <= INT32_MAX / 2 && D <= INT32_MAX / 2;
}

typedef void (*zwin_kernel_t)(const uint4 *, uint4 *, int, int);

template <bool DIL, int... KS> static zwin_kernel_t zwin_pick(int k, std::integer_sequence<int, KS...>)
{
  static constexpr zwin_kernel_t tab[] = {zwin_u8_kernel<DIL, KS + 1>...};
  return tab[k - 1];
}

template <bool DIL> int zwin_u8(const uint8_t *in, uint8_t *out, int64_t D, int64_t H, int64_t W, int k, hipStream_t s)
{
  if (k < 1 || k > VOL_ZWIN_KMAX) PM_FAIL(PM_E_ARG, "z window radius %d outside [1, %d]", k, VOL_ZWIN_KMAX);
  const int cpp = (int)(H * (W / 16));
  const dim3 grid((unsigned)((cpp + 255) / 256), (unsigned)((D + VOL_ZB - 1) / VOL_ZB));
  hipLaunchKernelGGL(zwin_pick<DIL>(k, std::make_integer_sequence<int, VOL_ZWIN_KMAX>{}), grid, dim3(256), 0, s,
                     (const uint4 *)in, (uint4 *)out, (int)D, cpp);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <bool DIL> int cross3_u8(const uint8_t *in, uint8_t *out, int64_t D, int64_t H, int64_t W, hipStream_t s)
{
  const int W16 = (int)(W / 16);
  const int64_t cpp = H * W16;
  const dim3 grid((unsigned)((cpp + 255) / 256), (unsigned)((D + VOL_ZB - 1) / VOL_ZB));
  hipLaunchKernelGGL((cross3_u8_kernel<DIL>), grid, dim3(256), 0, s, (const uint4 *)in, (uint4 *)out, (int)D, (int)H,
                     W16);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template int zwin_u8<true>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, int, hipStream_t);
template int zwin_u8<false>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, int, hipStream_t);
template int cross3_u8<true>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, hipStream_t);
template int cross3_u8<false>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, hipStream_t);

}  // namespace pm
