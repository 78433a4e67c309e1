// vol_u8.hip -- uint8 volume kernels that stream along z (D x H x W, rows of W % 16 == 0
// bytes, 16-byte aligned buffers).
//
//  * zwin_u8_kernel<DIL,K>: out(z) = max/min of in(z-K .. z+K) (planes outside the volume
//    are skipped): the z pass of SESquare3D^K (a (2K+1)^3 box = 2D square per plane, then
//    this) and SESegmentZ3D^K (ero_dil_static_se_t.hpp:214-265, 761-839).
//  * cross3_u8_kernel<DIL>: one pass of the 6-connected cross SECross3D (7 taps,
//    ero_dil_static_se_t.hpp:841-915); SECross3D^k is k passes.
//  * cube_u8_kernel<DIL,R>: the (2R+1)^3 box (SESquare3D^R) in one launch, R <= VOL_CUBE_RMAX:
//    the fused 2D square tile of sq_u8_tile.hpp whose region is loaded as the max / min over
//    planes z-R .. z+R (the box is separable, and the z window commutes with the 2D square).
// A thread owns one 16-byte chunk of a plane (4 words, consecutive lanes on consecutive
// chunks, so every load and store is a coalesced 128-bit access) and walks VOL_ZB planes.
// The cross keeps the chunk's planes z-1, z, z+1 in registers; the y +- 1 chunks and the two
// dwords left and right of the chunk come through L1/L2 (they are other lanes' own chunks).
// Bytes are combined in the high-byte pair form of sq_u8_tile.hpp: the word itself carries
// bytes 1 and 3 in the high bytes of its u16 halves, the word shifted left by 8 carries
// bytes 0 and 2, and one v_pk_max_u16 / v_pk_min_u16 handles two pixels.
#include <utility>

#include "pm_common.hpp"
#include "sq_u8_kernel.hpp"
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

// bytewise max / min of two words (high-byte u16 pairs, see sq_u8_tile.hpp)
template <bool DIL> __device__ __forceinline__ uint32_t bytes_op(uint32_t a, uint32_t b)
{
  return __builtin_amdgcn_perm(pk<DIL>(odd_hi(a), odd_hi(b)), pk<DIL>(even_hi(a), even_hi(b)), 0x07030501u);
}
template <bool DIL> __device__ __forceinline__ uint4 bytes_op(const uint4 &a, const uint4 &b)
{
  return make_uint4(bytes_op<DIL>(a.x, b.x), bytes_op<DIL>(a.y, b.y), bytes_op<DIL>(a.z, b.z), bytes_op<DIL>(a.w, b.w));
}

// Region of the 2D tile at (y0, x0) of plane z, as sq_tile_load (aligned rows, R16 = 16),
// each chunk the max / min over planes z-R .. z+R that lie in the volume: plane z's chunks
// first, then the other planes', all of a plane's loads in flight at once.
template <bool DIL, int R, int TW, int TH>
__device__ __forceinline__ void cube_region_load(uint8_t *smem, const uint8_t *__restrict__ vol, int D, int z, int H,
                                                 int W, int y0, int x0)
{
  constexpr int K = sq_k_of(R), R16 = sq_r16<K>(), AW = TW + 2 * R16, cpr = AW >> 4;
  constexpr int nchunk = (TH + 2 * R) * cpr;
  constexpr int MAXC = (nchunk + FTHREADS - 1) / FTHREADS;
  const uint32_t nv = DIL ? 0u : 0xFFFFFFFFu;
  const int gx0 = x0 - R16;
  const int64_t plane = (int64_t)H * W;
  const int tid = threadIdx.x;
  int64_t off[MAXC];  // byte offset of the chunk within a plane, -1 outside the image
#pragma unroll
  for (int i = 0; i < MAXC; i++) {
    const int c = tid + i * FTHREADS;
    const int ri = c / cpr, cc = c - ri * cpr;
    const int gy = y0 - R + ri, gx = gx0 + (cc << 4);
    off[i] = (c < nchunk && (unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) ? (int64_t)gy * W + gx : -1;
  }
  uint4 v[MAXC];
  const uint8_t *pz = vol + (int64_t)z * plane;
#pragma unroll
  for (int i = 0; i < MAXC; i++)
    v[i] = off[i] >= 0 ? *reinterpret_cast<const uint4 *>(pz + off[i]) : make_uint4(nv, nv, nv, nv);
#pragma unroll
  for (int dz = -R; dz <= R; dz++) {
    if (dz == 0 || (unsigned)(z + dz) >= (unsigned)D) continue;  // planes outside the volume: neutral
    const uint8_t *pq = vol + (int64_t)(z + dz) * plane;
    uint4 u[MAXC];
#pragma unroll
    for (int i = 0; i < MAXC; i++)
      u[i] = off[i] >= 0 ? *reinterpret_cast<const uint4 *>(pq + off[i]) : make_uint4(nv, nv, nv, nv);
#pragma unroll
    for (int i = 0; i < MAXC; i++) v[i] = bytes_op<DIL>(v[i], u[i]);
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

// Persistent over the (plane, tile) list with the XCD-aware order of sq_u8_kernel: an XCD
// walks a contiguous range of planes, so the z-neighbour planes of its tiles are read
// through its L2.  Rows of W % 16 == 0 bytes, 16-byte aligned buffers.
template <bool DIL, int R>
__global__ __launch_bounds__(FTHREADS) void cube_u8_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                           int D, int H, int W, int tiles_x, int tiles_per_img,
                                                           int ntiles)
{
  extern __shared__ __align__(16) uint8_t smem[];
  constexpr int K = sq_k_of(R), TW = 256, TH = sq_th(R);
  const TileSched ts = tile_sched(blockIdx.x, gridDim.x, ntiles);
  for (int tile = ts.first; tile < ts.end; tile += ts.step) {
    const int z = tile / tiles_per_img;
    const int t_in = tile - z * tiles_per_img;
    const int ty = t_in / tiles_x, tx = t_in - ty * tiles_x;
    const int y0 = ty * TH, x0 = tx * TW;
    cube_region_load<DIL, R, TW, TH>(smem, in, D, z, H, W, y0, x0);
    __syncthreads();
    sq_tile_passes<K, DIL, true, R, TW, TH>(smem, out + (int64_t)z * H * W, H, W, R, y0, x0);
    __syncthreads();  // A/T are rewritten by the next tile
  }
}

typedef void (*cube_kernel_t)(const uint8_t *, uint8_t *, int, int, int, int, int, int);

template <bool DIL> int cube_u8(const uint8_t *in, uint8_t *out, int64_t D, int64_t H, int64_t W, int r, hipStream_t s)
{
  static_assert(VOL_CUBE_RMAX == 1, "one kernel per radius");
  if (r < 1 || r > VOL_CUBE_RMAX) PM_FAIL(PM_E_ARG, "cube radius %d outside [1, %d]", r, VOL_CUBE_RMAX);
  const cube_kernel_t k = cube_u8_kernel<DIL, 1>;
  constexpr int TW = 256;
  const int th = sq_th(r);
  const int tiles_x = (int)((W + TW - 1) / TW), tiles_y = (int)((H + th - 1) / th);
  const int64_t per = (int64_t)tiles_x * tiles_y, ntiles = per * D;
  if (ntiles > INT32_MAX) PM_FAIL(PM_E_ARG, "volume too large for the cube kernel");
  const size_t lds = sq_tile_lds_bytes(r, TW, th);
  int dev = 0, cus = 0, occ = 0;
  PM_HIP(hipGetDevice(&dev));
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, FTHREADS, lds));
  int64_t g = (int64_t)(occ > 0 ? occ : 1) * cus;
  if (g >= 8) g &= ~(int64_t)7;
  if (g > ntiles) g = ntiles;
  hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(FTHREADS), lds, s, in, out, (int)D, (int)H, (int)W, tiles_x, (int)per,
                     (int)ntiles);
  PM_HIP(hipGetLastError());
  return PM_OK;
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
template int cube_u8<true>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, int, hipStream_t);
template int cube_u8<false>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, int, hipStream_t);

}  // namespace pm
