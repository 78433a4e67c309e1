// sq_u8_rs.hip -- the 3 x 3 square (Dilate/Erode(img, SESquare2D, 1), ero_dil_static.cpp:229-305)
// on uint8 images with 16-byte aligned rows, streamed through registers: no LDS, no barrier.
//
// A wave owns 64 consecutive 16-byte chunks of a band of SQ_RS_RUN rows; lane l owns chunk
// 64 * wc + l and walks down the band.  It keeps its last three rows in registers in the
// high-byte-pair form of sq_u8_tile.hpp (E = the word shifted left by 8: bytes 0 and 2 in the
// high bytes of its u16 halves; O = the word: bytes 1 and 3), so a row step is: one 16-byte
// load of the new row, the vertical max over the three rows (v_pk_max_u16 on E and O), the
// words beside the chunk from the neighbouring lanes through DPP (wave_shr:1 / wave_shl:1),
// the byte shifts by one (one v_perm_b32 each in this form: the left-shifted word's E is
// (left word's byte 3, byte 1) and its O is E; the right-shifted word's E is O and its O is
// (byte 2, right word's byte 0)), the horizontal max, and one 16-byte store.  Lanes 0 and 63
// also track the word just outside the wave (left of chunk 0, right of chunk 63), which DPP
// leaves in place where a lane has no source.  Every load and store of a wave is one
// contiguous 1 KiB row segment.  8192x8192, us: 31.1-31.8 (LDS tile kernel) -> 28.5-29.1;
// bands of 8 rows (8192 waves) beat 16 (29.5) and 32 (38.4); for r = 2 and 3 this scheme is
// no faster than the tile kernel (30.5 / 32.6 vs 31.0 / 31.2 us).
#include <type_traits>

#include "pm_common.hpp"
#include "sq_u8.hpp"
#include "sq_u8_tile.hpp"

namespace pm {

constexpr int SQ_RS_RUN = 8;  // output rows per wave

__device__ __forceinline__ uint32_t rs_sh(uint32_t hi, uint32_t lo) { return __builtin_amdgcn_perm(hi, lo, 0x05050303u); }

template <bool DIL>
__global__ __launch_bounds__(256) void sq_u8_rs_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, int H,
                                                       int W, int nwc, int nruns, int nwaves)
{
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= nwaves) return;  // whole waves
  // wave -> (image, band, column group), column groups fastest: the waves of one band run
  // side by side and meet in the same L2
  const int per_img = nruns * nwc;
  const int img = wave / per_img;
  const int rem = wave - img * per_img;
  const int run = rem / nwc, wc = rem - run * nwc;
  const uint8_t *__restrict__ src = in + (int64_t)img * H * W;
  uint8_t *__restrict__ dst = out + (int64_t)img * H * W;
  const int cpr = W >> 4;  // chunks per row
  const int c = wc * 64 + lane;
  const bool cin = c < cpr;
  const int xc = lane == 0 ? c - 1 : c + 1;  // lanes 0 / 63: the chunk outside the wave
  const bool xin = (lane == 0 || lane == 63) && (unsigned)xc < (unsigned)cpr;
  const int xoff = lane == 0 ? 12 : 0;       // its last word (left) or first word (right)
  const int y0 = run * SQ_RS_RUN;
  const uint32_t nv = DIL ? 0u : 0xFFFFFFFFu;
  uint32_t re[3][4], ro[3][4], xe[3], xo[3];
  auto load = [&](int yi, uint32_t(&e)[4], uint32_t(&o)[4], uint32_t &ex, uint32_t &ox) {
    uint4 v = make_uint4(nv, nv, nv, nv);
    uint32_t xw = nv;
    if ((unsigned)yi < (unsigned)H) {
      const uint8_t *row = src + (int64_t)yi * W;
      if (cin) v = *reinterpret_cast<const uint4 *>(row + 16 * c);
      if (xin) xw = *reinterpret_cast<const uint32_t *>(row + 16 * xc + xoff);
    }
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      e[i] = even_hi(w[i]);
      o[i] = odd_hi(w[i]);
    }
    ex = even_hi(xw);
    ox = odd_hi(xw);
  };
  load(y0 - 1, re[0], ro[0], xe[0], xo[0]);
  load(y0, re[1], ro[1], xe[1], xo[1]);
  const int nsteps = min(SQ_RS_RUN, H - y0);
  // output row y0 + j needs rows y0 + j - 1 .. y0 + j + 1; row y0 + j + 1 goes to slot
  // (j + 2) % 3, a compile-time index since the loop is unrolled by 3
  for (int base = 0; base < nsteps; base += 3) {
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const int j = base + u;
      if (j >= nsteps) break;
      const int slot = (u + 2) % 3;
      load(y0 + j + 1, re[slot], ro[slot], xe[slot], xo[slot]);
      uint32_t ve[4], vo[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        ve[i] = pk<DIL>(pk<DIL>(re[0][i], re[1][i]), re[2][i]);
        vo[i] = pk<DIL>(pk<DIL>(ro[0][i], ro[1][i]), ro[2][i]);
      }
      const uint32_t vxe = pk<DIL>(pk<DIL>(xe[0], xe[1]), xe[2]), vxo = pk<DIL>(pk<DIL>(xo[0], xo[1]), xo[2]);
      // the word left of word 0 (from lane l-1; its O is all the shift by one needs) and right of
      // word 3 (from lane l+1; its E); lanes 0 and 63 have no DPP source and keep the tracked word
      const uint32_t lo = __builtin_amdgcn_update_dpp(vxo, vo[3], 0x138, 0xf, 0xf, false);
      const uint32_t rie = __builtin_amdgcn_update_dpp(vxe, ve[0], 0x130, 0xf, 0xf, false);
      uint32_t res[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t ec = ve[i], oc = vo[i];
        const uint32_t ol = i == 0 ? lo : vo[i - 1];
        const uint32_t er = i == 3 ? rie : ve[i + 1];
        const uint32_t me = pk<DIL>(pk<DIL>(ec, rs_sh(oc, ol)), oc);  // centre, left, right
        const uint32_t mo = pk<DIL>(pk<DIL>(oc, ec), rs_sh(er, ec));
        res[i] = __builtin_amdgcn_perm(mo, me, 0x07030501u);
      }
      if (cin) *reinterpret_cast<uint4 *>(dst + (int64_t)(y0 + j) * W + 16 * c) = make_uint4(res[0], res[1], res[2], res[3]);
    }
  }
}

// ---------------------------------------------------------------- opening, closing, gradient
// The same streaming for the compositions of the 3 x 3 square.

// 3-tap horizontal op of a row of four words in (E, O) form; lo = O of the word left of word
// 0, re = E of the word right of word 3
template <bool DIL>
__device__ __forceinline__ void rs_h3(const uint32_t (&ve)[4], const uint32_t (&vo)[4], uint32_t lo, uint32_t re,
                                      uint32_t (&he)[4], uint32_t (&ho)[4])
{
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t ol = i == 0 ? lo : vo[i - 1], er = i == 3 ? re : ve[i + 1];
    he[i] = pk<DIL>(pk<DIL>(ve[i], rs_sh(vo[i], ol)), vo[i]);
    ho[i] = pk<DIL>(pk<DIL>(vo[i], ve[i]), rs_sh(er, ve[i]));
  }
}
// the same for one word (the word a lane 0 / 63 tracks outside its wave)
template <bool DIL> __device__ __forceinline__ void rs_h3w(uint32_t e, uint32_t o, uint32_t lo, uint32_t re, uint32_t &he, uint32_t &ho)
{
  he = pk<DIL>(pk<DIL>(e, rs_sh(o, lo)), o);
  ho = pk<DIL>(pk<DIL>(o, e), rs_sh(re, e));
}

// A row of the wave in (E, O) form: the lane's four words and the word it tracks outside the
// wave (lanes 0 / 63).
struct RsRow {
  uint32_t e[4], o[4], xe, xo;
};

// Opening (OPEN: the 3 x 3 erosion, then the 3 x 3 dilation) or closing, streamed: a ring of
// three input rows and a ring of three first-operator rows.  The first operator's rows and
// words outside the image hold the second operator's neutral value (the reference pads each
// operator of the composition with its own).
template <bool OPEN>
__global__ __launch_bounds__(256) void sq_u8_rs2_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, int H,
                                                        int W, int nwc, int nruns, int nwaves)
{
  constexpr bool D1 = !OPEN, D2 = OPEN;  // an opening erodes first
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= nwaves) return;
  const int per_img = nruns * nwc;
  const int img = wave / per_img;
  const int rem = wave - img * per_img;
  const int run = rem / nwc, wc = rem - run * nwc;
  const uint8_t *__restrict__ src = in + (int64_t)img * H * W;
  uint8_t *__restrict__ dst = out + (int64_t)img * H * W;
  const int cpr = W >> 4;
  const int c = wc * 64 + lane;
  const bool cin = c < cpr;
  const int xc = lane == 0 ? c - 1 : c + 1;
  const bool xin = (lane == 0 || lane == 63) && (unsigned)xc < (unsigned)cpr;
  const int xoff = lane == 0 ? 12 : 0;
  const int y0 = run * SQ_RS_RUN;
  const uint32_t n1 = D1 ? 0u : 0xFFFFFFFFu, n2 = D2 ? 0u : 0xFFFFFFFFu;
  RsRow I[3], S[3];
  auto load = [&](int yi, RsRow &r) {
    uint4 v = make_uint4(n1, n1, n1, n1);
    uint32_t xw = n1;
    if ((unsigned)yi < (unsigned)H) {
      const uint8_t *row = src + (int64_t)yi * W;
      if (cin) v = *reinterpret_cast<const uint4 *>(row + 16 * c);
      if (xin) xw = *reinterpret_cast<const uint32_t *>(row + 16 * xc + xoff);
    }
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      r.e[i] = even_hi(w[i]);
      r.o[i] = odd_hi(w[i]);
    }
    r.xe = even_hi(xw);
    r.xo = odd_hi(xw);
  };
  // first operator of row ys from the input ring (slots a, b, c in any order)
  auto first = [&](int ys, const RsRow &a, const RsRow &b, const RsRow &cc, RsRow &r) {
    uint32_t ve[4], vo[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ve[i] = pk<D1>(pk<D1>(a.e[i], b.e[i]), cc.e[i]);
      vo[i] = pk<D1>(pk<D1>(a.o[i], b.o[i]), cc.o[i]);
    }
    const uint32_t vxe = pk<D1>(pk<D1>(a.xe, b.xe), cc.xe), vxo = pk<D1>(pk<D1>(a.xo, b.xo), cc.xo);
    const uint32_t lo = __builtin_amdgcn_update_dpp(vxo, vo[3], 0x138, 0xf, 0xf, false);
    const uint32_t re = __builtin_amdgcn_update_dpp(vxe, ve[0], 0x130, 0xf, 0xf, false);
    rs_h3<D1>(ve, vo, lo, re, r.e, r.o);
    // the tracked word: lane 0's (left of the wave) needs its right neighbour, word 0, for its
    // last byte; lane 63's (right of the wave) its left neighbour, word 3, for its first byte
    rs_h3w<D1>(vxe, vxo, vo[3], ve[0], r.xe, r.xo);
    const bool rin = (unsigned)ys < (unsigned)H;
    const uint32_t ne = even_hi(n2);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      r.e[i] = rin && cin ? r.e[i] : ne;
      r.o[i] = rin && cin ? r.o[i] : n2;
    }
    r.xe = rin && xin ? r.xe : ne;
    r.xo = rin && xin ? r.xo : n2;
  };
  load(y0 - 2, I[0]);
  load(y0 - 1, I[1]);
  load(y0, I[2]);
  first(y0 - 1, I[0], I[1], I[2], S[0]);
  load(y0 + 1, I[0]);
  first(y0, I[1], I[2], I[0], S[1]);
  const int nsteps = min(SQ_RS_RUN, H - y0);
  // step j: input row y0 + j + 2 -> input slot (j + 1) % 3, first-operator row y0 + j + 1 ->
  // slot (j + 2) % 3, output row y0 + j from the three first-operator rows
  for (int base = 0; base < nsteps; base += 3) {
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const int j = base + u;
      if (j >= nsteps) break;
      load(y0 + j + 2, I[(u + 1) % 3]);
      first(y0 + j + 1, I[0], I[1], I[2], S[(u + 2) % 3]);
      uint32_t ve[4], vo[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        ve[i] = pk<D2>(pk<D2>(S[0].e[i], S[1].e[i]), S[2].e[i]);
        vo[i] = pk<D2>(pk<D2>(S[0].o[i], S[1].o[i]), S[2].o[i]);
      }
      const uint32_t vxe = pk<D2>(pk<D2>(S[0].xe, S[1].xe), S[2].xe), vxo = pk<D2>(pk<D2>(S[0].xo, S[1].xo), S[2].xo);
      const uint32_t lo = __builtin_amdgcn_update_dpp(vxo, vo[3], 0x138, 0xf, 0xf, false);
      const uint32_t re = __builtin_amdgcn_update_dpp(vxe, ve[0], 0x130, 0xf, 0xf, false);
      uint32_t he[4], ho[4];
      rs_h3<D2>(ve, vo, lo, re, he, ho);
      if (cin)
        *reinterpret_cast<uint4 *>(dst + (int64_t)(y0 + j) * W + 16 * c) =
            make_uint4(__builtin_amdgcn_perm(ho[0], he[0], 0x07030501u), __builtin_amdgcn_perm(ho[1], he[1], 0x07030501u),
                       __builtin_amdgcn_perm(ho[2], he[2], 0x07030501u), __builtin_amdgcn_perm(ho[3], he[3], 0x07030501u));
    }
  }
}

// Gradient (3 x 3 dilation minus 3 x 3 erosion) streamed: one load per row, a ring of three
// rows for the max (neutral 0 outside the image) and one for the min (neutral 255); max >= min
// bytewise, so the word subtraction never borrows.
__global__ __launch_bounds__(256) void sq_u8_rsg_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, int H,
                                                        int W, int nwc, int nruns, int nwaves)
{
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= nwaves) return;
  const int per_img = nruns * nwc;
  const int img = wave / per_img;
  const int rem = wave - img * per_img;
  const int run = rem / nwc, wc = rem - run * nwc;
  const uint8_t *__restrict__ src = in + (int64_t)img * H * W;
  uint8_t *__restrict__ dst = out + (int64_t)img * H * W;
  const int cpr = W >> 4;
  const int c = wc * 64 + lane;
  const bool cin = c < cpr;
  const int xc = lane == 0 ? c - 1 : c + 1;
  const bool xin = (lane == 0 || lane == 63) && (unsigned)xc < (unsigned)cpr;
  const int xoff = lane == 0 ? 12 : 0;
  const int y0 = run * SQ_RS_RUN;
  RsRow A[3], B[3];  // max path (outside: 0), min path (outside: 255)
  auto load = [&](int yi, RsRow &a, RsRow &b) {
    const bool rin = (unsigned)yi < (unsigned)H;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    uint32_t xw = 0u;
    if (rin) {
      const uint8_t *row = src + (int64_t)yi * W;
      if (cin) v = *reinterpret_cast<const uint4 *>(row + 16 * c);
      if (xin) xw = *reinterpret_cast<const uint32_t *>(row + 16 * xc + xoff);
    }
    const bool vin = rin && cin, xv = rin && xin;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      a.e[i] = even_hi(w[i]);
      a.o[i] = odd_hi(w[i]);
      b.e[i] = vin ? a.e[i] : 0xFFFFFF00u;
      b.o[i] = vin ? a.o[i] : 0xFFFFFFFFu;
    }
    a.xe = even_hi(xw);
    a.xo = odd_hi(xw);
    b.xe = xv ? a.xe : 0xFFFFFF00u;
    b.xo = xv ? a.xo : 0xFFFFFFFFu;
  };
  auto hop = [&](auto dil, const RsRow (&R)[3], uint32_t (&he)[4], uint32_t (&ho)[4]) {
    constexpr bool D = decltype(dil)::value;
    uint32_t ve[4], vo[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ve[i] = pk<D>(pk<D>(R[0].e[i], R[1].e[i]), R[2].e[i]);
      vo[i] = pk<D>(pk<D>(R[0].o[i], R[1].o[i]), R[2].o[i]);
    }
    const uint32_t vxe = pk<D>(pk<D>(R[0].xe, R[1].xe), R[2].xe), vxo = pk<D>(pk<D>(R[0].xo, R[1].xo), R[2].xo);
    const uint32_t lo = __builtin_amdgcn_update_dpp(vxo, vo[3], 0x138, 0xf, 0xf, false);
    const uint32_t re = __builtin_amdgcn_update_dpp(vxe, ve[0], 0x130, 0xf, 0xf, false);
    rs_h3<D>(ve, vo, lo, re, he, ho);
  };
  load(y0 - 1, A[0], B[0]);
  load(y0, A[1], B[1]);
  const int nsteps = min(SQ_RS_RUN, H - y0);
  for (int base = 0; base < nsteps; base += 3) {
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const int j = base + u;
      if (j >= nsteps) break;
      load(y0 + j + 1, A[(u + 2) % 3], B[(u + 2) % 3]);
      uint32_t xe[4], xo[4], ne[4], no[4];
      hop(std::integral_constant<bool, true>{}, A, xe, xo);
      hop(std::integral_constant<bool, false>{}, B, ne, no);
      uint32_t res[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
        res[i] = __builtin_amdgcn_perm(xo[i], xe[i], 0x07030501u) - __builtin_amdgcn_perm(no[i], ne[i], 0x07030501u);
      if (cin) *reinterpret_cast<uint4 *>(dst + (int64_t)(y0 + j) * W + 16 * c) = make_uint4(res[0], res[1], res[2], res[3]);
    }
  }
}

bool sq_u8_rs_applicable(const void *in, const void *out, int64_t H, int64_t W)
{
  return H > 0 && W > 0 && W % 16 == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0 && H <= INT32_MAX / 2 &&
         W <= INT32_MAX / 2;
}

template <bool DIL> int sq_u8_rs(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, hipStream_t s)
{
  if (B * H * W == 0) return PM_OK;
  const int64_t nwc = (W / 16 + 63) / 64, nruns = (H + SQ_RS_RUN - 1) / SQ_RS_RUN;
  const int64_t nwaves = B * nruns * nwc;
  if (nwaves > INT32_MAX - 3) PM_FAIL(PM_E_ARG, "image too large for the streamed 3x3 kernel");
  hipLaunchKernelGGL((sq_u8_rs_kernel<DIL>), dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, s, in, out, (int)H, (int)W,
                     (int)nwc, (int)nruns, (int)nwaves);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template int sq_u8_rs<true>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, hipStream_t);
template int sq_u8_rs<false>(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t, hipStream_t);

// op: 0 opening, 1 closing, 2 gradient
int sq_u8_rs_compose(int op, const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, hipStream_t s)
{
  if (B * H * W == 0) return PM_OK;
  const int64_t nwc = (W / 16 + 63) / 64, nruns = (H + SQ_RS_RUN - 1) / SQ_RS_RUN;
  const int64_t nwaves = B * nruns * nwc;
  if (nwaves > INT32_MAX - 3) PM_FAIL(PM_E_ARG, "image too large for the streamed 3x3 kernels");
  const dim3 grid((unsigned)((nwaves + 3) / 4));
  if (op == 0)
    hipLaunchKernelGGL((sq_u8_rs2_kernel<true>), grid, dim3(256), 0, s, in, out, (int)H, (int)W, (int)nwc, (int)nruns,
                       (int)nwaves);
  else if (op == 1)
    hipLaunchKernelGGL((sq_u8_rs2_kernel<false>), grid, dim3(256), 0, s, in, out, (int)H, (int)W, (int)nwc, (int)nruns,
                       (int)nwaves);
  else
    hipLaunchKernelGGL(sq_u8_rsg_kernel, grid, dim3(256), 0, s, in, out, (int)H, (int)W, (int)nwc, (int)nruns, (int)nwaves);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

}  // namespace pm
