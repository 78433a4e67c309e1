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

}  // namespace pm
