// vol_u8.hpp -- uint8 volume kernels streaming along z (vol_u8.hip).
#pragma once
#include "pm_common.hpp"

namespace pm {

constexpr int VOL_ZWIN_KMAX = 8;  // largest z radius of one zwin_u8 launch
// largest radius of the one-launch box (cube_u8).  512^3, us, one launch vs the 2D square
// then the z window: r = 1 85.7 vs 103, r = 2 105 vs 104, r = 3 130 vs 109 (the region load
// reads 2r + 1 planes)
constexpr int VOL_CUBE_RMAX = 1;

// True if the volume kernels can run on these buffers (W % 16 == 0, 16-byte aligned).
bool vol_u8_applicable(const void *in, const void *out, int64_t D, int64_t H, int64_t W);
// out(z) = max/min of in(z-k .. z+k), 1 <= k <= VOL_ZWIN_KMAX
template <bool DIL> int zwin_u8(const uint8_t *in, uint8_t *out, int64_t D, int64_t H, int64_t W, int k, hipStream_t s);
// (2r+1)^3 box in one launch, 1 <= r <= VOL_CUBE_RMAX
template <bool DIL> int cube_u8(const uint8_t *in, uint8_t *out, int64_t D, int64_t H, int64_t W, int r, hipStream_t s);
// one pass of the 6-connected cross
template <bool DIL> int cross3_u8(const uint8_t *in, uint8_t *out, int64_t D, int64_t H, int64_t W, hipStream_t s);

}  // namespace pm
