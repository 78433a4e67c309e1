// sq_u8.hpp -- fused uint8 square erosion/dilation (sq_u8.hip).
#pragma once
#include "pm_common.hpp"

namespace pm {

constexpr int64_t SQ_U8_RMAX = 32;  // largest radius handled by one launch

// True if the fused kernel can run on these buffers (any width and alignment; images whose
// rows are 16-byte aligned take the 128-bit load/store variant).
bool sq_u8_applicable(const void *in, const void *out, int64_t H, int64_t W);
// (2r+1)^2 square on B images of H x W; r > SQ_U8_RMAX runs several launches.
template <bool DIL> int sq_u8(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int64_t r, hipStream_t s);

// The 3 x 3 square (r = 1) on B images whose rows are 16-byte aligned, streamed through
// registers (sq_u8_rs.hip); sq_u8 takes it for such images.
bool sq_u8_rs_applicable(const void *in, const void *out, int64_t H, int64_t W);
template <bool DIL> int sq_u8_rs(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, hipStream_t s);
// The 3 x 3 opening (op 0), closing (1) or gradient (2) the same way, on such images.
int sq_u8_rs_compose(int op, const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, hipStream_t s);

// Dilate - Erode by the (2r+1)^2 square on B images of H x W in one launch,
// 1 <= r <= SQ_U8_GRAD_RMAX
constexpr int64_t SQ_U8_GRAD_RMAX = 32;
int sq_u8_gradient(const uint8_t *in, uint8_t *out, int64_t B, int64_t H, int64_t W, int64_t r, hipStream_t s);

}  // namespace pm
