// sq_u8_kea.hip -- fused square kernels: erosion, 16-byte aligned rows,
// every radius 1..32 and tile width 128 / 192 / 256 (sq_u8_kernel.hpp).
#include <utility>

#include "sq_u8_kernel.hpp"

namespace pm {

template <int... RS> static sq_kernel_t pick(int r, int tw, std::integer_sequence<int, RS...>)
{
  static constexpr sq_kernel_t w128[] = {sq_u8_kernel<sq_k_of(RS + 1), false, true, RS + 1, 128>...};
  static constexpr sq_kernel_t w192[] = {sq_u8_kernel<sq_k_of(RS + 1), false, true, RS + 1, 192>...};
  static constexpr sq_kernel_t w256[] = {sq_u8_kernel<sq_k_of(RS + 1), false, true, RS + 1, 256>...};
  return tw == 256 ? w256[r - 1] : tw == 192 ? w192[r - 1] : w128[r - 1];
}

template <> sq_kernel_t sq_kernel_for<false, true>(int r, int tw)
{
  return pick(r, tw, std::make_integer_sequence<int, 32>{});
}

}  // namespace pm
