// batch_r17.hip -- batched dilation + opening kernels for radii 17..24 (batch_tile.hpp).
#include <utility>

#include "batch_tile.hpp"

namespace pm {

template <int... RS> static batch_kernel_t pick(int r, int phase, std::integer_sequence<int, RS...>)
{
  static constexpr batch_kernel_t p0[] = {batch_sq_kernel<17 + RS, 0>...};
  static constexpr batch_kernel_t p1[] = {batch_sq_kernel<17 + RS, 1>...};
  return phase ? p1[r - 17] : p0[r - 17];
}

template <> batch_kernel_t batch_kernel_range<17>(int r, int phase)
{
  return pick(r, phase, std::make_integer_sequence<int, 8>{});
}

}  // namespace pm
