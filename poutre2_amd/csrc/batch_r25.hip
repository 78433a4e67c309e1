// batch_r25.hip -- batched dilation + opening kernels for radii 25..32 (batch_tile.hpp).
#include <utility>

#include "batch_tile.hpp"

namespace pm {

template <int... RS> static batch_kernel_t pick(int r, int phase, std::integer_sequence<int, RS...>)
{
  static constexpr batch_kernel_t p0[] = {batch_sq_kernel<25 + RS, 0>...};
  static constexpr batch_kernel_t p1[] = {batch_sq_kernel<25 + RS, 1>...};
  return phase ? p1[r - 25] : p0[r - 25];
}

template <> batch_kernel_t batch_kernel_range<25>(int r, int phase)
{
  return pick(r, phase, std::make_integer_sequence<int, 8>{});
}

}  // namespace pm
