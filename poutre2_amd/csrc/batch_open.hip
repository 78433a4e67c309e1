// batch_open.hip -- one-launch batched dilation + fused opening for radii 1..BATCH_FUSED_RMAX
// (batch_tile.hpp).
#include <utility>

#include "batch_tile.hpp"

namespace pm {

template <int... RS> static batch_fused_t pick(int r, std::integer_sequence<int, RS...>)
{
  static constexpr batch_fused_t tab[] = {batch_fused_kernel<1 + RS>...};
  return tab[r - 1];
}

batch_fused_t batch_fused_for(int r) { return pick(r, std::make_integer_sequence<int, BATCH_FUSED_RMAX>{}); }

}  // namespace pm
