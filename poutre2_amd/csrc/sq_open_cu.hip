// sq_open_cu.hip -- fused closing kernels, rows of any alignment, every radius
// 1..SQ_U8_OPEN_RMAX (sq_open.hpp).
#include <utility>

#include "sq_open.hpp"

namespace pm {

template <int... RS> static open_kernel_t pick(int r, std::integer_sequence<int, RS...>)
{
  static constexpr open_kernel_t tab[] = {sq_open_kernel<sq_k_open(RS + 1), false, false, RS + 1, OPEN_TW>...};
  return tab[r - 1];
}

template <> open_kernel_t sq_open_kernel_for<false, false>(int r)
{
  return pick(r, std::make_integer_sequence<int, (int)SQ_U8_OPEN_RMAX>{});
}

}  // namespace pm
