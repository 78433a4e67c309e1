// sq_u16_keu.hip -- uint16 square erosion kernels for rows of any alignment, every radius
// 1..SQ_U16_RMAX (sq_u16.hpp), in a file of their own so that the instantiations compile in
// parallel.
#include <utility>

#include "sq_u16.hpp"

namespace pm {

template <int... RS> static u16_kernel_t pick(int r, std::integer_sequence<int, RS...>)
{
  static constexpr u16_kernel_t tab[] = {sq_u16_kernel<false, false, RS + 1>...};
  return tab[r - 1];
}

template <> u16_kernel_t sq_u16_kernel_for<false, false>(int r)
{
  return pick(r, std::make_integer_sequence<int, SQ_U16_RMAX>{});
}

}  // namespace pm
