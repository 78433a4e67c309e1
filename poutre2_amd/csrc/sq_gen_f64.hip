// sq_gen_f64.hip -- fused square kernels for double images, every radius 1..SQ_GEN_RMAX
// (sq_gen.hpp).
#include <utility>

#include "sq_gen.hpp"

namespace pm {

template <bool DIL, int... RS> static gen_kernel_t<double, DIL> pick(int r, std::integer_sequence<int, RS...>)
{
  static constexpr gen_kernel_t<double, DIL> tab[] = {sq_gen_kernel<double, DIL, RS + 1>...};
  return tab[r - 1];
}

template <> gen_kernel_t<double, true> sq_gen_kernel_for<double, true>(int r)
{
  return pick<true>(r, std::make_integer_sequence<int, SQ_GEN_RMAX>{});
}
template <> gen_kernel_t<double, false> sq_gen_kernel_for<double, false>(int r)
{
  return pick<false>(r, std::make_integer_sequence<int, SQ_GEN_RMAX>{});
}

}  // namespace pm
