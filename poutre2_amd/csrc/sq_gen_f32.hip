// sq_gen_f32.hip -- fused square kernels for float images, every radius 1..SQ_GEN_RMAX
// (sq_gen.hpp).
#include <utility>

#include "sq_gen.hpp"

namespace pm {

template <bool DIL, int... RS> static gen_kernel_t<float, DIL> pick(int r, std::integer_sequence<int, RS...>)
{
  static constexpr gen_kernel_t<float, DIL> tab[] = {sq_gen_kernel<float, DIL, RS + 1>...};
  return tab[r - 1];
}

template <> gen_kernel_t<float, true> sq_gen_kernel_for<float, true>(int r)
{
  return pick<true>(r, std::make_integer_sequence<int, SQ_GEN_RMAX>{});
}
template <> gen_kernel_t<float, false> sq_gen_kernel_for<float, false>(int r)
{
  return pick<false>(r, std::make_integer_sequence<int, SQ_GEN_RMAX>{});
}

}  // namespace pm
