// sq_gen_i64.hip -- fused square kernels for int64_t images, every radius 1..SQ_GEN_RMAX
// (sq_gen.hpp).
#include <utility>

#include "sq_gen.hpp"

namespace pm {

template <bool DIL, int... RS> static gen_kernel_t<int64_t, DIL> pick(int r, std::integer_sequence<int, RS...>)
{
  static constexpr gen_kernel_t<int64_t, DIL> tab[] = {sq_gen_kernel<int64_t, DIL, RS + 1>...};
  return tab[r - 1];
}

template <> gen_kernel_t<int64_t, true> sq_gen_kernel_for<int64_t, true>(int r)
{
  return pick<true>(r, std::make_integer_sequence<int, SQ_GEN_RMAX>{});
}
template <> gen_kernel_t<int64_t, false> sq_gen_kernel_for<int64_t, false>(int r)
{
  return pick<false>(r, std::make_integer_sequence<int, SQ_GEN_RMAX>{});
}

}  // namespace pm
