// erodil.hpp -- internal interface of the flat erosion / dilation drivers (erodil.hip),
// used by the geodesy and batch translation units.
#pragma once
#include "pm_common.hpp"

namespace pm {

constexpr int NL_MAX = 64;
// neighbour list passed by value as a kernel argument ({dz,dy,dx} per tap)
struct NLArgs {
  int n;
  int dz[NL_MAX], dy[NL_MAX], dx[NL_MAX];
};

int se_rank(int se);
int se_nl(int se, NLArgs &a);
int check_se_rank(int se, int rank);

template <typename T, bool DIL> int static_iter(const T *in, T *out, Dims d, int rank, int se, int iter, hipStream_t s);
template <typename T, bool DIL>
int rect_pass(const T *in, T *out, int64_t B, int64_t H, int64_t W, int64_t ry, int64_t rx, hipStream_t s);
template <typename T, bool DIL> int nl_pass(const T *in, T *out, Dims d, const NLArgs &a, hipStream_t s);

}  // namespace pm
