// arith.hpp -- internal interface of the elementwise kernels (arith.hip).
#pragma once
#include "pm_common.hpp"

namespace pm {

enum { EW_SUP = 0, EW_INF = 1, EW_SSUB = 2, EW_SADD = 3 };
enum { EWC_SSUB = 0, EWC_SADD = 1 };

template <typename T, int OP> int ew2(const T *a, const T *b, T *o, int64_t n, hipStream_t s);
template <typename T, int OP> int ewc(const T *a, T c, T *o, int64_t n, hipStream_t s);
template <typename T, bool OPENING>
int dpo_marker(const T *ext, const T *tmp, const T *img, T *o, int64_t n, hipStream_t s);

}  // namespace pm
