// Shared helpers of the femhip HIP library: error state, launch checks and
// the NumPy-order quadrature reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/femhip.h"

namespace femhip {

void set_error(const std::string& msg);

#define FEMHIP_TRY_HIP(expr)                                                  \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::femhip::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return FEMHIP_EHIP;                                                     \
    }                                                                         \
  } while (0)

#define FEMHIP_CHECK_LAUNCH()                                      \
  do {                                                             \
    hipError_t _e = hipGetLastError();                             \
    if (_e != hipSuccess) {                                        \
      ::femhip::set_error(std::string("kernel launch: ") +         \
                          hipGetErrorString(_e));                  \
      return FEMHIP_EHIP;                                          \
    }                                                              \
  } while (0)

#define FEMHIP_REQUIRE(cond, code, msg) \
  do {                                  \
    if (!(cond)) {                      \
      ::femhip::set_error(msg);         \
      return code;                      \
    }                                   \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// NumPy's float64 add-reduction over a contiguous run of n terms
// (the 8-accumulator pairwise scheme of np.add.reduce), with the terms
// produced in index order by `term(q)`.  The recursion is resolved at
// compile time.  The caller adds the reduction's initial 0.0.
template <int OFF, int N, class F>
__device__ __forceinline__ double np_pairwise(const F& term) {
#pragma clang fp contract(off)
  if constexpr (N < 8) {
    double res = -0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) res = res + term(OFF + i);
    return res;
  } else if constexpr (N <= 128) {
    double r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = term(OFF + j);
    constexpr int NB = N - (N % 8);
#pragma unroll
    for (int i = 8; i < NB; i += 8) {
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = r[j] + term(OFF + i + j);
    }
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
#pragma unroll
    for (int i = NB; i < N; ++i) res = res + term(OFF + i);
    return res;
  } else {
    constexpr int n2 = (N / 2) - ((N / 2) % 8);
    return np_pairwise<OFF, n2>(term) + np_pairwise<OFF + n2, N - n2>(term);
  }
}

}  // namespace femhip
