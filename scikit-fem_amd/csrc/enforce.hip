// Dirichlet rows on a device CSR matrix (skfem/utils.py:327-400, enforce):
// the data of every row in D is zeroed (pattern kept), its diagonal set to
// `diag` (the diagonal slot must exist in the pattern), and b[D] = x[D].
// One 64-lane group per row of D; rows of D are distinct, so no two groups
// write the same slot.
#include "common.h"

namespace femhip {
namespace {

__global__ void __launch_bounds__(256)
enforce_rows_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                    double* __restrict__ data, const int32_t* __restrict__ D, int64_t nD,
                    double diag, double* __restrict__ b, const double* __restrict__ x,
                    int32_t* __restrict__ missing) {
  const int64_t g = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (g >= nD) return;
  const int32_t r = D[g];
  const int32_t s = indptr[r], e = indptr[r + 1];
  bool found = false;
  for (int32_t k = s + lane; k < e; k += 64) {
    const bool is_diag = indices[k] == r;
    data[k] = is_diag ? diag : 0.0;
    found |= is_diag;
  }
  if (__any(found) == 0 && lane == 0 && missing != nullptr) missing[0] = 1;
  if (lane == 0 && b != nullptr) b[r] = x != nullptr ? x[r] : 0.0;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_enforce_rows(const int32_t* indptr, const int32_t* indices, double* data,
                                   const int32_t* D, int64_t nD, double diag, double* b,
                                   const double* x, int32_t* missing, femhip_stream_t s) {
  FEMHIP_REQUIRE(nD >= 0, FEMHIP_EINVAL, "enforce: negative count");
  if (nD == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(indptr && indices && data && D, FEMHIP_EINVAL, "enforce: NULL argument");
  enforce_rows_kernel<<<(unsigned)ceil_div(nD * 64, 256), 256, 0,
                        reinterpret_cast<hipStream_t>(s)>>>(indptr, indices, data, D, nD, diag,
                                                            b, x, missing);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
