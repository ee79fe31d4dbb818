// A scalar CSR matrix expanded to the components of ElementVector
// (element_vector.py:36-48): the vector DOF of component c of scalar DOF s
// is dim*s + c (every DOF block of Dofs grows by the factor dim, in the
// entity-major order of dofs.py:264-334), and a form whose value for equal
// components is the scalar form's value, and exact zero otherwise, gives
//
//   row dim*r + c = scalar row r with columns dim*s + c, same values.
//
// vector_laplace (models/poisson.py:12-14, ddot(grad u, grad v)) is such a
// form: the einsum over the component and direction indices adds the scalar
// dot product's terms in the same order, all other terms being 0 * 0.
#include "common.h"

namespace femhip {
namespace {

constexpr int kExpandBlock = 256;

// one thread per (scalar row, component)
__global__ void __launch_bounds__(kExpandBlock)
expand_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
              const double* __restrict__ data, int64_t n, int dim,
              int32_t* __restrict__ oindptr, int32_t* __restrict__ oindices,
              double* __restrict__ odata) {
  const int64_t id = (int64_t)blockIdx.x * kExpandBlock + threadIdx.x;
  if (id == 0) oindptr[n * dim] = (int32_t)(dim * (int64_t)indptr[n]);
  if (id >= n * dim) return;
  const int64_t r = id / dim;
  const int c = (int)(id - r * dim);
  const int64_t s0 = indptr[r], len = indptr[r + 1] - s0;
  const int64_t o0 = dim * s0 + c * len;
  oindptr[id] = (int32_t)o0;
  for (int64_t k = 0; k < len; ++k) {
    oindices[o0 + k] = dim * indices[s0 + k] + c;
    if (data != nullptr) odata[o0 + k] = data[s0 + k];
  }
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_csr_expand_components(const int32_t* indptr, const int32_t* indices,
                                            const double* data, int64_t n, int64_t nnz,
                                            int32_t dim, int32_t* out_indptr,
                                            int32_t* out_indices, double* out_data,
                                            femhip_stream_t s) {
  FEMHIP_REQUIRE(indptr != nullptr && out_indptr != nullptr, FEMHIP_EINVAL,
                 "csr expand: NULL argument");
  FEMHIP_REQUIRE(dim >= 1 && dim <= 3 && n >= 0 && nnz >= 0, FEMHIP_EINVAL,
                 "csr expand: dim 1..3 and non-negative sizes");
  FEMHIP_REQUIRE((int64_t)dim * nnz <= INT32_MAX && (int64_t)dim * n < INT32_MAX,
                 FEMHIP_EOVERFLOW, "csr expand: the expanded matrix exceeds int32 indices");
  FEMHIP_REQUIRE(nnz == 0 || (indices != nullptr && out_indices != nullptr), FEMHIP_EINVAL,
                 "csr expand: NULL index arrays");
  FEMHIP_REQUIRE(data == nullptr || out_data != nullptr, FEMHIP_EINVAL,
                 "csr expand: values given but no output for them");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  const unsigned grid = (unsigned)ceil_div(n * dim > 0 ? n * dim : 1, kExpandBlock);
  expand_kernel<<<grid, kExpandBlock, 0, st>>>(indptr, indices, data, n, dim, out_indptr,
                                               out_indices, out_data);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
