// Global coordinates of the quadrature points and the mesh parameter of a
// cell basis (CellBasis.global_coordinates / mesh_parameters,
// cell_basis.py:129-141), the default form parameters w.x and w.h:
//
//   affine (MappingAffine.F, mapping_affine.py:183-192):
//     x_i(e, q) = (sum_j A_ij(e) X_jq) + b_i(e), the einsum sum in j order
//     from 0, A[:, j] = p[t[j+1]] - p[t[0]], b = p[t[0]]
//   hexahedra (MappingIsoparametric.F/Fmap, mapping_isoparametric.py:50-64,170-171):
//     x_i(e, q) = sum_k p_i[t_k(e)] phi_k(X_q) over the 8 trilinear basis
//     functions in order, from 0
//   h(e, q) = |detDF(e, q)| ** (1 / dim)  (the |detDF| of the mapping pass;
//     sqrt for dim 2, as NumPy's power takes the exponent 1/2; pow for 1/3)
//
// Compiled with -ffp-contract=off: x equals the reference's bit for bit.
#include "exact_local.h"

namespace femhip {
namespace {

constexpr int kCoordBlock = 256;

template <int D>
__global__ void __launch_bounds__(kCoordBlock)
affine_coords_kernel(const double* __restrict__ p, const int32_t* __restrict__ t, int64_t N,
                     int64_t E, int Q, const double* __restrict__ X, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * kCoordBlock + threadIdx.x;
  if (idx >= E * Q) return;
  const int64_t e = idx / Q;
  const int q = (int)(idx - e * Q);
  const int32_t n0 = t[e];
  double A[D][D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const int32_t nj = t[(int64_t)(j + 1) * E + e];
#pragma unroll
    for (int i = 0; i < D; ++i) A[i][j] = p[i * N + nj] - p[i * N + n0];
  }
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) s = s + A[i][j] * X[j * Q + q];
    out[(int64_t)i * E * Q + idx] = s + p[i * N + n0];
  }
}

__global__ void __launch_bounds__(kCoordBlock)
hex_coords_kernel(const double* __restrict__ p, const int32_t* __restrict__ t, int64_t N,
                  int64_t E, int Q, const double* __restrict__ phi1, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * kCoordBlock + threadIdx.x;
  if (idx >= E * Q) return;
  const int64_t e = idx / Q;
  const int q = (int)(idx - e * Q);
  double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int32_t nk = t[(int64_t)k * E + e];
    const double ph = phi1[k * Q + q];
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = x[i] + p[i * N + nk] * ph;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) out[(int64_t)i * E * Q + idx] = x[i];
}

// h = |detDF| ** (1 / dim); |detDF| per element (affine, G row D*D) or per
// point (hexahedra, G row 9)
__global__ void __launch_bounds__(kCoordBlock)
mesh_param_kernel(const double* __restrict__ absdet, bool per_point, int64_t E, int Q,
                  double expo, double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * kCoordBlock + threadIdx.x;
  if (idx >= E * Q) return;
  const double d = per_point ? absdet[idx] : absdet[idx / Q];
  // NumPy's ``**`` evaluates a scalar exponent of 1/2 as sqrt
  out[idx] = expo == 0.5 ? sqrt(d) : pow(d, expo);
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_global_coordinates(const femhip_geom_t* g, const double* X,
                                         const double* phi1, double* out, femhip_stream_t s) {
  FEMHIP_REQUIRE(g != nullptr && out != nullptr, FEMHIP_EINVAL,
                 "global coordinates: NULL argument");
  FEMHIP_REQUIRE(g->dim == 2 || g->dim == 3, FEMHIP_EUNSUPPORTED,
                 "global coordinates: 2-D or 3-D meshes only");
  const bool hex = g->kind == FEMHIP_HEX;
  FEMHIP_REQUIRE(hex ? phi1 != nullptr : X != nullptr, FEMHIP_EINVAL,
                 "global coordinates: need the points (simplices) or the trilinear basis "
                 "values (hexahedra)");
  const int64_t E = g->nelems;
  const int Q = g->nq;
  if (E == 0 || Q == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(g->p != nullptr && g->t != nullptr, FEMHIP_EINVAL,
                 "global coordinates: NULL mesh arrays");
  const unsigned grid = (unsigned)ceil_div(E * Q, kCoordBlock);
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (hex)
    hex_coords_kernel<<<grid, kCoordBlock, 0, st>>>(g->p, g->t, g->nnodes, E, Q, phi1, out);
  else if (g->dim == 2)
    affine_coords_kernel<2><<<grid, kCoordBlock, 0, st>>>(g->p, g->t, g->nnodes, E, Q, X, out);
  else
    affine_coords_kernel<3><<<grid, kCoordBlock, 0, st>>>(g->p, g->t, g->nnodes, E, Q, X, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_mesh_parameters(const femhip_geom_t* g, const double* G, double* out,
                                      femhip_stream_t s) {
  FEMHIP_REQUIRE(g != nullptr && G != nullptr && out != nullptr, FEMHIP_EINVAL,
                 "mesh parameters: NULL argument");
  FEMHIP_REQUIRE(g->dim == 2 || g->dim == 3, FEMHIP_EUNSUPPORTED,
                 "mesh parameters: 2-D or 3-D meshes only");
  const int64_t E = g->nelems;
  const int Q = g->nq;
  if (E == 0 || Q == 0) return FEMHIP_OK;
  const bool hex = g->kind == FEMHIP_HEX;
  const double* absdet = hex ? G + (int64_t)9 * E * Q : G + (int64_t)g->dim * g->dim * E;
  const unsigned grid = (unsigned)ceil_div(E * Q, kCoordBlock);
  mesh_param_kernel<<<grid, kCoordBlock, 0, reinterpret_cast<hipStream_t>(s)>>>(
      absdet, hex, E, Q, 1.0 / (double)g->dim, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
