// Basis.interpolate (abstract_basis.py:271-322) for scalar elements: a DOF
// vector w evaluated at every quadrature point,
//
//   value[e, q]   = 0 * (w_0 phi_0) + w_0 phi_0 + w_1 phi_1 + ...
//   grad[d, e, q] = the same sum with the global gradients of the basis,
//
// one term per local basis function in index order (the reference's
// `out = 0. * einsum(...)` followed by `out += einsum(...)`), with the global
// gradients formed as in ElementH1.gbasis (element_h1.py:10-24).  Compiled
// with -ffp-contract=off: the values equal the reference's bit for bit.
#include "exact_local.h"

namespace femhip {
namespace {

constexpr int kBlock = 256;

template <int D, bool AFF>
__global__ void __launch_bounds__(kBlock)
interpolate_kernel(const double* __restrict__ w, const int32_t* __restrict__ dofs, int nb,
                   int64_t E, int Q, const double* __restrict__ phi,
                   const double* __restrict__ dphi, const double* __restrict__ G,
                   double* __restrict__ value, double* __restrict__ grad) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= E * Q) return;
  const int64_t e = idx / Q;
  const int q = (int)(idx - e * Q);
  double inv[D][D];
  if (grad != nullptr) {
    const int64_t stride = AFF ? E : E * Q;
    const int64_t at = AFF ? e : e * Q + q;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) inv[i][j] = G[(i * D + j) * stride + at];
  }
  double v = 0.0, g[D];
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = 0.0;
  for (int i = 0; i < nb; ++i) {
    const double wi = w[dofs[(int64_t)i * E + e]];
    const double tv = wi * phi[i * Q + q];
    double tg[D];
    if (grad != nullptr) {
      exact::global_grad<D>(inv, dphi, i, q, Q, tg);
#pragma unroll
      for (int d = 0; d < D; ++d) tg[d] = wi * tg[d];
    }
    if (i == 0) {  // out = 0. * (w_0 phi_0)
      v = 0.0 * tv;
      if (grad != nullptr)
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = 0.0 * tg[d];
    }
    v = v + tv;
    if (grad != nullptr)
#pragma unroll
      for (int d = 0; d < D; ++d) g[d] = g[d] + tg[d];
  }
  value[idx] = v;
  if (grad != nullptr)
#pragma unroll
    for (int d = 0; d < D; ++d) grad[(int64_t)d * E * Q + idx] = g[d];
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_interpolate(const femhip_geom_t* g, const femhip_space_t* sp,
                                  const double* G, const double* w, double* value,
                                  double* grad, femhip_stream_t s) {
  FEMHIP_REQUIRE(g != nullptr && sp != nullptr && w != nullptr && value != nullptr,
                 FEMHIP_EINVAL, "interpolate: NULL argument");
  FEMHIP_REQUIRE(!sp->vector, FEMHIP_EUNSUPPORTED, "interpolate: scalar elements only");
  FEMHIP_REQUIRE(grad == nullptr || G != nullptr, FEMHIP_EINVAL,
                 "interpolate: gradients need the mapping");
  FEMHIP_REQUIRE(grad == nullptr || sp->dphi != nullptr, FEMHIP_EINVAL,
                 "interpolate: gradients need the reference gradients");
  FEMHIP_REQUIRE(sp->phi != nullptr && sp->element_dofs != nullptr, FEMHIP_EINVAL,
                 "interpolate: NULL basis tables");
  const int64_t E = g->nelems;
  const int Q = g->nq;
  if (E == 0 || Q == 0) return FEMHIP_OK;
  const unsigned grid = (unsigned)ceil_div(E * Q, kBlock);
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  const bool aff = g->kind != FEMHIP_HEX;
  if (g->dim == 3) {
    if (aff)
      interpolate_kernel<3, true><<<grid, kBlock, 0, st>>>(w, sp->element_dofs, sp->nbfun, E, Q, sp->phi,
                                                           sp->dphi, G, value, grad);
    else
      interpolate_kernel<3, false><<<grid, kBlock, 0, st>>>(w, sp->element_dofs, sp->nbfun, E, Q,
                                                            sp->phi, sp->dphi, G, value, grad);
  } else if (g->dim == 2) {
    interpolate_kernel<2, true><<<grid, kBlock, 0, st>>>(w, sp->element_dofs, sp->nbfun, E, Q, sp->phi,
                                                         sp->dphi, G, value, grad);
  } else {
    set_error("interpolate: 2-D or 3-D meshes only");
    return FEMHIP_EUNSUPPORTED;
  }
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
