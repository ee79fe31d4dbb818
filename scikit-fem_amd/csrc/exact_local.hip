// Mapping and local element matrices evaluated in the reference's
// floating-point operation order.
//
// The CSR pattern of the reference depends on which local entries come out
// as exactly 0.0 (eliminate_zeros runs before duplicates are summed), so
// these kernels reproduce NumPy's arithmetic bit for bit: IEEE multiply,
// add and divide with no FMA contraction (this file is compiled with
// -ffp-contract=off and every function carries the pragma), einsum's
// sequential accumulation from 0 over summed indices, and NumPy's pairwise
// add-reduction over the quadrature points.  They serve the symbolic phase
// (zero detection) and the generic numeric path of every element/form.
//
// Work decomposition: the mapping kernel runs one thread per element
// (affine) or per (element, quadrature point) (isoparametric) and writes
// invDF and |detDF| in structure-of-arrays form; the local kernel runs one
// thread per (element, local entry) with the entry on blockIdx.y, so the
// reference tables are read with wave-uniform (scalar) loads and the output
// contrib[(j*nb + i)*E + e] is written fully coalesced.
#include "exact_local.h"

namespace femhip {
namespace {

using exact::kBlock;

// ---------------------------------------------------------------------------
// mapping
// ---------------------------------------------------------------------------
// MappingAffine._init_Ab/_init_invA, mapping_affine.py:55-131.
template <int D>
__global__ void __launch_bounds__(kBlock)
affine_map_kernel(const double* __restrict__ p, const int32_t* __restrict__ t,
                  int64_t N, int64_t E, double* __restrict__ G) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  double inv[D][D];
  const double det = exact::affine_inverse<D>(p, t, N, E, e, inv);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) G[(int64_t)(i * D + j) * E + e] = inv[i][j];
  G[(int64_t)D * D * E + e] = fabs(det);
}

// MappingIsoparametric._J/detDF/invDF, mapping_isoparametric.py:112-226,
// with the mesh's own trilinear element as the mapping.
__global__ void __launch_bounds__(kBlock)
hex_map_kernel(const double* __restrict__ p, const int32_t* __restrict__ t,
               const double* __restrict__ gdphi, int64_t N, int64_t E, int Q,
               double* __restrict__ G, int32_t* __restrict__ status) {
#pragma clang fp contract(off)
  const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t EQ = E * Q;
  if (id >= EQ) return;
  const int64_t e = id / Q;
  const int q = (int)(id - e * Q);
  double J[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) J[i][j] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int32_t nk = t[(int64_t)k * E + e];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double pk = p[i * N + nk];
#pragma unroll
      for (int j = 0; j < 3; ++j) J[i][j] = J[i][j] + pk * gdphi[(k * 3 + j) * Q + q];
    }
  }
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  if (det == 0.0 && status != nullptr) status[0] = 1;
  G[0 * EQ + id] = (-J[1][2] * J[2][1] + J[1][1] * J[2][2]) / det;
  G[3 * EQ + id] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  G[6 * EQ + id] = (-J[1][1] * J[2][0] + J[1][0] * J[2][1]) / det;
  G[1 * EQ + id] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  G[4 * EQ + id] = (-J[0][2] * J[2][0] + J[0][0] * J[2][2]) / det;
  G[7 * EQ + id] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  G[2 * EQ + id] = (-J[0][2] * J[1][1] + J[0][1] * J[1][2]) / det;
  G[5 * EQ + id] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  G[8 * EQ + id] = (-J[0][1] * J[1][0] + J[0][0] * J[1][1]) / det;
  G[9 * EQ + id] = fabs(det);
}

}  // namespace

int dispatch_bilinear(int kind, int Q, const exact::LocalArgs& a, int fk, bool vec,
                      hipStream_t s) {
  if (kind == FEMHIP_TRI) return exact::dispatch_bilinear_tri(Q, a, fk, vec, s);
  if (kind == FEMHIP_TET) return exact::dispatch_bilinear_tet(Q, a, fk, vec, s);
  return exact::dispatch_bilinear_hex(Q, a, fk, vec, s);
}

int dispatch_load(int kind, int Q, const exact::LocalArgs& a, hipStream_t s) {
  if (kind == FEMHIP_TRI) return exact::dispatch_load_tri(Q, a, s);
  if (kind == FEMHIP_TET) return exact::dispatch_load_tet(Q, a, s);
  return exact::dispatch_load_hex(Q, a, s);
}

namespace {

int check_geom(const femhip_geom_t* g) {
  FEMHIP_REQUIRE(g != nullptr, FEMHIP_EINVAL, "geometry is NULL");
  FEMHIP_REQUIRE(g->kind >= FEMHIP_TRI && g->kind <= FEMHIP_HEX,
                 FEMHIP_EUNSUPPORTED, "unknown element kind");
  const int want_dim = g->kind == FEMHIP_TRI ? 2 : 3;
  const int want_nv = g->kind == FEMHIP_TRI ? 3 : (g->kind == FEMHIP_TET ? 4 : 8);
  FEMHIP_REQUIRE(g->dim == want_dim && g->nverts == want_nv, FEMHIP_EINVAL,
                 "dim / nverts do not match the element kind");
  FEMHIP_REQUIRE(g->nelems >= 0 && g->nnodes >= 0 && g->nq > 0, FEMHIP_EINVAL,
                 "negative size or no quadrature points");
  FEMHIP_REQUIRE(g->nnodes < INT32_MAX, FEMHIP_EOVERFLOW,
                 "node count exceeds int32");
  FEMHIP_REQUIRE(g->nelems == 0 || (g->p && g->t && g->W), FEMHIP_EINVAL,
                 "NULL p, t or W");
  FEMHIP_REQUIRE(g->kind != FEMHIP_HEX || g->nelems == 0 || g->gdphi,
                 FEMHIP_EINVAL, "hexahedral mapping needs gdphi");
  FEMHIP_REQUIRE(g->nelems * (int64_t)g->nq < ((int64_t)1 << 40), FEMHIP_EOVERFLOW,
                 "element count too large");
  return FEMHIP_OK;
}

int check_space(const femhip_geom_t* g, const femhip_space_t* sp) {
  FEMHIP_REQUIRE(sp != nullptr, FEMHIP_EINVAL, "space is NULL");
  FEMHIP_REQUIRE(sp->nbs > 0 && sp->nbfun == sp->nbs * (sp->vector ? g->dim : 1),
                 FEMHIP_EINVAL, "inconsistent nbs / nbfun / vector");
  FEMHIP_REQUIRE((int64_t)sp->nbfun * sp->nbfun <= 65535, FEMHIP_EUNSUPPORTED,
                 "too many local entries");
  FEMHIP_REQUIRE(g->nelems == 0 || (sp->phi && sp->dphi), FEMHIP_EINVAL,
                 "NULL reference tables");
  return FEMHIP_OK;
}

}  // namespace

}  // namespace femhip

using namespace femhip;

extern "C" int64_t femhip_geometry_size(const femhip_geom_t* g) {
  if (check_geom(g) != FEMHIP_OK) return -1;
  if (g->kind == FEMHIP_HEX) return (int64_t)10 * g->nelems * g->nq;
  return (int64_t)(g->dim * g->dim + 1) * g->nelems;
}

extern "C" int femhip_geometry(const femhip_geom_t* g, double* G, int32_t* status,
                               femhip_stream_t s) {
  int rc = check_geom(g);
  if (rc) return rc;
  if (g->nelems == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(G != nullptr, FEMHIP_EINVAL, "NULL output G");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (g->kind == FEMHIP_HEX) {
    const int64_t n = g->nelems * g->nq;
    hex_map_kernel<<<(unsigned)ceil_div(n, kBlock), kBlock, 0, st>>>(
        g->p, g->t, g->gdphi, g->nnodes, g->nelems, g->nq, G, status);
  } else if (g->dim == 2) {
    affine_map_kernel<2><<<(unsigned)ceil_div(g->nelems, kBlock), kBlock, 0, st>>>(
        g->p, g->t, g->nnodes, g->nelems, G);
  } else {
    affine_map_kernel<3><<<(unsigned)ceil_div(g->nelems, kBlock), kBlock, 0, st>>>(
        g->p, g->t, g->nnodes, g->nelems, G);
  }
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_local_bilinear(const femhip_geom_t* g, const femhip_space_t* sp,
                                     const femhip_form_t* f, const double* G,
                                     double* contrib, femhip_stream_t s) {
  int rc = check_geom(g);
  if (rc) return rc;
  if ((rc = check_space(g, sp))) return rc;
  FEMHIP_REQUIRE(f != nullptr, FEMHIP_EINVAL, "form is NULL");
  FEMHIP_REQUIRE(f->kind != FEMHIP_UNIT_LOAD, FEMHIP_EINVAL,
                 "unit_load is a linear form");
  FEMHIP_REQUIRE(f->kind != FEMHIP_COEF_LOAD, FEMHIP_EINVAL, "coef_load is a linear form");
  FEMHIP_REQUIRE((f->kind != FEMHIP_KAPPA_LAPLACE && f->kind != FEMHIP_COEF_MASS) ||
                     (f->kappa != nullptr && (f->kappa_mode == 1 || f->kappa_mode == 2)),
                 FEMHIP_EINVAL, "kappa_laplace / coef_mass need the coefficient (mode 1 or 2)");
  if (g->nelems == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(G != nullptr && contrib != nullptr, FEMHIP_EINVAL, "NULL G or output");
  exact::LocalArgs a{G, g->W, sp->phi, sp->dphi, f->kappa, contrib, g->nelems, sp->nbfun,
              f->kappa_mode, f->lam, f->two_mu};
  return dispatch_bilinear(g->kind, g->nq, a, f->kind, sp->vector != 0,
                           reinterpret_cast<hipStream_t>(s));
}

extern "C" int femhip_local_linear(const femhip_geom_t* g, const femhip_space_t* sp,
                                   const femhip_form_t* f, const double* G,
                                   double* contrib, femhip_stream_t s) {
  int rc = check_geom(g);
  if (rc) return rc;
  if ((rc = check_space(g, sp))) return rc;
  FEMHIP_REQUIRE(f != nullptr && (f->kind == FEMHIP_UNIT_LOAD || f->kind == FEMHIP_COEF_LOAD),
                 FEMHIP_EUNSUPPORTED, "the built-in linear forms are unit_load and coef_load");
  FEMHIP_REQUIRE(!sp->vector, FEMHIP_EUNSUPPORTED, "unit_load needs a scalar element");
  const bool coef = f->kind == FEMHIP_COEF_LOAD;
  FEMHIP_REQUIRE(!coef || (f->kappa != nullptr && (f->kappa_mode == 1 || f->kappa_mode == 2)),
                 FEMHIP_EINVAL, "coef_load needs the coefficient (mode 1 or 2)");
  if (g->nelems == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(G != nullptr && contrib != nullptr, FEMHIP_EINVAL, "NULL G or output");
  exact::LocalArgs a{G, g->W, sp->phi, sp->dphi, coef ? f->kappa : nullptr, contrib, g->nelems,
              sp->nbfun, coef ? f->kappa_mode : 0, 0.0, 0.0};
  return dispatch_load(g->kind, g->nq, a, reinterpret_cast<hipStream_t>(s));
}
