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
  exact::LocalArgs a{G,        g->W,          sp->phi,   sp->dphi, f->kappa,
                     contrib,  g->nelems,     sp->nbfun, f->kappa_mode,
                     f->lam,   f->two_mu,     sp->phi,   sp->dphi, sp->nbfun};
  return dispatch_bilinear(g->kind, g->nq, a, f->kind, sp->vector != 0,
                           reinterpret_cast<hipStream_t>(s));
}

extern "C" int femhip_local_bilinear_mixed(const femhip_geom_t* g, const femhip_space_t* u,
                                           const femhip_space_t* v, const femhip_form_t* f,
                                           const double* G, double* contrib,
                                           femhip_stream_t s) {
  int rc = check_geom(g);
  if (rc) return rc;
  if ((rc = check_space(g, u))) return rc;
  if ((rc = check_space(g, v))) return rc;
  FEMHIP_REQUIRE(u->vector == v->vector, FEMHIP_EUNSUPPORTED,
                 "trial and test spaces must both be scalar or both ElementVector");
  FEMHIP_REQUIRE((int64_t)u->nbfun * v->nbfun <= 65535, FEMHIP_EUNSUPPORTED,
                 "too many local entries");
  FEMHIP_REQUIRE(f != nullptr, FEMHIP_EINVAL, "form is NULL");
  FEMHIP_REQUIRE(f->kind != FEMHIP_UNIT_LOAD && f->kind != FEMHIP_COEF_LOAD, FEMHIP_EINVAL,
                 "unit_load / coef_load are linear forms");
  FEMHIP_REQUIRE((f->kind != FEMHIP_KAPPA_LAPLACE && f->kind != FEMHIP_COEF_MASS) ||
                     (f->kappa != nullptr && (f->kappa_mode == 1 || f->kappa_mode == 2)),
                 FEMHIP_EINVAL, "kappa_laplace / coef_mass need the coefficient (mode 1 or 2)");
  if (g->nelems == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(G != nullptr && contrib != nullptr, FEMHIP_EINVAL, "NULL G or output");
  exact::LocalArgs a{G,        g->W,          u->phi,   u->dphi, f->kappa,
                     contrib,  g->nelems,     u->nbfun, f->kappa_mode,
                     f->lam,   f->two_mu,     v->phi,   v->dphi, v->nbfun};
  return dispatch_bilinear(g->kind, g->nq, a, f->kind, u->vector != 0,
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
  exact::LocalArgs a{G,        g->W,      sp->phi,   sp->dphi, coef ? f->kappa : nullptr,
                     contrib,  g->nelems, sp->nbfun, coef ? f->kappa_mode : 0,
                     0.0,      0.0,       sp->phi,   sp->dphi, sp->nbfun};
  return dispatch_load(g->kind, g->nq, a, reinterpret_cast<hipStream_t>(s));
}

// ---------------------------------------------------------------------------
// recheck of the fused path's suspects
// ---------------------------------------------------------------------------
#include "fused_check.h"

namespace femhip {
namespace {

struct RecheckArgs {
  const femhip_fused_check_t* c;   // unused on the device
  const int32_t* t;
  const double* p;
  const double* dphi;
  const double* phi;   // P2 mass rechecks only
  const double* W;
  const double* kappa;
  const uint64_t* queue;
  const uint64_t* mask;
  int64_t cap;
  int64_t E;
  int64_t N;
  int64_t nzeros;
  femhip_fused_state_t* st;
  int32_t* verdict;
};

constexpr int kCheckBlocks = 128;

// Entry (e, i, j) of a Laplace / kappa-Laplace / mass matrix of NB local
// functions in the reference's arithmetic (the same operations as
// local_row_kernel: the affine inverse of mapping_affine.py:55-131,
// einsum-ordered gradients, (kappa grad u) . grad v summed from 0 or
// u * v, times dx, NumPy's pairwise sum over the points) compared with 0.0;
// counts the zero entries and those whose zero-ness differs from the plan's
// mask bit of contribution (j * NB + i) * E + e; the last workgroup to
// finish forms the verdict.
template <int D, int Q, int FK, int NB = D + 1>
__global__ void __launch_bounds__(kBlock) fused_recheck_kernel(RecheckArgs a) {
#pragma clang fp contract(off)
  __shared__ unsigned long long sm[kBlock / 64], sz[kBlock / 64];
  femhip_fused_state_t* st = a.st;
  const int64_t nq = st->nsusp;
  const int64_t n = nq < a.cap ? nq : a.cap;
  unsigned long long mism = 0, zeros = 0;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kBlock) {
    const uint64_t q = a.queue[k];
    const int64_t e = (int64_t)(uint32_t)q;
    const int i = (int)((q >> 32) & 0xffu), j = (int)((q >> 40) & 0xffu);
    double inv[D][D];
    const double adet = fabs(exact::affine_inverse<D>(a.p, a.t, a.N, a.E, e, inv));
    const double kap = FK == FEMHIP_KAPPA_LAPLACE ? a.kappa[e] : 1.0;
    auto term = [&](int qq) -> double {
#pragma clang fp contract(off)
      const double dx = adet * a.W[qq];
      if constexpr (FK == FEMHIP_MASS) return (a.phi[j * Q + qq] * a.phi[i * Q + qq]) * dx;
      double gi[D], gj[D];
      exact::global_grad<D>(inv, a.dphi, i, qq, Q, gi);
      exact::global_grad<D>(inv, a.dphi, j, qq, Q, gj);
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if constexpr (FK == FEMHIP_KAPPA_LAPLACE)
          s = s + (kap * gj[d]) * gi[d];
        else
          s = s + gj[d] * gi[d];
      }
      return s * dx;
    };
    const double r = 0.0 + np_pairwise<0, Q>(term);
    const bool zero = r == 0.0;
    const uint64_t c = (uint64_t)((int64_t)(j * NB + i) * a.E + e);
    const bool bit = (a.mask[c >> 6] >> (c & 63)) & 1ull;
    mism += zero != bit ? 1 : 0;
    zeros += zero ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mism += __shfl_xor(mism, o, 64);
    zeros += __shfl_xor(zeros, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sm[threadIdx.x >> 6] = mism;
    sz[threadIdx.x >> 6] = zeros;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < kBlock / 64; ++w) {
    mism += sm[w];
    zeros += sz[w];
  }
  if (mism) atomicAdd(reinterpret_cast<unsigned long long*>(&st->mismatch), mism);
  if (zeros) atomicAdd(reinterpret_cast<unsigned long long*>(&st->zeros), zeros);
  // Every workgroup's thread 0 counts itself in st->checked after its adds
  // (release); the one that completes the count (acquire: it then sees
  // every workgroup's adds) forms the verdict and clears the counters and
  // flags for the next pass.  The count returns to zero with them, so no
  // fill is needed between calls.
  const int32_t before = __hip_atomic_fetch_add(&st->checked, 1, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_AGENT);
  if (before + 1 != (int32_t)gridDim.x) return;
  const int64_t M = __hip_atomic_load(&st->mismatch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int64_t Z = __hip_atomic_load(&st->zeros, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int32_t flags = __hip_atomic_load(&st->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int32_t* verdict = a.verdict;
  verdict[0] = (flags == 0 && M == 0 && Z == a.nzeros) ? 1 : 0;
  verdict[1] = (int32_t)(M < INT32_MAX ? M : INT32_MAX);
  verdict[2] = (int32_t)(Z < INT32_MAX ? Z : INT32_MAX);
  verdict[3] = flags;
  st->nsusp = 0;
  st->flags = 0;
  st->mismatch = 0;
  st->zeros = 0;
  st->checked = 0;
}

// the mass form has no zero entries on valid elements: the verdict only
// needs the flags and an empty zero set
__global__ void fused_verdict_kernel(femhip_fused_state_t* st, int64_t nzeros,
                                     int32_t* verdict) {
  const int32_t flags = st->flags;
  const int64_t nq = st->nsusp;
  verdict[0] = (flags == 0 && nq == 0 && nzeros == 0) ? 1 : 0;
  verdict[1] = 0;
  verdict[2] = 0;
  verdict[3] = flags;
  st->nsusp = 0;
  st->flags = 0;
}

template <int D, int Q>
int recheck_form(const RecheckArgs& a, int form, hipStream_t s) {
  if (form == FEMHIP_LAPLACE)
    fused_recheck_kernel<D, Q, FEMHIP_LAPLACE><<<kCheckBlocks, kBlock, 0, s>>>(a);
  else
    fused_recheck_kernel<D, Q, FEMHIP_KAPPA_LAPLACE><<<kCheckBlocks, kBlock, 0, s>>>(a);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

}  // namespace

int fused_check_run(const femhip_fused_check_t* c, femhip_fused_state_t* st, int form, int dim,
                    const double* p, int64_t nnodes, const double* kappa, hipStream_t s) {
  FEMHIP_REQUIRE(c->t && c->dphi && c->W && c->zero_mask && c->nelems > 0, FEMHIP_EINVAL,
                 "fused check: NULL table or no elements");
  if (form == FEMHIP_MASS) {
    fused_verdict_kernel<<<1, 1, 0, s>>>(st, c->nzeros, c->verdict);
    FEMHIP_CHECK_LAUNCH();
    return FEMHIP_OK;
  }
  FEMHIP_REQUIRE(form == FEMHIP_LAPLACE || form == FEMHIP_KAPPA_LAPLACE, FEMHIP_EUNSUPPORTED,
                 "fused check: laplace, mass or kappa_laplace");
  RecheckArgs a{c,        c->t,         p,  c->dphi,  nullptr,  c->W,      kappa, c->queue,
                c->zero_mask, c->capacity, c->nelems, nnodes, c->nzeros, st, c->verdict};
  if (dim == 2) {
    switch (c->Q) {
      case 1: return recheck_form<2, 1>(a, form, s);
      case 3: return recheck_form<2, 3>(a, form, s);
      case 4: return recheck_form<2, 4>(a, form, s);
      case 6: return recheck_form<2, 6>(a, form, s);
    }
  } else {
    switch (c->Q) {
      case 1: return recheck_form<3, 1>(a, form, s);
      case 4: return recheck_form<3, 4>(a, form, s);
      case 5: return recheck_form<3, 5>(a, form, s);
      case 11: return recheck_form<3, 11>(a, form, s);
    }
  }
  set_error("fused check: unsupported number of quadrature points");
  return FEMHIP_EUNSUPPORTED;
}

// P2 simplices (rows plans): the suspects of the last first pass rechecked
// against the plan's zero mask, and the verdict
template <int D, int NB, int Q>
int rows_recheck_form(const RecheckArgs& a, int form, hipStream_t s) {
  if (form == FEMHIP_LAPLACE)
    fused_recheck_kernel<D, Q, FEMHIP_LAPLACE, NB><<<kCheckBlocks, kBlock, 0, s>>>(a);
  else if (form == FEMHIP_KAPPA_LAPLACE)
    fused_recheck_kernel<D, Q, FEMHIP_KAPPA_LAPLACE, NB><<<kCheckBlocks, kBlock, 0, s>>>(a);
  else
    fused_recheck_kernel<D, Q, FEMHIP_MASS, NB><<<kCheckBlocks, kBlock, 0, s>>>(a);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

int rows_check_run(const femhip_fused_check_t* c, femhip_fused_state_t* st, int form,
                   const femhip_geom_t* g, const femhip_space_t* sp, const double* kappa,
                   hipStream_t s) {
  FEMHIP_REQUIRE(c->zero_mask && c->queue && c->capacity > 0 && g->nelems > 0, FEMHIP_EINVAL,
                 "rows check: NULL zero mask or queue");
  RecheckArgs a{c,          g->t,      g->p,       sp->dphi, sp->phi,  g->W,
                kappa,      c->queue,  c->zero_mask, c->capacity, g->nelems, g->nnodes,
                c->nzeros,  st,        c->verdict};
  if (g->dim == 3 && sp->nbs == 10) {
    if (g->nq == 11) return rows_recheck_form<3, 10, 11>(a, form, s);
    if (g->nq == 4) return rows_recheck_form<3, 10, 4>(a, form, s);
  } else if (g->dim == 2 && sp->nbs == 6) {
    if (g->nq == 6) return rows_recheck_form<2, 6, 6>(a, form, s);
    if (g->nq == 3) return rows_recheck_form<2, 6, 3>(a, form, s);
  }
  set_error("rows check: unsupported element or number of quadrature points");
  return FEMHIP_EUNSUPPORTED;
}

int fused_check_empty(const femhip_fused_check_t* c, femhip_fused_state_t* st, hipStream_t s) {
  fused_verdict_kernel<<<1, 1, 0, s>>>(st, c->nzeros, c->verdict);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

}  // namespace femhip
