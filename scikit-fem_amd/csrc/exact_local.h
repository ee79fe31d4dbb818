// Templates of the exact-order local element kernels, instantiated per
// element family in exact_tri.hip, exact_tet.hip and exact_hex.hip (split so
// the heavy instantiations compile in parallel).  Every translation unit that
// includes this header is compiled with -ffp-contract=off.
#pragma once

#include "common.h"

namespace femhip {
namespace exact {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// local entries
// ---------------------------------------------------------------------------
struct LocalArgs {
  const double* G;
  const double* W;
  const double* phi;
  const double* dphi;
  const double* kappa;
  double* out;
  int64_t E;
  int nb;               // trial functions (u); the only space of a one-space form
  int kappa_mode;
  double lam, two_mu;
  // test functions (v): the same tables as u unless the form has a test
  // basis of its own (bilinear_form.py:58-128 with vbasis != ubasis)
  const double* phiv;
  const double* dphiv;
  int nbv;
};

// Inverse Jacobian and |det| at (e, q): affine values are per element,
// isoparametric ones per quadrature point.
template <int D, bool AFF, int Q>
struct MapAt {
  double inv[D][D];
  double adet;
  __device__ __forceinline__ void load(const double* __restrict__ G, int64_t E,
                                       int64_t e, int q) {
    const int64_t stride = AFF ? E : E * Q;
    const int64_t at = AFF ? e : e * Q + q;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) inv[i][j] = G[(i * D + j) * stride + at];
    adet = G[(D * D) * stride + at];
  }
};

// MappingAffine._init_Ab/_init_invA (mapping_affine.py:55-131) for element e:
// A[:, j] = p[t[j+1, e]] - p[t[0, e]], the determinant by cofactor expansion
// along the first row and every entry of invA as the reference writes it.
// Returns the signed determinant.  Used by every kernel that needs the affine
// mapping of an element (the geometry pass and the facet kernels).
template <int D>
__device__ __forceinline__ double affine_inverse(const double* __restrict__ p,
                                                 const int32_t* __restrict__ t, int64_t N,
                                                 int64_t E, int64_t e, double (&inv)[D][D]) {
#pragma clang fp contract(off)
  const int32_t n0 = t[e];
  double A[D][D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const int32_t nj = t[(int64_t)(j + 1) * E + e];
#pragma unroll
    for (int i = 0; i < D; ++i) A[i][j] = p[i * N + nj] - p[i * N + n0];
  }
  if constexpr (D == 2) {
    const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    inv[0][0] = A[1][1] / det;
    inv[0][1] = -A[0][1] / det;
    inv[1][0] = -A[1][0] / det;
    inv[1][1] = A[0][0] / det;
    return det;
  } else {
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    inv[0][0] = (-A[1][2] * A[2][1] + A[1][1] * A[2][2]) / det;
    inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    inv[2][0] = (-A[1][1] * A[2][0] + A[1][0] * A[2][1]) / det;
    inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    inv[1][1] = (-A[0][2] * A[2][0] + A[0][0] * A[2][2]) / det;
    inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    inv[0][2] = (-A[0][2] * A[1][1] + A[0][1] * A[1][2]) / det;
    inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    inv[2][2] = (-A[0][1] * A[1][0] + A[0][0] * A[1][1]) / det;
    return det;
  }
}

// grad[j] = sum_i invDF[i, j] * dphi[i]   (einsum 'ijkl,il->jkl',
// element_h1.py:17), accumulated from 0 in index order.
template <int D>
__device__ __forceinline__ void global_grad(const double (&inv)[D][D],
                                            const double* __restrict__ dphi,
                                            int fn, int q, int Q,
                                            double (&g)[D]) {
#pragma clang fp contract(off)
  double d[D];
#pragma unroll
  for (int i = 0; i < D; ++i) d[i] = dphi[(fn * D + i) * Q + q];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) s = s + inv[i][j] * d[i];
    g[j] = s;
  }
}

// ddot(C(sym_grad u), sym_grad v) for u = phi_a e_cu, v = phi_b e_cv
// (models/elasticity.py:35-53, helpers.py:71-150, element_vector.py:36-48).
template <int D>
__device__ __forceinline__ double elasticity_value(const double (&gu)[D], int cu,
                                                   const double (&gv)[D], int cv,
                                                   double lam, double two_mu) {
#pragma clang fp contract(off)
  double Gu[D][D], Gv[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      Gu[i][j] = (i == cu) ? gu[j] : 0.0;
      Gv[i][j] = (i == cv) ? gv[j] : 0.0;
    }
  double Eu[D][D], Ev[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      Eu[i][j] = 0.5 * (Gu[i][j] + Gu[j][i]);
      Ev[i][j] = 0.5 * (Gv[i][j] + Gv[j][i]);
    }
  double tr = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) tr = tr + Eu[i][i];
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const double c = two_mu * Eu[i][j] + lam * ((i == j) ? tr : 0.0 * tr);
      s = s + c * Ev[i][j];
    }
  return s;
}

// Simplices: the nb*nb entry blocks of one block of kBlock elements are
// adjacent in the grid, so the elements' mapping is fetched once into L2 and
// read there by every entry.  Hexahedra (Q points x 10 values of mapping per
// element, up to 27 KB): blocks of kIsoBlock elements, and workgroup b, which
// the dispatcher places on XCD b % 8, takes an element block that only that
// XCD works on, so one XCD's L2 holds the mappings its entries read.
constexpr int kIsoBlock = 64;
constexpr int kXcds = 8;
template <int D, bool AFF, int Q, int FK, bool VEC>
__global__ void __launch_bounds__(kBlock) local_bilinear_kernel(LocalArgs a) {
#pragma clang fp contract(off)
  const int64_t E = a.E;
  const unsigned nb2 = (unsigned)(a.nb * a.nbv);
  int64_t e;
  int entry;  // = j * nbv + i
  if constexpr (AFF) {
    e = (int64_t)(blockIdx.x / nb2) * kBlock + threadIdx.x;
    entry = (int)(blockIdx.x % nb2);
  } else {
    const unsigned xcd = blockIdx.x % kXcds, r = blockIdx.x / kXcds;
    entry = (int)(r % nb2);
    e = ((int64_t)(r / nb2) * kXcds + xcd) * kIsoBlock + threadIdx.x;
  }
  if (e >= E) return;
  const int j = entry / a.nbv;
  const int i = entry - j * a.nbv;
  const int fu = VEC ? j / D : j, cu = VEC ? j - (j / D) * D : 0;
  const int fv = VEC ? i / D : i, cv = VEC ? i - (i / D) * D : 0;

  MapAt<D, AFF, Q> m;
  if constexpr (AFF) m.load(a.G, E, e, 0);

  auto term = [&](int q) -> double {
#pragma clang fp contract(off)
    if constexpr (!AFF) m.load(a.G, E, e, q);
    const double dx = m.adet * a.W[q];
    double val;
    if constexpr (FK == FEMHIP_MASS) {
      val = a.phi[fu * Q + q] * a.phiv[fv * Q + q];
    } else if constexpr (FK == FEMHIP_COEF_MASS) {
      // w.c * u * v, evaluated left to right
      const double k = (a.kappa_mode == 1) ? a.kappa[e] : a.kappa[e * Q + q];
      val = (k * a.phi[fu * Q + q]) * a.phiv[fv * Q + q];
    } else {
      double gu[D], gv[D];
      global_grad<D>(m.inv, a.dphi, fu, q, Q, gu);
      global_grad<D>(m.inv, a.dphiv, fv, q, Q, gv);
      if constexpr (FK == FEMHIP_LAPLACE) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) s = s + gu[d] * gv[d];
        val = s;
      } else if constexpr (FK == FEMHIP_KAPPA_LAPLACE) {
        const double k = (a.kappa_mode == 1) ? a.kappa[e] : a.kappa[e * Q + q];
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) s = s + (k * gu[d]) * gv[d];
        val = s;
      } else {
        val = elasticity_value<D>(gu, cu, gv, cv, a.lam, a.two_mu);
      }
    }
    return val * dx;
  };
  const double r = 0.0 + np_pairwise<0, Q>(term);
  a.out[(int64_t)entry * E + e] = r;
}

// Affine scalar forms, one thread per (element, test function i): the
// element's mapping and the test function's values or gradients at every
// point are formed once for the row, and each entry (j, i) is the same sum
// as in local_bilinear_kernel.  laplace and mass integrands are symmetric in
// (u, v) bit for bit, so only j >= i is computed and stored at both (j, i)
// and (i, j); kappa_laplace ((k grad u) . grad v) and coef_mass ((c u) v)
// are not, and compute the whole row.
template <int D, int Q, int FK>
__global__ void __launch_bounds__(kBlock) local_row_kernel(LocalArgs a) {
#pragma clang fp contract(off)
  constexpr bool GRAD = FK == FEMHIP_LAPLACE || FK == FEMHIP_KAPPA_LAPLACE;
  constexpr bool SYM = FK == FEMHIP_LAPLACE || FK == FEMHIP_MASS;
  constexpr bool COEF = FK == FEMHIP_KAPPA_LAPLACE || FK == FEMHIP_COEF_MASS;
  const int64_t E = a.E;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  const int i = blockIdx.y;
  const int nb = a.nb;
  MapAt<D, true, Q> m;
  m.load(a.G, E, e, 0);
  double vi[Q][GRAD ? D : 1];
  double k[COEF ? Q : 1];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if constexpr (GRAD)
      global_grad<D>(m.inv, a.dphi, i, q, Q, vi[q]);
    else
      vi[q][0] = a.phi[i * Q + q];
    if constexpr (COEF) k[q] = (a.kappa_mode == 1) ? a.kappa[e] : a.kappa[e * Q + q];
  }
  for (int j = SYM ? i : 0; j < nb; ++j) {
    auto term = [&](int q) -> double {
#pragma clang fp contract(off)
      const double dx = m.adet * a.W[q];
      double val;
      if constexpr (GRAD) {
        double gj[D];
        global_grad<D>(m.inv, a.dphi, j, q, Q, gj);
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          if constexpr (FK == FEMHIP_KAPPA_LAPLACE)
            s = s + (k[q] * gj[d]) * vi[q][d];
          else
            s = s + gj[d] * vi[q][d];
        }
        val = s;
      } else if constexpr (FK == FEMHIP_COEF_MASS) {
        val = (k[q] * a.phi[j * Q + q]) * vi[q][0];
      } else {
        val = a.phi[j * Q + q] * vi[q][0];
      }
      return val * dx;
    };
    const double r = 0.0 + np_pairwise<0, Q>(term);
    a.out[(int64_t)(j * nb + i) * E + e] = r;
    if (SYM && j != i) a.out[(int64_t)(i * nb + j) * E + e] = r;
  }
}

template <int D, bool AFF, int Q, bool COEF>
__global__ void __launch_bounds__(kBlock) local_load_kernel(LocalArgs a) {
#pragma clang fp contract(off)
  // isoparametric: the nb row blocks of one element block are adjacent in
  // the grid (the per-point mapping is read from L2, as in the entry kernel)
  const int64_t E = a.E;
  const unsigned eb = AFF ? blockIdx.x : blockIdx.x / (unsigned)a.nb;
  const int i = AFF ? (int)blockIdx.y : (int)(blockIdx.x % (unsigned)a.nb);
  const int64_t e = (int64_t)eb * kBlock + threadIdx.x;
  if (e >= E) return;
  MapAt<D, AFF, Q> m;
  if constexpr (AFF) m.load(a.G, E, e, 0);
  auto term = [&](int q) -> double {
#pragma clang fp contract(off)
    if constexpr (!AFF) m.load(a.G, E, e, q);
    const double dx = m.adet * a.W[q];
    if constexpr (COEF) {  // w.c * v
      const double k = (a.kappa_mode == 1) ? a.kappa[e] : a.kappa[e * Q + q];
      return (k * a.phi[i * Q + q]) * dx;
    }
    return a.phi[i * Q + q] * dx;
  };
  a.out[(int64_t)i * E + e] = 0.0 + np_pairwise<0, Q>(term);
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <int D, bool AFF, int Q>
int launch_bilinear_q(const LocalArgs& a, int fk, bool vec, hipStream_t s) {
  if constexpr (AFF) {
    // scalar forms of one space on simplices: one thread per (element, row)
    if (!vec && fk != FEMHIP_ELASTICITY && a.phiv == a.phi && a.nbv == a.nb) {
      const dim3 rgrid((unsigned)ceil_div(a.E, kBlock), (unsigned)a.nb);
      switch (fk) {
        case FEMHIP_LAPLACE:
          local_row_kernel<D, Q, FEMHIP_LAPLACE><<<rgrid, kBlock, 0, s>>>(a);
          break;
        case FEMHIP_MASS:
          local_row_kernel<D, Q, FEMHIP_MASS><<<rgrid, kBlock, 0, s>>>(a);
          break;
        case FEMHIP_KAPPA_LAPLACE:
          local_row_kernel<D, Q, FEMHIP_KAPPA_LAPLACE><<<rgrid, kBlock, 0, s>>>(a);
          break;
        case FEMHIP_COEF_MASS:
          local_row_kernel<D, Q, FEMHIP_COEF_MASS><<<rgrid, kBlock, 0, s>>>(a);
          break;
        default:
          set_error("scalar elements support laplace, mass, kappa_laplace and coef_mass");
          return FEMHIP_EUNSUPPORTED;
      }
      FEMHIP_CHECK_LAUNCH();
      return FEMHIP_OK;
    }
  }
  const int64_t nblk = AFF ? ceil_div(a.E, kBlock)
                           : ceil_div(ceil_div(a.E, kIsoBlock), kXcds) * kXcds;
  FEMHIP_REQUIRE(nblk * a.nb * a.nbv <= INT32_MAX, FEMHIP_EOVERFLOW,
                 "local entries: grid exceeds 2^31 blocks");
  const unsigned grid = (unsigned)(nblk * a.nb * a.nbv);
  const int threads = AFF ? kBlock : kIsoBlock;
  if (vec) {
    FEMHIP_REQUIRE(fk == FEMHIP_ELASTICITY, FEMHIP_EUNSUPPORTED,
                   "vector elements support only linear_elasticity");
    local_bilinear_kernel<D, AFF, Q, FEMHIP_ELASTICITY, true><<<grid, threads, 0, s>>>(a);
  } else {
    switch (fk) {
      case FEMHIP_LAPLACE:
        local_bilinear_kernel<D, AFF, Q, FEMHIP_LAPLACE, false><<<grid, threads, 0, s>>>(a);
        break;
      case FEMHIP_MASS:
        local_bilinear_kernel<D, AFF, Q, FEMHIP_MASS, false><<<grid, threads, 0, s>>>(a);
        break;
      case FEMHIP_KAPPA_LAPLACE:
        local_bilinear_kernel<D, AFF, Q, FEMHIP_KAPPA_LAPLACE, false><<<grid, threads, 0, s>>>(a);
        break;
      case FEMHIP_COEF_MASS:
        local_bilinear_kernel<D, AFF, Q, FEMHIP_COEF_MASS, false><<<grid, threads, 0, s>>>(a);
        break;
      default:
        set_error("scalar elements support laplace, mass, kappa_laplace and coef_mass");
        return FEMHIP_EUNSUPPORTED;
    }
  }
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

template <int D, bool AFF, int Q>
int launch_load_q(const LocalArgs& a, hipStream_t s) {
  const dim3 grid = AFF ? dim3((unsigned)ceil_div(a.E, kBlock), (unsigned)a.nb)
                        : dim3((unsigned)(ceil_div(a.E, kBlock) * a.nb));
  if (a.kappa != nullptr)
    local_load_kernel<D, AFF, Q, true><<<grid, kBlock, 0, s>>>(a);
  else
    local_load_kernel<D, AFF, Q, false><<<grid, kBlock, 0, s>>>(a);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}


#define FEMHIP_QCASE(D, AFF, QQ, CALL) \
  case QQ:                             \
    return CALL<D, AFF, QQ>

int dispatch_bilinear_tri(int Q, const LocalArgs& a, int fk, bool vec, hipStream_t s);
int dispatch_bilinear_tet(int Q, const LocalArgs& a, int fk, bool vec, hipStream_t s);
int dispatch_bilinear_hex(int Q, const LocalArgs& a, int fk, bool vec, hipStream_t s);
int dispatch_load_tri(int Q, const LocalArgs& a, hipStream_t s);
int dispatch_load_tet(int Q, const LocalArgs& a, hipStream_t s);
int dispatch_load_hex(int Q, const LocalArgs& a, hipStream_t s);

}  // namespace exact
}  // namespace femhip
