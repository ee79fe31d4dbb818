// Boundary forms on a FacetBasis (facet_basis.py:23-110) for P1/P2
// triangles and tetrahedra: the local entries of the facet mass u*v and the
// facet load 1*v (models/poisson.py), and the facet normals, in the
// reference's arithmetic, so that the exact zeros -- the basis functions of
// the vertex opposite a facet, evaluated at the facet's quadrature points --
// come out as the reference's do.
//
// Per facet f (vertices facets[:, f], owning element e = tind[f]) and facet
// quadrature point X_q:
//   x   = B X_q + c              MappingAffine.G, mapping_affine.py:234-245
//         (B[:, j] = p[f_{j+1}] - p[f_0], c = p[f_0]; einsum in index order)
//   Y   = invA_e (x - b_e)       MappingAffine.invF, mapping_affine.py:195-203
//         (invA_e, b_e = p[t[0, e]] from the element's vertices,
//         mapping_affine.py:55-131)
//   dx  = |detB| W_q             detB: mapping_affine.py:158-181
//   phi_i(Y) by the element's polynomial in the reference's operation order
//         (element_tri_p1.py, element_tri_p2.py, element_tet_p1.py,
//         element_tet_p2.py)
// and, as BilinearForm._kernel / LinearForm._kernel do, the entries are
// np.sum((phi_j phi_i) dx, axis=1) and np.sum((1.0 phi_i) dx, axis=1).
//
// One thread per facet computes the facet's geometry once and all of its
// entries; the output is facet-minor (entry k of facet f at k * nf + f), so
// every store of a wave is one contiguous run.  Compiled with
// -ffp-contract=off.
#include "exact_local.h"

namespace femhip {
namespace {

constexpr int kFacetBlock = 256;

template <int D, int DEG>
constexpr int nbasis() {
  return D == 2 ? (DEG == 1 ? 3 : 6) : (DEG == 1 ? 4 : 10);
}

// P1 / P2 simplex basis values at Y (reference expressions, left to right)
template <int D, int DEG>
__device__ __forceinline__ double lbasis(int i, const double (&Y)[D]) {
#pragma clang fp contract(off)
  if constexpr (D == 2) {
    const double x = Y[0], y = Y[1];
    if constexpr (DEG == 1) {
      if (i == 0) return 1. - x - y;
      return i == 1 ? x : y;
    } else {
      switch (i) {
        case 0: return 1. - 3. * x - 3. * y + 2. * (x * x) + 4. * x * y + 2. * (y * y);
        case 1: return 2. * (x * x) - x;
        case 2: return 2. * (y * y) - y;
        case 3: return 4. * x - 4. * (x * x) - 4. * x * y;
        case 4: return 4. * x * y;
        default: return 4. * y - 4. * x * y - 4. * (y * y);
      }
    }
  } else {
    const double x = Y[0], y = Y[1], z = Y[2];
    if constexpr (DEG == 1) {
      switch (i) {
        case 0: return 1. - x - y - z;
        case 1: return x;
        case 2: return y;
        default: return z;
      }
    } else {
      switch (i) {
        case 0:
          return 1. - 3. * x + 2. * (x * x) - 3. * y + 4. * x * y + 2. * (y * y) - 3. * z +
                 4. * x * z + 4. * y * z + 2. * (z * z);
        case 1: return -1. * x + 2. * (x * x);
        case 2: return -1. * y + 2. * (y * y);
        case 3: return -1. * z + 2. * (z * z);
        case 4: return 4. * x - 4. * (x * x) - 4. * x * y - 4. * x * z;
        case 5: return 4. * x * y;
        case 6: return 0. + 4. * y - 4. * x * y - 4. * (y * y) - 4. * y * z;
        case 7: return 0. + 4. * z - 4. * x * z - 4. * y * z - 4. * (z * z);
        case 8: return 0. + 4. * x * z;
        default: return 0. + 4. * y * z;
      }
    }
  }
}

// reference gradients of the P1 / P2 simplex basis at Y (the dphi of the
// same element files, left to right; the 0*x components are 0)
template <int D, int DEG>
__device__ __forceinline__ void ldbasis(int i, const double (&Y)[D], double (&g)[D]) {
#pragma clang fp contract(off)
  if constexpr (DEG == 1) {
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = (i == 0) ? -1. : (i == d + 1 ? 1. : 0.);
  } else if constexpr (D == 2) {
    const double x = Y[0], y = Y[1];
    switch (i) {
      case 0: g[0] = -3. + 4. * x + 4. * y; g[1] = -3. + 4. * x + 4. * y; break;
      case 1: g[0] = 4. * x - 1.; g[1] = 0.; break;
      case 2: g[0] = 0.; g[1] = 4. * y - 1.; break;
      case 3: g[0] = 4. - 8. * x - 4. * y; g[1] = -4. * x; break;
      case 4: g[0] = 4. * y; g[1] = 4. * x; break;
      default: g[0] = -4. * y; g[1] = 4. - 4. * x - 8. * y; break;
    }
  } else {
    const double x = Y[0], y = Y[1], z = Y[2];
    switch (i) {
      case 0: {
        const double v = -3. + 4. * x + 4. * y + 4. * z;
        g[0] = v; g[1] = v; g[2] = v;
        break;
      }
      case 1: g[0] = -1. + 4. * x; g[1] = 0.; g[2] = 0.; break;
      case 2: g[0] = 0.; g[1] = -1. + 4. * y; g[2] = 0.; break;
      case 3: g[0] = 0.; g[1] = 0.; g[2] = -1. + 4. * z; break;
      case 4: g[0] = 4. - 8. * x - 4. * y - 4. * z; g[1] = -4. * x; g[2] = -4. * x; break;
      case 5: g[0] = 4. * y; g[1] = 4. * x; g[2] = 0.; break;
      case 6: g[0] = -4. * y; g[1] = 4. - 4. * x - 8. * y - 4. * z; g[2] = -4. * y; break;
      case 7: g[0] = -4. * z; g[1] = -4. * z; g[2] = 4. - 4. * x - 4. * y - 8. * z; break;
      case 8: g[0] = 4. * z; g[1] = 0.; g[2] = 4. * x; break;
      default: g[0] = 0.; g[1] = 4. * z; g[2] = 4. * y; break;
    }
  }
}

struct FacetArgs {
  const double* p;        // [D][nnodes]
  const int32_t* facets;  // [D][nf] vertices of the facets
  const int32_t* t;       // [D+1][E] element vertices
  const int32_t* tind;    // [nf] owning element
  const double* X;        // [D-1][Q] facet quadrature points
  const double* W;        // [Q]
  double* out;            // mass [(j nb + i) nf + f], load [i nf + f]
  int64_t nnodes, E, nf;
  const double* coef;     // COEF_MASS / COEF_LOAD: [nf] (mode 1) or [nf][Q] (mode 2)
  int coef_mode;
};

// the facet mapping x = B X + c and its area scaling detB
// (MappingAffine._init_boundary_mapping, mapping_affine.py:155-181)
template <int D>
__device__ __forceinline__ double facet_frame(const FacetArgs& a, int64_t f, double (&B)[D][D - 1],
                                              double (&c)[D]) {
#pragma clang fp contract(off)
  const int64_t N = a.nnodes;
  int32_t fv[D];
#pragma unroll
  for (int k = 0; k < D; ++k) fv[k] = a.facets[(int64_t)k * a.nf + f];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    c[i] = a.p[i * N + fv[0]];
#pragma unroll
    for (int j = 0; j < D - 1; ++j) B[i][j] = a.p[i * N + fv[j + 1]] - a.p[i * N + fv[0]];
  }
  if constexpr (D == 2) {
    return sqrt(B[0][0] * B[0][0] + B[1][0] * B[1][0]);
  } else {
    const double t0 = B[1][0] * B[2][1] - B[2][0] * B[1][1];
    const double t1 = -B[0][0] * B[2][1] + B[2][0] * B[0][1];
    const double t2 = B[0][0] * B[1][1] - B[1][0] * B[0][1];
    return sqrt(t0 * t0 + t1 * t1 + t2 * t2);
  }
}

// the facet's geometry: Y (reference coordinates of the facet quadrature
// points in the owning element) and dx
template <int D, int Q>
struct FacetPoints {
  double Y[Q][D];
  double dx[Q];
  double inv[D][D];  // the owning element's invA
  __device__ __forceinline__ void init(const FacetArgs& a, int64_t f) {
#pragma clang fp contract(off)
    const int64_t N = a.nnodes;
    double B[D][D - 1], c[D];
    const double detB = facet_frame<D>(a, f, B, c);
    const int64_t e = a.tind[f];
    double b[D];
    exact::affine_inverse<D>(a.p, a.t, N, a.E, e, inv);
    const int32_t n0 = a.t[e];
#pragma unroll
    for (int i = 0; i < D; ++i) b[i] = a.p[i * N + n0];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      double y[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D - 1; ++j) s = s + B[i][j] * a.X[j * Q + q];
        y[i] = (s + c[i]) - b[i];
      }
#pragma unroll
      for (int i = 0; i < D; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s = s + inv[i][j] * y[j];
        Y[q][i] = s;
      }
      dx[q] = fabs(detB) * a.W[q];
    }
  }
};

template <int D, int DEG, int Q, int FORM>
__global__ void __launch_bounds__(kFacetBlock) facet_local_kernel(FacetArgs a) {
#pragma clang fp contract(off)
  constexpr int NB = nbasis<D, DEG>();
  const int64_t f = (int64_t)blockIdx.x * kFacetBlock + threadIdx.x;
  if (f >= a.nf) return;
  FacetPoints<D, Q> fp;
  fp.init(a, f);
  double phi[NB][Q];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int i = 0; i < NB; ++i) phi[i][q] = lbasis<D, DEG>(i, fp.Y[q]);
  double k[Q];
  if constexpr (FORM == FEMHIP_COEF_MASS || FORM == FEMHIP_COEF_LOAD) {
#pragma unroll
    for (int q = 0; q < Q; ++q) k[q] = a.coef_mode == 1 ? a.coef[f] : a.coef[f * Q + q];
  }
  const int64_t nf = a.nf;
  if constexpr (FORM == FEMHIP_MASS) {
    // (phi_j phi_i) dx == (phi_i phi_j) dx exactly: each unordered pair once
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int i = 0; i <= j; ++i) {
        auto term = [&](int q) -> double {
#pragma clang fp contract(off)
          return (phi[j][q] * phi[i][q]) * fp.dx[q];
        };
        const double v = 0.0 + np_pairwise<0, Q>(term);
        a.out[(int64_t)(j * NB + i) * nf + f] = v;
        if (i != j) a.out[(int64_t)(i * NB + j) * nf + f] = v;
      }
  } else if constexpr (FORM == FEMHIP_COEF_MASS) {
    // w.c * u * v = ((c phi_j) phi_i): not symmetric in the last bit
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        auto term = [&](int q) -> double {
#pragma clang fp contract(off)
          return ((k[q] * phi[j][q]) * phi[i][q]) * fp.dx[q];
        };
        a.out[(int64_t)(j * NB + i) * nf + f] = 0.0 + np_pairwise<0, Q>(term);
      }
  } else {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      auto term = [&](int q) -> double {
#pragma clang fp contract(off)
        if constexpr (FORM == FEMHIP_COEF_LOAD) return (k[q] * phi[i][q]) * fp.dx[q];
        return (1.0 * phi[i][q]) * fp.dx[q];
      };
      a.out[(int64_t)i * nf + f] = 0.0 + np_pairwise<0, Q>(term);
    }
  }
}

// AbstractBasis.interpolate (abstract_basis.py:271-322) on the facets: a DOF
// vector w at the facet quadrature points, value and global gradient, one
// term per local basis function in index order after the 0 * (first term)
// start; the gradient of phi_i is sum_k invA[k][j] dphi_i[k] (ElementH1.gbasis,
// element_h1.py:10-24, einsum 'ijkl,ikl->jkl' in k order).
template <int D, int DEG, int Q>
__global__ void __launch_bounds__(kFacetBlock)
facet_interpolate_kernel(FacetArgs a, const int32_t* __restrict__ dofs, const double* __restrict__ w,
                         double* __restrict__ value, double* __restrict__ grad) {
#pragma clang fp contract(off)
  constexpr int NB = nbasis<D, DEG>();
  const int64_t f = (int64_t)blockIdx.x * kFacetBlock + threadIdx.x;
  if (f >= a.nf) return;
  FacetPoints<D, Q> fp;
  fp.init(a, f);
  const int64_t nf = a.nf;
  double v[Q], g[Q][D];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const double wi = w[dofs[(int64_t)i * nf + f]];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double tv = wi * lbasis<D, DEG>(i, fp.Y[q]);
      double dref[D], tg[D];
      if (grad != nullptr) {
        ldbasis<D, DEG>(i, fp.Y[q], dref);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
#pragma unroll
          for (int kk = 0; kk < D; ++kk) s = s + fp.inv[kk][j] * dref[kk];
          tg[j] = wi * s;
        }
      }
      if (i == 0) {
        v[q] = 0.0 * tv;
        if (grad != nullptr)
#pragma unroll
          for (int j = 0; j < D; ++j) g[q][j] = 0.0 * tg[j];
      }
      v[q] = v[q] + tv;
      if (grad != nullptr)
#pragma unroll
        for (int j = 0; j < D; ++j) g[q][j] = g[q][j] + tg[j];
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    value[f * Q + q] = v[q];
    if (grad != nullptr)
#pragma unroll
      for (int j = 0; j < D; ++j) grad[((int64_t)j * nf + f) * Q + q] = g[q][j];
  }
}

// the facet quadrature points in global coordinates, x = B X + c
// (MappingAffine.G, mapping_affine.py:234-245; FacetBasis.global_coordinates,
// facet_basis.py:126-131), and the mesh parameter |detB| ** (1 / (dim - 1))
// (FacetBasis.mesh_parameters, facet_basis.py:133-139)
template <int D>
__global__ void __launch_bounds__(kFacetBlock)
facet_coords_kernel(FacetArgs a, int Q, double* __restrict__ x, double* __restrict__ h) {
#pragma clang fp contract(off)
  const int64_t f = (int64_t)blockIdx.x * kFacetBlock + threadIdx.x;
  if (f >= a.nf) return;
  double B[D][D - 1], c[D];
  const double detB = facet_frame<D>(a, f, B, c);
  for (int q = 0; q < Q; ++q) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < D - 1; ++j) s = s + B[i][j] * a.X[j * Q + q];
      x[((int64_t)i * a.nf + f) * Q + q] = s + c[i];
    }
  }
  if (h != nullptr) {
    // NumPy's ``**`` evaluates the scalar exponents 1 and 1/2 as a copy and
    // as sqrt
    const double v = D == 2 ? fabs(detB) : sqrt(fabs(detB));
    for (int q = 0; q < Q; ++q) h[f * Q + q] = v;
  }
}

// outward unit normals seen from element tn[f] through its local facet
// floc[f] (MappingAffine.normals, mapping_affine.py:255-281): the reference
// cell's facet normal N (RefTri/RefTet facet order), n_j = sum_i invA[i][j] N_i
// (einsum 'ijkl,ik->jkl', i in order), |n|^2 summed over j in order, and
// n_j * (1 / |n|); the same value at every quadrature point.
template <int D>
__global__ void __launch_bounds__(kFacetBlock)
facet_normals_kernel(const double* __restrict__ p, const int32_t* __restrict__ t,
                     const int32_t* __restrict__ tn, const int32_t* __restrict__ floc,
                     int64_t N, int64_t E, int64_t nf, int Q, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t f = (int64_t)blockIdx.x * kFacetBlock + threadIdx.x;
  if (f >= nf) return;
  const int64_t e = tn[f];
  const int k = floc[f];
  double inv[D][D];
  exact::affine_inverse<D>(p, t, N, E, e, inv);
  double Nr[D];
  if constexpr (D == 2) {
    // facets (0,1), (1,2), (0,2)
    Nr[0] = k == 0 ? 0. : (k == 1 ? 1. : -1.);
    Nr[1] = k == 0 ? -1. : (k == 1 ? 1. : 0.);
  } else {
    // facets (0,1,2), (0,1,3), (0,2,3), (1,2,3)
    Nr[0] = k == 2 ? -1. : (k == 3 ? 1. : 0.);
    Nr[1] = k == 1 ? -1. : (k == 3 ? 1. : 0.);
    Nr[2] = k == 0 ? -1. : (k == 3 ? 1. : 0.);
  }
  double n[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) acc = acc + inv[i][j] * Nr[i];
    n[j] = acc;
  }
  double l2 = n[0] * n[0];
#pragma unroll
  for (int j = 1; j < D; ++j) l2 = l2 + n[j] * n[j];
  const double r = 1. / sqrt(l2);
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const double v = n[j] * r;
    for (int q = 0; q < Q; ++q) out[((int64_t)j * nf + f) * Q + q] = v;
  }
}

// form >= 0: local entries of that form; form < 0: interpolation
struct FacetInterp {
  const int32_t* dofs;
  const double* w;
  double* value;
  double* grad;
};

template <int D, int DEG, int Q>
int launch(const FacetArgs& a, int form, const FacetInterp* ip, hipStream_t s) {
  const unsigned grid = (unsigned)ceil_div(a.nf, kFacetBlock);
  if (ip != nullptr)
    facet_interpolate_kernel<D, DEG, Q><<<grid, kFacetBlock, 0, s>>>(a, ip->dofs, ip->w, ip->value,
                                                                     ip->grad);
  else if (form == FEMHIP_MASS)
    facet_local_kernel<D, DEG, Q, FEMHIP_MASS><<<grid, kFacetBlock, 0, s>>>(a);
  else if (form == FEMHIP_COEF_MASS)
    facet_local_kernel<D, DEG, Q, FEMHIP_COEF_MASS><<<grid, kFacetBlock, 0, s>>>(a);
  else if (form == FEMHIP_COEF_LOAD)
    facet_local_kernel<D, DEG, Q, FEMHIP_COEF_LOAD><<<grid, kFacetBlock, 0, s>>>(a);
  else
    facet_local_kernel<D, DEG, Q, FEMHIP_UNIT_LOAD><<<grid, kFacetBlock, 0, s>>>(a);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

template <int D, int DEG>
int dispatch_q(const FacetArgs& a, int form, const FacetInterp* ip, int Q, hipStream_t s) {
  switch (Q) {
    case 1: return launch<D, DEG, 1>(a, form, ip, s);
    case 2: return launch<D, DEG, 2>(a, form, ip, s);
    case 3: return launch<D, DEG, 3>(a, form, ip, s);
    case 4: return launch<D, DEG, 4>(a, form, ip, s);
    case 5: return launch<D, DEG, 5>(a, form, ip, s);
    case 6: return launch<D, DEG, 6>(a, form, ip, s);
    case 7: return launch<D, DEG, 7>(a, form, ip, s);
    default:
      set_error("facet forms: 1 to 7 facet quadrature points");
      return FEMHIP_EUNSUPPORTED;
  }
}

int dispatch(const femhip_facet_geom_t* g, const FacetArgs& a, int form, const FacetInterp* ip,
             hipStream_t st) {
  if (g->dim == 2)
    return g->degree == 1 ? dispatch_q<2, 1>(a, form, ip, g->nq, st)
                          : dispatch_q<2, 2>(a, form, ip, g->nq, st);
  return g->degree == 1 ? dispatch_q<3, 1>(a, form, ip, g->nq, st)
                        : dispatch_q<3, 2>(a, form, ip, g->nq, st);
}

int check_facet_geom(const femhip_facet_geom_t* g) {
  FEMHIP_REQUIRE(g != nullptr, FEMHIP_EINVAL, "facet forms: NULL geometry");
  FEMHIP_REQUIRE((g->kind == FEMHIP_TRI && g->dim == 2) || (g->kind == FEMHIP_TET && g->dim == 3),
                 FEMHIP_EUNSUPPORTED, "facet forms: triangles or tetrahedra only");
  FEMHIP_REQUIRE(g->degree == 1 || g->degree == 2, FEMHIP_EUNSUPPORTED,
                 "facet forms: P1 or P2 only");
  FEMHIP_REQUIRE(g->nfacets >= 0 && g->nnodes >= 0 && g->nelems >= 0 && g->nq >= 1,
                 FEMHIP_EINVAL, "facet forms: negative size or no quadrature points");
  if (g->nfacets > 0)
    FEMHIP_REQUIRE(g->p && g->facets && g->t && g->tind && g->X && g->W, FEMHIP_EINVAL,
                   "facet forms: NULL array");
  return FEMHIP_OK;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_facet_local(const femhip_facet_geom_t* g, int32_t form, const double* coef,
                                  int32_t coef_mode, double* out, femhip_stream_t s) {
  int rc = check_facet_geom(g);
  if (rc) return rc;
  FEMHIP_REQUIRE(form == FEMHIP_MASS || form == FEMHIP_UNIT_LOAD || form == FEMHIP_COEF_MASS ||
                     form == FEMHIP_COEF_LOAD,
                 FEMHIP_EUNSUPPORTED, "facet forms: mass, unit_load, coef_mass, coef_load");
  const bool coefd = form == FEMHIP_COEF_MASS || form == FEMHIP_COEF_LOAD;
  FEMHIP_REQUIRE(!coefd || (coef_mode == 1 || coef_mode == 2), FEMHIP_EINVAL,
                 "facet forms: coefficient mode 1 (per facet) or 2 (per point)");
  if (g->nfacets == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(out != nullptr && (!coefd || coef != nullptr), FEMHIP_EINVAL,
                 "facet forms: NULL output or coefficient");
  FacetArgs a{g->p, g->facets, g->t, g->tind, g->X, g->W, out, g->nnodes, g->nelems,
              g->nfacets, coefd ? coef : nullptr, coef_mode};
  return dispatch(g, a, form, nullptr, reinterpret_cast<hipStream_t>(s));
}

extern "C" int femhip_facet_interpolate(const femhip_facet_geom_t* g, const int32_t* dofs,
                                        const double* w, double* value, double* grad,
                                        femhip_stream_t s) {
  int rc = check_facet_geom(g);
  if (rc) return rc;
  if (g->nfacets == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(dofs != nullptr && w != nullptr && value != nullptr, FEMHIP_EINVAL,
                 "facet interpolate: NULL argument");
  FacetArgs a{g->p, g->facets, g->t, g->tind, g->X, g->W, nullptr, g->nnodes, g->nelems,
              g->nfacets, nullptr, 0};
  const FacetInterp ip{dofs, w, value, grad};
  return dispatch(g, a, -1, &ip, reinterpret_cast<hipStream_t>(s));
}

extern "C" int femhip_facet_normals(const femhip_facet_geom_t* g, const int32_t* tind_normals,
                                    const int32_t* floc, double* out, femhip_stream_t s) {
  int rc = check_facet_geom(g);
  if (rc) return rc;
  if (g->nfacets == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(tind_normals != nullptr && floc != nullptr && out != nullptr, FEMHIP_EINVAL,
                 "facet normals: NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  const unsigned grid = (unsigned)ceil_div(g->nfacets, kFacetBlock);
  if (g->dim == 2)
    facet_normals_kernel<2><<<grid, kFacetBlock, 0, st>>>(g->p, g->t, tind_normals, floc,
                                                          g->nnodes, g->nelems, g->nfacets,
                                                          g->nq, out);
  else
    facet_normals_kernel<3><<<grid, kFacetBlock, 0, st>>>(g->p, g->t, tind_normals, floc,
                                                          g->nnodes, g->nelems, g->nfacets,
                                                          g->nq, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_facet_global_coordinates(const femhip_facet_geom_t* g, double* x, double* h,
                                               femhip_stream_t s) {
  int rc = check_facet_geom(g);
  if (rc) return rc;
  if (g->nfacets == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(x != nullptr, FEMHIP_EINVAL, "facet coordinates: NULL output");
  FacetArgs a{g->p, g->facets, g->t, g->tind, g->X, g->W, nullptr, g->nnodes, g->nelems,
              g->nfacets, nullptr, 0};
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  const unsigned grid = (unsigned)ceil_div(g->nfacets, kFacetBlock);
  if (g->dim == 2)
    facet_coords_kernel<2><<<grid, kFacetBlock, 0, st>>>(a, g->nq, x, h);
  else
    facet_coords_kernel<3><<<grid, kFacetBlock, 0, st>>>(a, g->nq, x, h);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
