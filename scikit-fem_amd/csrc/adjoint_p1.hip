// Adjoint of P1 (kappa-)Laplace + unit-load assembly on simplices.
//
// Given upstream gradients on the CSR values (grad_data, on the pattern the
// forward produced) and on the load vector (grad_load), one thread per
// element recomputes the element geometry -- gradients g_a of the
// barycentric functions and the volume V = |det D| / d! -- instead of
// storing it, and returns
//   dL/dkappa_e = sum_ab Gs_ab V g_a.g_b            (Gs = sym. upstream)
//   dL/dx_a     = V [ kappa (tr(H) I - 2 H) + beta I ] g_a
// with H = sum_ab Gs_ab g_a g_b^T and beta = sum_c grad_load[t_c] / (d+1).
// This follows from K = kappa V Lambda Lambda^T, Lambda = B D^-1,
// dV/dD = V D^-T and d(D^-1) = -D^-1 dD D^-1.  The per-vertex gradients are
// written as contributions gX[d][a*E + e]; the caller sums them per node
// with the deterministic rows plan (femhip_segment_sum).
#include "common.h"

namespace femhip {
namespace {

constexpr int kBlock = 256;

template <int D>
__global__ void __launch_bounds__(kBlock)
p1_adjoint_kernel(int64_t E, int64_t N, const double* __restrict__ p,
                  const int32_t* __restrict__ t, const double* __restrict__ kappa,
                  const int32_t* __restrict__ slot_of, const double* __restrict__ gdata,
                  const double* __restrict__ gload, double* __restrict__ dkappa,
                  double* __restrict__ gX) {
  constexpr int NV = D + 1;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= E) return;
  int32_t v[NV];
  double x[NV][D];
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    v[a] = t[(int64_t)a * E + e];
#pragma unroll
    for (int d = 0; d < D; ++d) x[a][d] = p[d * N + v[a]];
  }
  // gradients g_a = n_a / det and V
  double g[NV][D], V;
  if constexpr (D == 3) {
    double c1[3], c2[3], c3[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      c1[d] = x[1][d] - x[0][d];
      c2[d] = x[2][d] - x[0][d];
      c3[d] = x[3][d] - x[0][d];
    }
    double n[4][3];
    n[1][0] = c2[1] * c3[2] - c2[2] * c3[1];
    n[1][1] = c2[2] * c3[0] - c2[0] * c3[2];
    n[1][2] = c2[0] * c3[1] - c2[1] * c3[0];
    n[2][0] = c3[1] * c1[2] - c3[2] * c1[1];
    n[2][1] = c3[2] * c1[0] - c3[0] * c1[2];
    n[2][2] = c3[0] * c1[1] - c3[1] * c1[0];
    n[3][0] = c1[1] * c2[2] - c1[2] * c2[1];
    n[3][1] = c1[2] * c2[0] - c1[0] * c2[2];
    n[3][2] = c1[0] * c2[1] - c1[1] * c2[0];
#pragma unroll
    for (int d = 0; d < 3; ++d) n[0][d] = -(n[1][d] + n[2][d] + n[3][d]);
    const double det = c1[0] * n[1][0] + c1[1] * n[1][1] + c1[2] * n[1][2];
    const double r = 1.0 / det;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) g[a][d] = n[a][d] * r;
    V = fabs(det) * (1.0 / 6.0);
  } else {
    const double c1x = x[1][0] - x[0][0], c1y = x[1][1] - x[0][1];
    const double c2x = x[2][0] - x[0][0], c2y = x[2][1] - x[0][1];
    const double det = c1x * c2y - c1y * c2x;
    const double r = 1.0 / det;
    g[1][0] = c2y * r;
    g[1][1] = -c2x * r;
    g[2][0] = -c1y * r;
    g[2][1] = c1x * r;
    g[0][0] = -(g[1][0] + g[2][0]);
    g[0][1] = -(g[1][1] + g[2][1]);
    V = fabs(det) * 0.5;
  }
  const double kap = kappa ? kappa[e] : 1.0;
  double H[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) H[i][j] = 0.0;
  double dk = 0.0;
  if (gdata != nullptr) {
    // upstream values of the element's entries (j = b trial, i = a test),
    // each slot read once; Gs is symmetric, so the pairs (a, b) and (b, a)
    // are taken together
    double G[NV][NV];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
      for (int b = 0; b < NV; ++b) {
        const int32_t sl = slot_of[((int64_t)b * NV + a) * E + e];
        G[a][b] = sl >= 0 ? gdata[sl] : 0.0;
      }
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
      for (int b = a; b < NV; ++b) {
        const double w = (a == b) ? G[a][a] : G[a][b] + G[b][a];  // 2 Gs off the diagonal
        double gg = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) gg += g[a][d] * g[b][d];
        dk += w * gg;
        const double h = (a == b) ? w : 0.5 * w;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j)
            H[i][j] += (a == b) ? h * g[a][i] * g[a][j]
                                : h * (g[a][i] * g[b][j] + g[b][i] * g[a][j]);
      }
  }
  if (dkappa != nullptr) dkappa[e] = V * dk;
  double beta = 0.0;
  if (gload != nullptr) {
#pragma unroll
    for (int a = 0; a < NV; ++a) beta += gload[v[a]];
    beta *= 1.0 / (double)NV;
  }
  double tr = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) tr += H[i][i];
  double M[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j)
      M[i][j] = V * (kap * ((i == j ? tr : 0.0) - 2.0 * H[i][j]) + (i == j ? beta : 0.0));
  const int64_t CE = (int64_t)NV * E;
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) s += M[i][j] * g[a][j];
      gX[(int64_t)i * CE + (int64_t)a * E + e] = s;
    }
}

// out[c * out_stride + s] = sum over k in [seg[s], seg[s+1]) of
// contrib[c * in_stride + perm[k]], c < ncomp, summed in perm order.
__global__ void __launch_bounds__(kBlock)
segment_sum_multi_kernel(const double* __restrict__ contrib, int32_t ncomp, int64_t in_stride,
                         const int32_t* __restrict__ perm, const int32_t* __restrict__ seg,
                         int64_t nseg, double* __restrict__ out, int64_t out_stride) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= nseg) return;
  double acc[3] = {0.0, 0.0, 0.0};
  const int32_t b = seg[s], en = seg[s + 1];
  int32_t k = b;
  // four contributions' gathers are issued together; every component is
  // summed in the same order as one at a time
  for (; k + 4 <= en; k += 4) {
    int64_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = perm[k + u];
    double v[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < 3; ++i) v[u][i] = i < ncomp ? contrib[i * in_stride + c[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (i < ncomp) acc[i] += v[u][i];
  }
  for (; k < en; ++k) {
    const int64_t c = perm[k];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (i < ncomp) acc[i] += contrib[i * in_stride + c];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (i < ncomp) out[i * out_stride + s] = acc[i];
}

__global__ void __launch_bounds__(kBlock)
invert_plan_kernel(const int32_t* __restrict__ perm, const int32_t* __restrict__ seg,
                   int64_t nseg, int32_t* __restrict__ slot_of) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= nseg) return;
  for (int32_t k = seg[s]; k < seg[s + 1]; ++k) slot_of[perm[k]] = (int32_t)s;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_invert_plan(const int32_t* perm, const int32_t* seg, int64_t nseg,
                                  int32_t* slot_of, femhip_stream_t s) {
  FEMHIP_REQUIRE(nseg >= 0, FEMHIP_EINVAL, "negative segment count");
  if (nseg == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(perm && seg && slot_of, FEMHIP_EINVAL, "NULL argument");
  invert_plan_kernel<<<(unsigned)ceil_div(nseg, kBlock), kBlock, 0,
                       reinterpret_cast<hipStream_t>(s)>>>(perm, seg, nseg, slot_of);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_p1_adjoint(int32_t dim, int64_t nelems, int64_t nnodes, const double* p,
                                 const int32_t* t, const double* kappa,
                                 const int32_t* slot_of, const double* grad_data,
                                 const double* grad_load, double* dkappa, double* gX,
                                 femhip_stream_t s) {
  FEMHIP_REQUIRE(dim == 2 || dim == 3, FEMHIP_EUNSUPPORTED, "adjoint: P1 tri/tet only");
  FEMHIP_REQUIRE(nelems >= 0 && nnodes > 0, FEMHIP_EINVAL, "adjoint: bad sizes");
  if (nelems == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(p && t && gX, FEMHIP_EINVAL, "adjoint: NULL p, t or gX");
  FEMHIP_REQUIRE(grad_data == nullptr || slot_of != nullptr, FEMHIP_EINVAL,
                 "adjoint: grad_data needs slot_of");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  const unsigned grid = (unsigned)ceil_div(nelems, kBlock);
  if (dim == 3)
    p1_adjoint_kernel<3><<<grid, kBlock, 0, st>>>(nelems, nnodes, p, t, kappa, slot_of,
                                                  grad_data, grad_load, dkappa, gX);
  else
    p1_adjoint_kernel<2><<<grid, kBlock, 0, st>>>(nelems, nnodes, p, t, kappa, slot_of,
                                                  grad_data, grad_load, dkappa, gX);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_segment_sum_multi(const double* contrib, int32_t ncomp, int64_t in_stride,
                                        const int32_t* perm, const int32_t* seg, int64_t nseg,
                                        double* out, int64_t out_stride, femhip_stream_t s) {
  FEMHIP_REQUIRE(ncomp >= 1 && ncomp <= 3, FEMHIP_EINVAL, "segment_sum_multi: 1 to 3 components");
  FEMHIP_REQUIRE(nseg >= 0 && in_stride >= 0 && out_stride >= nseg, FEMHIP_EINVAL,
                 "segment_sum_multi: bad sizes");
  if (nseg == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(contrib && perm && seg && out, FEMHIP_EINVAL, "segment_sum_multi: NULL argument");
  segment_sum_multi_kernel<<<(unsigned)ceil_div(nseg, kBlock), kBlock, 0,
                             reinterpret_cast<hipStream_t>(s)>>>(contrib, ncomp, in_stride, perm,
                                                                 seg, nseg, out, out_stride);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
