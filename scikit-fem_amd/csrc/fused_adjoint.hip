// Adjoint of the fused P1 (kappa-)Laplace + unit-load assembly, owner-computes
// on the forward's plan (fused_plan.cpp).
//
// Given the upstream gradients on the CSR values (grad_data, on the
// connectivity pattern the plan was built for) and on the load vector
// (grad_load), the adjoint of K_e = kappa_e V Lambda Lambda^T (Lambda_a =
// g_a, the gradients of the barycentric functions, V = |det| / d!) and of
// b_a = V / (d+1) is, per element,
//   dL/dkappa_e = V tr(H),   H = sum_ab Gs_ab g_a g_b^T   (Gs: symmetric part
//                                                          of the upstream)
//   dL/dx_a    += V [ kappa (tr(H) I - 2 H) + beta I ] g_a,
//                 beta = sum_c grad_load[c] / (d+1)
// (dV/dD = V D^-T and d(D^-1) = -D^-1 dD D^-1 for the edge matrix D).
//
// One workgroup per partition, as in the forward: the partition's local
// nodes (owned + halo) have their coordinates, CSR row starts, upstream
// diagonal values and upstream load staged in LDS; every element touching an
// owned node is computed by the lane holding its record, which reads the
// element's twelve (2-D: six) off-diagonal upstream values from global memory
// (owned and halo rows alike: the record carries the column offsets of every
// vertex's row) and adds the node gradients of its owned vertices into LDS.
// The node sums are int64 fixed point with one exponent per launch
// (fused_common.h), so they are deterministic and no per-element gradient
// array is written.  dL/dkappa_e is written by the partition owning the
// element's vertex 0.
#include "fused_common.h"

namespace femhip {
namespace {

struct AdjArgs {
  const int32_t* parts;
  const int32_t* nodemap;
  const uint32_t* recs;
  const int32_t* rec_elem;
  const int32_t* indptr;     // CSR row starts of the pattern
  const int32_t* diag_slot;  // CSR slot of each row's diagonal
  const double* p;
  const double* kappa;       // per element, or NULL (kappa = 1)
  const double* gdata;       // upstream on the CSR values, or NULL
  const double* gload;       // upstream on the load vector, or NULL
  double* dkappa;            // per element, or NULL
  double* dX;                // [dim][nnodes]
  femhip_fused_state_t* st;
  int64_t nnodes;
  int32_t nparts;
  int32_t max_nodes;
  int32_t max_rows;
  int32_t mode;
};

constexpr int kAdjThreads = 512;

// LDS: [dX accumulators (max_rows x D int64) | coordinates (max_nodes x D) |
//  upstream diagonal (max_nodes) | upstream load (max_nodes) | reduction |
//  row starts (max_nodes int32)]
template <int D, int NT>
struct AdjLds {
  unsigned long long* acc;
  double* X;
  double* gdg;
  double* gl;
  double* red;
  int32_t* grs;
  __device__ AdjLds(double* smem, const AdjArgs& a) {
    acc = reinterpret_cast<unsigned long long*>(smem);
    X = smem + (size_t)D * a.max_rows;
    gdg = X + (size_t)D * a.max_nodes;
    gl = gdg + a.max_nodes;
    red = gl + a.max_nodes;
    grs = reinterpret_cast<int32_t*>(red + 2 * (NT / 64));
  }
};

size_t adjoint_lds_bytes(const femhip_fused_plan_info_t& I, int dim) {
  return 8 * ((size_t)dim * I.max_rows + (size_t)dim * I.max_nodes + 2 * (size_t)I.max_nodes +
              2 * (kAdjThreads / 64)) +
         4 * (size_t)I.max_nodes;
}

// Gradients g_a of the barycentric functions (n_a / det, n_a the
// cross-product normals; 2-D: rotated edges) and V = |det| / d!.
template <int D>
__device__ __forceinline__ void simplex_gradients(const double (&x)[D + 1][D],
                                                  double (&g)[D + 1][D], double& V) {
  if constexpr (D == 3) {
    double c[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      c[0][d] = x[1][d] - x[0][d];
      c[1][d] = x[2][d] - x[0][d];
      c[2][d] = x[3][d] - x[0][d];
    }
    double n[4][3];
    n[1][0] = c[1][1] * c[2][2] - c[1][2] * c[2][1];
    n[1][1] = c[1][2] * c[2][0] - c[1][0] * c[2][2];
    n[1][2] = c[1][0] * c[2][1] - c[1][1] * c[2][0];
    n[2][0] = c[2][1] * c[0][2] - c[2][2] * c[0][1];
    n[2][1] = c[2][2] * c[0][0] - c[2][0] * c[0][2];
    n[2][2] = c[2][0] * c[0][1] - c[2][1] * c[0][0];
    n[3][0] = c[0][1] * c[1][2] - c[0][2] * c[1][1];
    n[3][1] = c[0][2] * c[1][0] - c[0][0] * c[1][2];
    n[3][2] = c[0][0] * c[1][1] - c[0][1] * c[1][0];
#pragma unroll
    for (int d = 0; d < 3; ++d) n[0][d] = -(n[1][d] + n[2][d] + n[3][d]);
    const double det = c[0][0] * n[1][0] + c[0][1] * n[1][1] + c[0][2] * n[1][2];
    const double r = recip(det);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) g[a][d] = n[a][d] * r;
    V = fabs(det) * (1.0 / 6.0);
  } else {
    const double c1x = x[1][0] - x[0][0], c1y = x[1][1] - x[0][1];
    const double c2x = x[2][0] - x[0][0], c2y = x[2][1] - x[0][1];
    const double det = c1x * c2y - c1y * c2x;
    const double r = recip(det);
    g[1][0] = c2y * r;
    g[1][1] = -c2x * r;
    g[2][0] = -c1y * r;
    g[2][1] = c1x * r;
    g[0][0] = -(g[1][0] + g[2][0]);
    g[0][1] = -(g[1][1] + g[2][1]);
    V = fabs(det) * 0.5;
  }
}

template <int D, int RW, int IDB, int OFB, int NT>
__global__ void __launch_bounds__(NT) p1_fused_adjoint_kernel(AdjArgs a) {
  constexpr int NV = D + 1;
  extern __shared__ double smem[];
  AdjLds<D, NT> L(smem, a);
  const int tid = threadIdx.x;
  femhip_fused_state_t* st = a.st;
  const bool first = a.mode == 0;
  const FxPass pass = fx_begin<false>(st, a.mode);
  if (!pass.run) return;  // uniform
  const bool scatter = pass.scatter;
  const double sc = ldexp(1.0, pass.s);
  double mk = 0.0;

  for (int k = 0;; ++k) {
    const int part = next_partition(a.nparts, k);
    if (part < 0) break;
    if (k > 0) __syncthreads();  // LDS reuse
    const int32_t* hdr = a.parts + 8 * (int64_t)part;
    const int node_start = hdr[0], Ln = hdr[1], R = hdr[3];
    const int rec_start = hdr[5], nrec = hdr[6];

    // staging: per local node its coordinates, CSR row start, upstream
    // diagonal and upstream load; zeroed accumulators of the owned nodes
    for (int l = tid; l < Ln; l += NT) {
      const int64_t g = a.nodemap[node_start + l];
#pragma unroll
      for (int d = 0; d < D; ++d) L.X[D * l + d] = a.p[d * a.nnodes + g];
      L.grs[l] = a.indptr[g];
      L.gdg[l] = a.gdata ? a.gdata[a.diag_slot[g]] : 0.0;
      L.gl[l] = a.gload ? a.gload[g] : 0.0;
    }
    if (scatter)
      for (int i = tid; i < D * R; i += NT) L.acc[i] = 0ull;
    __syncthreads();

    // records and element ids: chunk c + 1 loads while chunk c is processed
    // (unconditional loads, an index past the records clamped to the last
    // one, so every path has the same loads in flight)
    struct Chunk {
      uint32_t w[RW];
      int32_t e;
    };
    auto fetch = [&](int c, Chunk& ch) {
      const int r = c * NT + tid;
      const int64_t ri = rec_start + (r < nrec ? r : nrec - 1);
#pragma unroll
      for (int q = 0; q < RW; ++q) ch.w[q] = __builtin_nontemporal_load(a.recs + ri * RW + q);
      ch.e = a.rec_elem[ri];
    };
    auto process = [&](const Chunk& ch, int c) {
      if (c * NT + tid < nrec) {
        const uint32_t (&w)[RW] = ch.w;
        const int32_t e = ch.e;
        int lv[NV];
        lv[0] = bits<0, IDB, RW>(w);
        lv[1] = bits<IDB, IDB, RW>(w);
        lv[2] = bits<2 * IDB, IDB, RW>(w);
        if constexpr (NV == 4) lv[3] = bits<3 * IDB, IDB, RW>(w);
        uint32_t off[NV][NV - 1];
        decode_fields<NV * IDB, OFB, RW>(w, off[0], std::make_integer_sequence<int, NV - 1>{});
        decode_fields<NV * IDB + (NV - 1) * OFB, OFB, RW>(
            w, off[1], std::make_integer_sequence<int, NV - 1>{});
        decode_fields<NV * IDB + 2 * (NV - 1) * OFB, OFB, RW>(
            w, off[2], std::make_integer_sequence<int, NV - 1>{});
        if constexpr (NV == 4)
          decode_fields<NV * IDB + 3 * (NV - 1) * OFB, OFB, RW>(
              w, off[3], std::make_integer_sequence<int, NV - 1>{});
        // upstream values of the element's entries: row a, column b
        double G[NV][NV];
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) {
          const int64_t rs = L.grs[lv[aa]];
          int q = 0;
#pragma unroll
          for (int b = 0; b < NV; ++b) {
            if (b == aa) continue;
            G[aa][b] = a.gdata ? a.gdata[rs + off[aa][q]] : 0.0;
            ++q;
          }
          G[aa][aa] = L.gdg[lv[aa]];
        }
        const double kap = a.kappa ? a.kappa[e] : 1.0;
        double x[NV][D];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int d = 0; d < D; ++d) x[v][d] = L.X[D * lv[v] + d];
        double g[NV][D], V;
        simplex_gradients<D>(x, g, V);
        // H = sum_ab Gs_ab g_a g_b^T via w_a = sum_b Gs_ab g_b; H is
        // symmetric, so only its upper triangle is formed
        double Gs[NV][NV];
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) {
          Gs[aa][aa] = G[aa][aa];
#pragma unroll
          for (int b = aa + 1; b < NV; ++b) Gs[aa][b] = Gs[b][aa] = 0.5 * (G[aa][b] + G[b][aa]);
        }
        double H[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = i; j < D; ++j) H[i][j] = 0.0;
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) {
          double wv[D];
#pragma unroll
          for (int d = 0; d < D; ++d) wv[d] = Gs[aa][0] * g[0][d];
#pragma unroll
          for (int b = 1; b < NV; ++b)
#pragma unroll
            for (int d = 0; d < D; ++d) wv[d] = __fma_rn(Gs[aa][b], g[b][d], wv[d]);
#pragma unroll
          for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = i; j < D; ++j) H[i][j] = __fma_rn(g[aa][i], wv[j], H[i][j]);
        }
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < i; ++j) H[i][j] = H[j][i];
        double tr = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) tr += H[i][i];
        if (a.dkappa && lv[0] < R) a.dkappa[e] = V * tr;
        double beta = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) beta += L.gl[lv[v]];
        beta *= 1.0 / (double)NV;
        double M[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j)
            M[i][j] = V * (kap * ((i == j ? tr : 0.0) - 2.0 * H[i][j]) +
                           (i == j ? beta : 0.0));
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) {
          if (lv[aa] >= R) continue;  // halo vertex: summed by its owner
#pragma unroll
          for (int i = 0; i < D; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) s = __fma_rn(M[i][j], g[aa][j], s);
            mk = max_abs(mk, s);
            if (scatter) lds_add(L.acc + D * lv[aa] + i, to_fixed(s, sc));
          }
        }
      }
    };
    if (nrec > 0) {
      const int nchunks = (nrec + NT - 1) / NT;
      Chunk cur, nxt;
      fetch(0, cur);
      for (int c = 0; c < nchunks; c += 2) {
        fetch(c + 1, nxt);
        process(cur, c);
        if (c + 1 >= nchunks) break;
        fetch(c + 2, cur);
        process(nxt, c + 1);
      }
    }
    if (!scatter) continue;
    __syncthreads();  // every contribution has landed
    const double inv = ldexp(1.0, -pass.s);
    for (int i = tid; i < D * R; i += NT) {
      const int r = i / D, d = i - r * D;
      a.dX[d * a.nnodes + a.nodemap[node_start + r]] = from_fixed((long long)L.acc[i], inv);
    }
  }  // partitions

  if (first) fx_end<NT, false>(st, mk, 0.0, false, false, L.red);
}

template <int D, int RW, int IDB, int OFB>
void launch_adjoint(const AdjArgs& a, size_t lds, unsigned grid, hipStream_t s) {
  p1_fused_adjoint_kernel<D, RW, IDB, OFB, kAdjThreads><<<grid, kAdjThreads, lds, s>>>(a);
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" size_t femhip_fused_adjoint_lds_bytes(const femhip_fused_plan_info_t* I,
                                                 int32_t dim) {
  return I ? adjoint_lds_bytes(*I, dim) : 0;
}

extern "C" int femhip_fused_adjoint(const femhip_fused_dev_t* plan, femhip_fused_state_t* state,
                                    int32_t dim, const double* p, int64_t nnodes,
                                    const int32_t* indptr, const int32_t* diag_slot,
                                    const double* kappa, const double* grad_data,
                                    const double* grad_load, double* dkappa, double* dX,
                                    femhip_stream_t s) {
  FEMHIP_REQUIRE(plan != nullptr && state != nullptr && p != nullptr && dX != nullptr,
                 FEMHIP_EINVAL, "fused adjoint: NULL argument");
  FEMHIP_REQUIRE(reinterpret_cast<uintptr_t>(state) % 128 == 0, FEMHIP_EINVAL,
                 "fused adjoint: state must be 128-byte aligned");
  FEMHIP_REQUIRE(dim == 2 || dim == 3, FEMHIP_EUNSUPPORTED, "fused adjoint: P1 tri/tet only");
  FEMHIP_REQUIRE(nnodes > 0, FEMHIP_EINVAL, "fused adjoint: no nodes");
  FEMHIP_REQUIRE(grad_data == nullptr || (indptr != nullptr && diag_slot != nullptr),
                 FEMHIP_EINVAL, "fused adjoint: grad_data needs indptr and diag_slot");
  FEMHIP_REQUIRE(dkappa == nullptr || plan->rec_elem != nullptr, FEMHIP_EINVAL,
                 "fused adjoint: dkappa needs rec_elem");
  FEMHIP_REQUIRE(kappa == nullptr || plan->rec_elem != nullptr, FEMHIP_EINVAL,
                 "fused adjoint: kappa needs rec_elem");
  const femhip_fused_plan_info_t& I = plan->info;
  FEMHIP_REQUIRE(I.pruned == 0, FEMHIP_EUNSUPPORTED,
                 "fused adjoint: the plan's pattern must hold every element pair "
                 "(the connectivity pattern)");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (I.nparts == 0) return FEMHIP_OK;
  AdjArgs a{plan->parts, plan->nodemap, plan->recs, plan->rec_elem, indptr, diag_slot, p,
            kappa,       grad_data,     grad_load,  dkappa,          dX,     state,     nnodes,
            I.nparts,    I.max_nodes,   I.max_rows, 0};
  const size_t lds = adjoint_lds_bytes(I, dim);
  FEMHIP_REQUIRE(lds <= 160 * 1024, FEMHIP_EINVAL, "fused adjoint: plan exceeds LDS");
  for (int mode = 0; mode < 2; ++mode) {
    a.mode = mode;
    const unsigned grid = (unsigned)(mode == 0 || I.nparts < (int)kRepeatGrid ? I.nparts
                                                                               : kRepeatGrid);
    if (dim == 3 && I.rec_words == 3 && I.idbits == 12 && I.offbits == 4)
      launch_adjoint<3, 3, 12, 4>(a, lds, grid, st);
    else if (dim == 3 && I.rec_words == 4 && I.idbits == 16 && I.offbits == 5)
      launch_adjoint<3, 4, 16, 5>(a, lds, grid, st);
    else if (dim == 2 && I.rec_words == 2 && I.idbits == 12 && I.offbits == 4)
      launch_adjoint<2, 2, 12, 4>(a, lds, grid, st);
    else if (dim == 2 && I.rec_words == 3 && I.idbits == 16 && I.offbits == 5)
      launch_adjoint<2, 3, 16, 5>(a, lds, grid, st);
    else {
      set_error("fused adjoint: unsupported record format");
      return FEMHIP_EUNSUPPORTED;
    }
    FEMHIP_CHECK_LAUNCH();
  }
  return FEMHIP_OK;
}
