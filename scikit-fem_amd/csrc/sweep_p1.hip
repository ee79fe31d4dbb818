// Fused P1 assembly, layer-sweep variant (plan from sweep_plan.cpp).
//
// A workgroup owns a block of rows and walks the block's levels l = 0, 1,
// ...: LDS holds two level buffers (ring of two) of CSR row accumulators,
// row descriptors, node coordinates and load accumulators.  Step l
//   1. stages level l+1 into buffer (l+1)&1 (zeroed rows, coordinates);
//   2. computes every element whose lowest vertex level is l (one element
//      per lane, local matrix in registers) and adds its entries into the
//      rows of levels l and l+1 it owns, one colour per barrier-separated
//      round (deterministic, no atomics);
//   3. streams the now complete rows of level l to HBM.
// Records carry per vertex (level parity, window index); owned vertices are
// those whose window index is below the level's owned-row count.
#include <utility>

#include "common.h"

namespace femhip {
namespace {

struct SweepArgs {
  const int32_t* blocks;
  const int32_t* levels;
  const int32_t* nodemap;
  const int32_t* rows;
  const int32_t* chunks;
  const uint32_t* recs;
  const int32_t* rec_elem;
  const double* p;
  const double* kappa;
  double* out;
  double* load;
  int64_t nnodes;
  int32_t nblocks;
  int32_t max_slots;   // per level buffer
  int32_t max_nodes;   // per level buffer
  int32_t max_rows;    // per level buffer
};

template <int POS, int NB, int RW>
__device__ __forceinline__ uint32_t sbits(const uint32_t (&w)[RW]) {
  constexpr int wi = POS >> 5, sh = POS & 31;
  if constexpr (sh + NB <= 32) {
    return (w[wi] >> sh) & ((1u << NB) - 1u);
  } else {
    const uint64_t v = (uint64_t)w[wi] | ((uint64_t)w[wi + 1] << 32);
    return (uint32_t)(v >> sh) & ((1u << NB) - 1u);
  }
}

template <int BASE, int OFB, int RW, int... F>
__device__ __forceinline__ void sdecode(const uint32_t (&w)[RW], uint32_t* o,
                                        std::integer_sequence<int, F...>) {
  ((o[F] = sbits<BASE + F * OFB, OFB, RW>(w)), ...);
}

template <int D, int FORM>
__device__ __forceinline__ double sweep_element(const double (&x)[D + 1][D], double coef,
                                                double (&K)[D + 1][D + 1]) {
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
    if constexpr (FORM == FEMHIP_MASS) {
      const double m = fabs(det) * (1.0 / 120.0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) K[i][j] = (i == j) ? 2.0 * m : m;
    } else {
      const double s = coef / (6.0 * fabs(det));
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
          const double v = s * (n[i][0] * n[j][0] + n[i][1] * n[j][1] + n[i][2] * n[j][2]);
          K[i][j] = v;
          K[j][i] = v;
        }
    }
    return fabs(det);
  } else {
    const double c1x = x[1][0] - x[0][0], c1y = x[1][1] - x[0][1];
    const double c2x = x[2][0] - x[0][0], c2y = x[2][1] - x[0][1];
    const double det = c1x * c2y - c1y * c2x;
    double n[3][2] = {{0, 0}, {c2y, -c2x}, {-c1y, c1x}};
    n[0][0] = -(n[1][0] + n[2][0]);
    n[0][1] = -(n[1][1] + n[2][1]);
    if constexpr (FORM == FEMHIP_MASS) {
      const double m = fabs(det) * (1.0 / 24.0);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) K[i][j] = (i == j) ? 2.0 * m : m;
    } else {
      const double s = coef / (2.0 * fabs(det));
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
          const double v = s * (n[i][0] * n[j][0] + n[i][1] * n[j][1]);
          K[i][j] = v;
          K[j][i] = v;
        }
    }
    return fabs(det);
  }
}

__device__ __forceinline__ int sweep_remap(int i, int n) {
  const int q = n / 8, r = n % 8;
  const int x = i % 8, j = i / 8;
  return x * q + (x < r ? x : r) + j;
}

template <int D, int OFB, int FORM, bool LOAD, int NT>
__global__ void __launch_bounds__(NT) p1_sweep_kernel(SweepArgs a) {
  constexpr int NV = D + 1;
  constexpr int RW = (NV * 11 + NV * (NV - 1) * OFB + 4 + 31) / 32;
  constexpr uint32_t SENT = (1u << OFB) - 1u;
  constexpr int CS = (D == 3) ? 4 : 2;
  extern __shared__ double smem[];
  const int MS = (a.max_slots + 1) & ~1;
  double* acc = smem;                                  // [2][MS]
  double* X = acc + 2 * MS;                            // [2][max_nodes][CS]
  double* bacc = X + 2 * CS * a.max_nodes;             // [2][max_rows] (LOAD)
  uint32_t* rinfo = reinterpret_cast<uint32_t*>(bacc + (LOAD ? 2 * a.max_rows : 0));
  double* const dummy = acc + 2 * MS - 1;              // slot padding... see below
  const int tid = threadIdx.x;

  for (int bi = blockIdx.x; bi < a.nblocks; bi += gridDim.x) {
    const int blk = sweep_remap(bi, a.nblocks);
    const int ls = a.blocks[4 * blk], nl = a.blocks[4 * blk + 1];
    int Rr[2] = {0, 0};
    __syncthreads();
    auto prep = [&](int l) {
      const int32_t* L = a.levels + 6 * (int64_t)(ls + l);
      const int node_off = L[0], nn = L[1], row_off = L[2], nr = L[3], ns = L[4];
      const int b = l & 1;
      for (int i = tid; i < ns; i += NT) acc[b * MS + i] = 0.0;
      for (int i = tid; i < nn; i += NT) {
        const int64_t g = a.nodemap[node_off + i];
#pragma unroll
        for (int d = 0; d < D; ++d) X[(b * a.max_nodes + i) * CS + d] = a.p[d * a.nnodes + g];
      }
      for (int r = tid; r < nr; r += NT) {
        rinfo[b * a.max_rows + r] = (uint32_t)a.rows[2 * (int64_t)(row_off + r) + 1];
        if constexpr (LOAD) bacc[b * a.max_rows + r] = 0.0;
      }
      Rr[b] = nr;
    };
    prep(0);
    for (int l = 0; l < nl; ++l) {
      if (l + 1 < nl) prep(l + 1);
      else Rr[(l + 1) & 1] = 0;
      __syncthreads();
      const int c0 = a.levels[6 * (int64_t)(ls + l) + 5];
      const int c1 = a.levels[6 * (int64_t)(ls + l + 1) + 5];
      for (int c = c0; c < c1; ++c) {
        const int32_t rec0 = a.chunks[2 * (int64_t)c];
        const uint32_t cw = (uint32_t)a.chunks[2 * (int64_t)c + 1];
        const int count = (int)(cw & 0xffffu), nrounds = (int)(cw >> 16);
        const bool active = tid < count;
        int par[NV], idx[NV];
        uint32_t offs[NV * (NV - 1)];
        double K[NV][NV];
        double vol = 0.0;
        int rc = -1;
        if (active) {
          uint32_t w[RW];
          const uint32_t* rp = a.recs + (int64_t)(rec0 + tid) * RW;
#pragma unroll
          for (int k = 0; k < RW; ++k) w[k] = __builtin_nontemporal_load(rp + k);
          double x[NV][D];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            uint32_t id;
            if (v == 0) id = sbits<0, 11, RW>(w);
            else if (v == 1) id = sbits<11, 11, RW>(w);
            else if (v == 2) id = sbits<22, 11, RW>(w);
            else id = sbits<33, 11, RW>(w);
            par[v] = (int)(id >> 10);
            idx[v] = (int)(id & 1023u);
            const double* xp = X + (par[v] * a.max_nodes + idx[v]) * CS;
            const double2 xy = *reinterpret_cast<const double2*>(xp);
            x[v][0] = xy.x;
            x[v][1] = xy.y;
            if constexpr (D == 3) x[v][2] = xp[2];
          }
          sdecode<NV * 11, OFB, RW>(w, offs, std::make_integer_sequence<int, NV * (NV - 1)>{});
          rc = (int)sbits<NV * 11 + NV * (NV - 1) * OFB, 4, RW>(w);
          const double coef = (FORM == FEMHIP_KAPPA_LAPLACE) ? a.kappa[a.rec_elem[rec0 + tid]] : 1.0;
          vol = sweep_element<D, FORM>(x, coef, K);
        }
        for (int r = 0; r < nrounds; ++r) {
          __syncthreads();
          if (rc == r) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
              if (idx[i] >= Rr[par[i]]) continue;   // halo vertex: not our row
              const int b = par[i];
              if constexpr (LOAD)
                bacc[b * a.max_rows + idx[i]] += vol * (D == 3 ? 1.0 / 24.0 : 1.0 / 6.0);
              const uint32_t info = rinfo[b * a.max_rows + idx[i]];
              double* row = acc + b * MS + (info & 0xffffu);
              const uint32_t dg = (info >> 16) & 0xffu;
              double* ptr[NV];
              ptr[i] = (dg != 0xffu) ? row + dg : dummy;
              int q = 0;
#pragma unroll
              for (int j = 0; j < NV; ++j) {
                if (j == i) continue;
                const uint32_t off = offs[i * (NV - 1) + q];
                ++q;
                ptr[j] = (off != SENT) ? row + off : dummy;
              }
              double v[NV];
#pragma unroll
              for (int j = 0; j < NV; ++j) v[j] = *ptr[j];
#pragma unroll
              for (int j = 0; j < NV; ++j) *ptr[j] = v[j] + K[i][j];
            }
          }
        }
      }
      __syncthreads();
      // flush level l: 16 lanes per row
      {
        const int32_t* L = a.levels + 6 * (int64_t)(ls + l);
        const int node_off = L[0], row_off = L[2], nr = L[3];
        const int b = l & 1;
        const int lane = tid & 15, grp = tid >> 4;
        for (int r = grp; r < nr; r += NT / 16) {
          const int64_t gstart = a.rows[2 * (int64_t)(row_off + r)];
          const uint32_t info = rinfo[b * a.max_rows + r];
          const int len = (int)(info >> 24), base = (int)(info & 0xffffu);
          for (int k = lane; k < len; k += 16)
            __builtin_nontemporal_store(acc[b * MS + base + k], a.out + gstart + k);
        }
        if constexpr (LOAD)
          for (int r = tid; r < nr; r += NT) a.load[a.nodemap[node_off + r]] = bacc[b * a.max_rows + r];
      }
      __syncthreads();
    }
  }
}

template <int D, int OFB, bool LOAD, int NT>
int sweep_launch_form(const SweepArgs& a, int form, size_t lds, int groups, hipStream_t s) {
  dim3 grid((unsigned)groups), block(NT);
  switch (form) {
    case FEMHIP_LAPLACE:
      p1_sweep_kernel<D, OFB, FEMHIP_LAPLACE, LOAD, NT><<<grid, block, lds, s>>>(a);
      break;
    case FEMHIP_MASS:
      p1_sweep_kernel<D, OFB, FEMHIP_MASS, LOAD, NT><<<grid, block, lds, s>>>(a);
      break;
    case FEMHIP_KAPPA_LAPLACE:
      p1_sweep_kernel<D, OFB, FEMHIP_KAPPA_LAPLACE, LOAD, NT><<<grid, block, lds, s>>>(a);
      break;
    default:
      set_error("sweep: laplace, mass or kappa_laplace only");
      return FEMHIP_EUNSUPPORTED;
  }
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

template <bool LOAD, int NT>
int sweep_dispatch(const SweepArgs& a, int dim, int offbits, int form, size_t lds, int groups,
                   hipStream_t s) {
  if (dim == 3 && offbits == 4) return sweep_launch_form<3, 4, LOAD, NT>(a, form, lds, groups, s);
  if (dim == 3 && offbits == 5) return sweep_launch_form<3, 5, LOAD, NT>(a, form, lds, groups, s);
  if (dim == 2 && offbits == 4) return sweep_launch_form<2, 4, LOAD, NT>(a, form, lds, groups, s);
  if (dim == 2 && offbits == 5) return sweep_launch_form<2, 5, LOAD, NT>(a, form, lds, groups, s);
  set_error("sweep: unsupported record format");
  return FEMHIP_EUNSUPPORTED;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" size_t femhip_sweep_lds_bytes(const femhip_sweep_plan_info_t* I, int32_t dim,
                                         int32_t with_load) {
  if (!I) return 0;
  const size_t MS = ((size_t)I->max_level_slots + 1 + 1) & ~(size_t)1;
  return 8 * 2 * MS + 8 * 2 * (size_t)(dim == 3 ? 4 : 2) * I->max_level_nodes +
         (with_load ? 8 * 2 * (size_t)I->max_level_rows : 0) + 4 * 2 * (size_t)I->max_level_rows;
}

extern "C" int femhip_sweep_assemble(const femhip_sweep_dev_t* plan, int32_t form, int32_t dim,
                                     const double* p, int64_t nnodes, const double* kappa,
                                     double* data, double* load, int32_t groups,
                                     femhip_stream_t s) {
  FEMHIP_REQUIRE(plan != nullptr && p != nullptr && data != nullptr, FEMHIP_EINVAL,
                 "sweep: NULL argument");
  const femhip_sweep_plan_info_t& I = plan->info;
  FEMHIP_REQUIRE(I.chunk == 512 || I.chunk == 1024, FEMHIP_EINVAL,
                 "sweep: plan chunk must be 512 or 1024");
  FEMHIP_REQUIRE(form != FEMHIP_KAPPA_LAPLACE || (kappa && plan->rec_elem), FEMHIP_EINVAL,
                 "sweep: kappa_laplace needs kappa and rec_elem");
  FEMHIP_REQUIRE(groups >= 0, FEMHIP_EINVAL, "sweep: negative group count");
  if (I.nblocks == 0) return FEMHIP_OK;
  const int ngroups = (groups == 0 || groups > I.nblocks) ? I.nblocks : groups;
  // one spare slot after the level buffers receives skipped contributions
  SweepArgs a{plan->blocks, plan->levels, plan->nodemap, plan->rows, plan->chunks,
              plan->recs, plan->rec_elem, p, kappa, data, load, nnodes, I.nblocks,
              I.max_level_slots + 1, I.max_level_nodes, I.max_level_rows};
  const size_t lds = femhip_sweep_lds_bytes(&I, dim, load != nullptr);
  FEMHIP_REQUIRE(lds <= 160 * 1024, FEMHIP_EINVAL, "sweep: plan exceeds LDS");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (I.chunk == 1024) {
    if (load) return sweep_dispatch<true, 1024>(a, dim, I.offbits, form, lds, ngroups, st);
    return sweep_dispatch<false, 1024>(a, dim, I.offbits, form, lds, ngroups, st);
  }
  if (load) return sweep_dispatch<true, 512>(a, dim, I.offbits, form, lds, ngroups, st);
  return sweep_dispatch<false, 512>(a, dim, I.offbits, form, lds, ngroups, st);
}
