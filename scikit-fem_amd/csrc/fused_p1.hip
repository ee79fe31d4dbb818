// Fused P1 assembly: one launch from node coordinates to CSR values.
//
// Workgroup = one row partition of the plan built by fused_plan.cpp.
//   1. zero the partition's CSR rows in LDS, stage the coordinates of its
//      local nodes (owned + halo) in LDS;
//   2. per chunk of element records: every lane decodes one record and
//      computes the element's local matrix in registers (the affine map,
//      its determinant and the cross-product gradients, FP64);
//   3. the chunk's colours are applied in rounds separated by barriers --
//      within a round no two lanes touch the same owned row, so each LDS
//      slot is summed in a fixed order (deterministic, no atomics);
//   4. the rows are written to HBM once, 16 lanes per row, with
//      non-temporal stores.
// Bytes per element visit: one 8-16 byte record; per local node: one map
// entry + its coordinates; per CSR slot: one 8-byte store.
#include <utility>

#include "common.h"

namespace femhip {
namespace {

constexpr int kFusedBlock = 512;

struct FusedArgs {
  const int32_t* parts;
  const int32_t* nodemap;
  const int32_t* rows;
  const int32_t* chunks;
  const uint32_t* recs;
  const int32_t* rec_elem;
  const double* p;
  const double* kappa;
  double* out;
  double* load;      // optional: unit-load vector, one entry per node
  int64_t nnodes;
  int32_t nparts;
  int32_t max_slots;
  int32_t max_nodes;
  int32_t max_rows;
};

template <int POS, int NB, int RW>
__device__ __forceinline__ uint32_t bits(const uint32_t (&w)[RW]) {
  static_assert(NB <= 32, "field too wide");
  constexpr int wi = POS >> 5, sh = POS & 31;
  if constexpr (sh + NB <= 32) {
    return (w[wi] >> sh) & ((1u << NB) - 1u);
  } else {
    const uint64_t v = (uint64_t)w[wi] | ((uint64_t)w[wi + 1] << 32);
    return (uint32_t)(v >> sh) & ((1u << NB) - 1u);
  }
}

template <int BASE, int OFB, int RW, int... F>
__device__ __forceinline__ void decode_offsets(const uint32_t (&w)[RW], uint32_t* o,
                                               std::integer_sequence<int, F...>) {
  ((o[F] = bits<BASE + F * OFB, OFB, RW>(w)), ...);
}

// Morton-adjacent partitions on the same XCD (workgroups are dealt to the
// eight XCDs round-robin).
__device__ __forceinline__ int xcd_remap(int i, int n) {
  const int q = n / 8, r = n % 8;
  const int x = i % 8, j = i / 8;
  return x * q + (x < r ? x : r) + j;
}

// Local matrix of a P1 simplex from its vertex coordinates in LDS: Laplace K_ab = coef * (n_a . n_b) / (d! |det|) with n_a the
// cross-product (2-D: rotated-edge) gradients scaled by det; mass
// |det| / (d! (d+1)(d+2)) * (1 + delta_ab).
template <int D, int FORM>
__device__ __forceinline__ double element_matrix(const double* X, const int (&lv)[D + 1],
                                                 double coef, double (&K)[D + 1][D + 1]) {
  // coordinates are staged AoS, 32 B (3-D: x y z pad) or 16 B per node
  double x[D + 1][D];
#pragma unroll
  for (int v = 0; v < D + 1; ++v) {
    const double2 xy = reinterpret_cast<const double2*>(X)[(D == 3 ? 2 : 1) * lv[v]];
    x[v][0] = xy.x;
    x[v][1] = xy.y;
    if constexpr (D == 3) x[v][2] = X[4 * lv[v] + 2];
  }
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

template <int D, int RW, int IDB, int OFB, int FORM, bool LOAD>
__global__ void __launch_bounds__(kFusedBlock) p1_fused_kernel(FusedArgs a) {
  constexpr int NV = D + 1;
  constexpr uint32_t SENT = (1u << OFB) - 1u;
  extern __shared__ double smem[];
  double* acc = smem;
  double* X = acc + ((a.max_slots + 3) & ~1);  // 16-byte aligned for b128 loads
  double* bacc = X + (D == 3 ? 4 : 2) * a.max_nodes;
  uint32_t* rinfo = reinterpret_cast<uint32_t*>(bacc + (LOAD ? a.max_rows : 0));

  // persistent: each workgroup walks partitions blockIdx.x, +gridDim.x, ...
  for (int pi = blockIdx.x; pi < a.nparts; pi += gridDim.x) {
  if (pi != (int)blockIdx.x) __syncthreads();  // LDS reuse
  const int part = xcd_remap(pi, a.nparts);
  const int32_t* hdr = a.parts + 8 * (int64_t)part;
  const int node_start = hdr[0], L = hdr[1], row_start = hdr[2], R = hdr[3];
  const int nslots = hdr[4], chunk_start = hdr[5], nchunks = hdr[6];
  const int tid = threadIdx.x;

  for (int i = tid; i < nslots; i += kFusedBlock) acc[i] = 0.0;
  constexpr int CS = (D == 3) ? 4 : 2;
  for (int l = tid; l < L; l += kFusedBlock) {
    const int64_t g = a.nodemap[node_start + l];
#pragma unroll
    for (int d = 0; d < D; ++d) X[CS * l + d] = a.p[d * a.nnodes + g];
  }
  for (int r = tid; r < R; r += kFusedBlock) {
    rinfo[r] = (uint32_t)a.rows[2 * (int64_t)(row_start + r) + 1];
    if constexpr (LOAD) bacc[r] = 0.0;
  }
  __syncthreads();  // staged coordinates are read before the first round

  // records of the next chunk are fetched while the current one is summed
  uint32_t wn[RW];
  int32_t rec_n = 0;
  uint32_t cw_n = 0;
  auto fetch = [&](int c) {
    rec_n = a.chunks[2 * (int64_t)(chunk_start + c)];
    cw_n = (uint32_t)a.chunks[2 * (int64_t)(chunk_start + c) + 1];
    if (tid < (int)(cw_n & 0xffffu)) {
      const uint32_t* rp = a.recs + (int64_t)(rec_n + tid) * RW;
#pragma unroll
      for (int k = 0; k < RW; ++k) wn[k] = __builtin_nontemporal_load(rp + k);
    }
  };
  if (nchunks > 0) fetch(0);
  double* const dummy = acc + a.max_slots;  // sink for skipped contributions

  for (int c = 0; c < nchunks; ++c) {
    const int32_t rec0 = rec_n;
    const uint32_t cw = cw_n;
    const int count = (int)(cw & 0xffffu), nrounds = (int)(cw >> 16);
    const bool active = tid < count;
    uint32_t w[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) w[k] = wn[k];
    if (c + 1 < nchunks) fetch(c + 1);
    int lv[NV];
    uint32_t offs[NV * (NV - 1)];
    double K[NV][NV];
    double vol = 0.0;
    int rc = -1;
    if (active) {
      lv[0] = bits<0, IDB, RW>(w);
      lv[1] = bits<IDB, IDB, RW>(w);
      lv[2] = bits<2 * IDB, IDB, RW>(w);
      if constexpr (NV == 4) lv[3] = bits<3 * IDB, IDB, RW>(w);
      rc = (int)bits<NV * IDB + NV * (NV - 1) * OFB, 4, RW>(w);
      decode_offsets<NV * IDB, OFB, RW>(w, offs, std::make_integer_sequence<int, NV * (NV - 1)>{});
      vol = element_matrix<D, FORM>(X, lv, FORM == FEMHIP_KAPPA_LAPLACE ? a.kappa[a.rec_elem[rec0 + tid]] : 1.0, K);
    }
    for (int r = 0; r < nrounds; ++r) {
      __syncthreads();
      if (rc == r) {
        // per owned vertex: its row's slots (all distinct), loads first,
        // then adds and stores; rows of halo vertices are skipped
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          if (lv[i] >= R) continue;
          // unit load: int phi_i = |det| / (d! (d+1))
          if constexpr (LOAD) bacc[lv[i]] += vol * (D == 3 ? 1.0 / 24.0 : 1.0 / 6.0);
          const uint32_t info = rinfo[lv[i]];
          double* row = acc + (info & 0xffffu);
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
  // write back: 16 lanes per row
  const int lane = tid & 15, grp = tid >> 4;
  for (int r = grp; r < R; r += kFusedBlock / 16) {
    const int64_t gstart = a.rows[2 * (int64_t)(row_start + r)];
    const uint32_t info = rinfo[r];
    const int len = (int)(info >> 24), base = (int)(info & 0xffffu);
    for (int k = lane; k < len; k += 16)
      __builtin_nontemporal_store(acc[base + k], a.out + gstart + k);
  }
  if constexpr (LOAD)
    for (int r = tid; r < R; r += kFusedBlock) a.load[a.nodemap[node_start + r]] = bacc[r];
  }  // partitions
}

template <int D, int RW, int IDB, int OFB, bool LOAD>
int launch_form(const FusedArgs& a, int form, size_t lds, int groups, hipStream_t s) {
  dim3 grid((unsigned)groups), block(kFusedBlock);
  switch (form) {
    case FEMHIP_LAPLACE:
      p1_fused_kernel<D, RW, IDB, OFB, FEMHIP_LAPLACE, LOAD><<<grid, block, lds, s>>>(a);
      break;
    case FEMHIP_MASS:
      p1_fused_kernel<D, RW, IDB, OFB, FEMHIP_MASS, LOAD><<<grid, block, lds, s>>>(a);
      break;
    case FEMHIP_KAPPA_LAPLACE:
      p1_fused_kernel<D, RW, IDB, OFB, FEMHIP_KAPPA_LAPLACE, LOAD><<<grid, block, lds, s>>>(a);
      break;
    default:
      set_error("fused path: laplace, mass or kappa_laplace only");
      return FEMHIP_EUNSUPPORTED;
  }
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

template <bool LOAD>
int dispatch_fused(const FusedArgs& a, const femhip_fused_plan_info_t& I, int form, int dim,
                   size_t lds, int groups, hipStream_t st) {
  if (dim == 3 && I.rec_words == 3 && I.idbits == 11 && I.offbits == 4)
    return launch_form<3, 3, 11, 4, LOAD>(a, form, lds, groups, st);
  if (dim == 3 && I.rec_words == 4 && I.idbits == 16 && I.offbits == 5)
    return launch_form<3, 4, 16, 5, LOAD>(a, form, lds, groups, st);
  if (dim == 2 && I.rec_words == 2 && I.idbits == 12 && I.offbits == 4)
    return launch_form<2, 2, 12, 4, LOAD>(a, form, lds, groups, st);
  if (dim == 2 && I.rec_words == 3 && I.idbits == 16 && I.offbits == 5)
    return launch_form<2, 3, 16, 5, LOAD>(a, form, lds, groups, st);
  set_error("fused: unsupported record format");
  return FEMHIP_EUNSUPPORTED;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_fused_assemble(const femhip_fused_dev_t* plan, int32_t form, int32_t dim,
                                     const double* p, int64_t nnodes, const double* kappa,
                                     double* data, double* load, int32_t groups,
                                     femhip_stream_t s) {
  FEMHIP_REQUIRE(plan != nullptr && p != nullptr && data != nullptr, FEMHIP_EINVAL,
                 "fused: NULL argument");
  const femhip_fused_plan_info_t& I = plan->info;
  FEMHIP_REQUIRE(I.block == kFusedBlock, FEMHIP_EINVAL, "fused: plan built for another block");
  FEMHIP_REQUIRE(form != FEMHIP_KAPPA_LAPLACE || (kappa && plan->rec_elem), FEMHIP_EINVAL,
                 "fused: kappa_laplace needs kappa and rec_elem");
  FEMHIP_REQUIRE(groups >= 0, FEMHIP_EINVAL, "fused: negative group count");
  if (I.nparts == 0) return FEMHIP_OK;
  const int ngroups = (groups == 0 || groups > I.nparts) ? I.nparts : groups;
  FusedArgs a{plan->parts, plan->nodemap, plan->rows, plan->chunks, plan->recs,
              plan->rec_elem, p, kappa, data, load, nnodes, I.nparts, I.max_slots,
              I.max_nodes, I.max_rows};
  const size_t lds = 8 * (((size_t)I.max_slots + 3) & ~(size_t)1) +
                     8 * (size_t)(dim == 3 ? 4 : 2) * I.max_nodes +
                     (load ? 8 * (size_t)I.max_rows : 0) + 4 * (size_t)I.max_rows;
  FEMHIP_REQUIRE(lds <= 160 * 1024, FEMHIP_EINVAL, "fused: plan exceeds LDS");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (load) return dispatch_fused<true>(a, I, form, dim, lds, ngroups, st);
  return dispatch_fused<false>(a, I, form, dim, lds, ngroups, st);
}
