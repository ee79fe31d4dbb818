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
  int32_t* scale;    // fixed-point exponents per (form, partition), or NULL
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

// A field inside one word as one v_bfe_u32.  Written as inline assembly so
// that a following shift by 3 (a slot index to a byte offset) folds into
// the address add (v_lshl_add_u32) instead of being merged into a
// shift-and-mask pair; fields that straddle two words use bits().
template <int POS, int NB, int RW>
__device__ __forceinline__ uint32_t field(const uint32_t (&w)[RW]) {
  constexpr int wi = POS >> 5, sh = POS & 31;
  if constexpr (sh + NB <= 32 && sh > 0) {
    uint32_t r;
    asm("v_bfe_u32 %0, %1, %2, %3" : "=v"(r) : "v"(w[wi]), "I"(sh), "I"(NB));
    return r;
  } else {
    return bits<POS, NB, RW>(w);
  }
}

template <int BASE, int OFB, int RW, int... F>
__device__ __forceinline__ void decode_fields(const uint32_t (&w)[RW], uint32_t* o,
                                              std::integer_sequence<int, F...>) {
  ((o[F] = field<BASE + F * OFB, OFB, RW>(w)), ...);
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

// 1/d: hardware reciprocal estimate refined by two Newton steps
// r <- r + r (1 - d r) (error squared per step; ~1 ulp at the end).
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __fma_rn(-d, r, 1.0);
  r = __fma_rn(r, e, r);
  e = __fma_rn(-d, r, 1.0);
  return __fma_rn(r, e, r);
}

// Local matrix of a P1 simplex from its vertex coordinates in LDS: Laplace K_ab = coef * (n_a . n_b) / (d! |det|) with n_a the
// cross-product (2-D: rotated-edge) gradients scaled by det; mass
// |det| / (d! (d+1)(d+2)) * (1 + delta_ab).
template <int D, int FORM>
__device__ __forceinline__ double element_matrix(const double* X, const int (&lv)[D + 1],
                                                 double coef, double (&K)[D + 1][D + 1]) {
  // coordinates are staged AoS: 24 B (x y z) or 16 B (x y) per node
  double x[D + 1][D];
#pragma unroll
  for (int v = 0; v < D + 1; ++v) {
    if constexpr (D == 3) {
      x[v][0] = X[3 * lv[v]];
      x[v][1] = X[3 * lv[v] + 1];
      x[v][2] = X[3 * lv[v] + 2];
    } else {
      const double2 xy = reinterpret_cast<const double2*>(X)[lv[v]];
      x[v][0] = xy.x;
      x[v][1] = xy.y;
    }
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
    const double det = c1[0] * n[1][0] + c1[1] * n[1][1] + c1[2] * n[1][2];
    if constexpr (FORM == FEMHIP_MASS) {
      const double m = fabs(det) * (1.0 / 120.0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) K[i][j] = (i == j) ? 2.0 * m : m;
    } else {
      // vertices 1..3 directly; row and column 0 from the zero row sums
      // (n_0 = -(n_1 + n_2 + n_3))
      const double s = coef * recip(6.0 * fabs(det));
#pragma unroll
      for (int i = 1; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
          const double v = s * (n[i][0] * n[j][0] + n[i][1] * n[j][1] + n[i][2] * n[j][2]);
          K[i][j] = v;
          K[j][i] = v;
        }
#pragma unroll
      for (int j = 1; j < 4; ++j) K[0][j] = K[j][0] = -(K[j][1] + K[j][2] + K[j][3]);
      K[0][0] = -(K[0][1] + K[0][2] + K[0][3]);
    }
    return fabs(det);
  } else {
    const double c1x = x[1][0] - x[0][0], c1y = x[1][1] - x[0][1];
    const double c2x = x[2][0] - x[0][0], c2y = x[2][1] - x[0][1];
    const double det = c1x * c2y - c1y * c2x;
    const double n[3][2] = {{0, 0}, {c2y, -c2x}, {-c1y, c1x}};
    if constexpr (FORM == FEMHIP_MASS) {
      const double m = fabs(det) * (1.0 / 24.0);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) K[i][j] = (i == j) ? 2.0 * m : m;
    } else {
      // vertices 1, 2 directly; row and column 0 from the zero row sums
      const double s = coef * recip(2.0 * fabs(det));
#pragma unroll
      for (int i = 1; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
          const double v = s * (n[i][0] * n[j][0] + n[i][1] * n[j][1]);
          K[i][j] = v;
          K[j][i] = v;
        }
#pragma unroll
      for (int j = 1; j < 3; ++j) K[0][j] = K[j][0] = -(K[j][1] + K[j][2]);
      K[0][0] = -(K[0][1] + K[0][2]);
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
  double* bacc = X + D * a.max_nodes;
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
  for (int l = tid; l < L; l += kFusedBlock) {
    const int64_t g = a.nodemap[node_start + l];
#pragma unroll
    for (int d = 0; d < D; ++d) X[D * l + d] = a.p[d * a.nnodes + g];
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
  double* const dummy = acc + a.max_slots;  // sink for pruned slots

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

// ---------------------------------------------------------------------
// Fixed-point variant.  A contribution v enters the partition's LDS row
// accumulators as the integer round(v * 2^s) (int64, two's complement),
// summed with LDS integer atomics: integer addition is associative, so the
// sums -- and the values written -- do not depend on the order in which
// lanes and waves arrive (deterministic without colour rounds or barriers).
//
// s is chosen per partition (and form) from M = the largest |K_aa| over the
// partition's elements: s = 47 - ilogb(M), so every contribution
// |K_ab| <= M (the local matrices are semidefinite) is below 2^48 after
// scaling.  The rounding error of one contribution is at most 2^-s / 2 <=
// M * 2^-48, and M <= max|A| (a diagonal entry of A sums same-sign K_aa),
// so a slot with n contributions is within n * 3.6e-15 * max|A| of its
// exact sum.  The exponent of the previous call is kept in `scale`; the
// pass measures M while it accumulates and is repeated with the exponent
// of the measured M when the two differ (first call, or a changed input),
// so the values are a function of the input alone.
// ---------------------------------------------------------------------
constexpr int kFxUnknown = INT32_MIN;
constexpr double kFxMagic = 6755399441055744.0;  // 1.5 * 2^52

// round(v * scale) as int64 for |v * scale| < 2^51: the fma rounds the
// scaled value to an integer held in the low mantissa bits of the sum.
__device__ __forceinline__ unsigned long long to_fixed(double v, double scale) {
  const double y = __fma_rn(v, scale, kFxMagic);
  return (unsigned long long)(__double_as_longlong(y) - __double_as_longlong(kFxMagic));
}

__device__ __forceinline__ double from_fixed(unsigned long long v, double inv_scale) {
  return (double)(long long)v * inv_scale;
}

__device__ __forceinline__ int fx_exponent(double m) {
  if (!(m > 0.0) || !(m < INFINITY)) return 0;
  const int s = 47 - ilogb(m);
  return s > 1000 ? 1000 : (s < -1000 ? -1000 : s);
}

__device__ __forceinline__ void lds_add(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// max(m, |v|) in one instruction (the |v| as a source modifier); the
// operands are finite, so the NaN canonicalisation of fmax is not needed
__device__ __forceinline__ double max_abs(double m, double v) {
  double r;
  asm("v_max_f64 %0, %1, |%2|" : "=v"(r) : "v"(m), "v"(v));
  return r;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// One pass over a partition's element records: the local matrices, the
// largest |K_aa| and load entry (mk, mv), and -- when SCATTER -- the
// fixed-point contributions added into the LDS rows (pruned slots go to a
// per-lane sink).
// Element records are fetched kPrefetch chunks ahead of their use (deeper
// rings measured no faster on the 1M-DoF tetrahedral mesh: the element
// loop is bound by LDS and issue, not by the record loads).
#ifndef FEMHIP_FUSED_PREFETCH
#define FEMHIP_FUSED_PREFETCH 1
#endif
constexpr int kPrefetch = FEMHIP_FUSED_PREFETCH;
static_assert(kPrefetch >= 1 && kPrefetch <= 3, "record prefetch depth 1 to 3");

template <int RW>
struct RecRing {
  // depth 1 alternates between two buffers (no register copies)
  static constexpr int kSlots = kPrefetch > 1 ? kPrefetch : 2;
  uint32_t w[kSlots][RW];
  int32_t rec[kSlots];
  int count[kSlots];
  // records of chunk c of the partition's chunk list into slot S (lane tid:
  // record tid of the chunk)
  template <int S>
  __device__ __forceinline__ void load(const FusedArgs& a, int chunk_start, int c, int tid) {
    rec[S] = a.chunks[2 * (int64_t)(chunk_start + c)];
    count[S] = a.chunks[2 * (int64_t)(chunk_start + c) + 1] & 0xffff;
    if (tid < count[S]) {
      const uint32_t* rp = a.recs + (int64_t)(rec[S] + tid) * RW;
#pragma unroll
      for (int k = 0; k < RW; ++k) w[S][k] = __builtin_nontemporal_load(rp + k);
    }
  }
  // chunks 0 .. kPrefetch-1
  __device__ __forceinline__ void prime(const FusedArgs& a, int chunk_start, int nchunks,
                                        int tid) {
    if (nchunks > 0) load<0>(a, chunk_start, 0, tid);
    if constexpr (kPrefetch > 1)
      if (nchunks > 1) load<(kPrefetch > 1 ? 1 : 0)>(a, chunk_start, 1, tid);
    if constexpr (kPrefetch > 2)
      if (nchunks > 2) load<kPrefetch - 1>(a, chunk_start, 2, tid);
  }
};

// One pass over a partition's element records: the local matrices, the
// largest |K_aa| and load entry (mk, mv), and -- when SCATTER -- the
// fixed-point contributions added into the LDS rows.  `ring` holds chunks
// 0..kPrefetch-1 on entry.
template <int D, int RW, int IDB, int OFB, int FORM, bool LOAD, bool SCATTER>
__device__ __forceinline__ void fx_pass(const FusedArgs& a, int chunk_start, int nchunks,
                                        int R, const double* X, const uint32_t* rinfo,
                                        unsigned long long* acc, unsigned long long* bacc,
                                        double sck, double scv, RecRing<RW>& ring,
                                        double& mk, double& mv) {
  constexpr int NV = D + 1;
  constexpr double LOADC = (D == 3) ? 1.0 / 24.0 : 1.0 / 6.0;  // int phi = |det| LOADC
  const int tid = threadIdx.x;

  auto element = [&](const uint32_t (&w)[RW], int32_t rec0) {
    int lv[NV];
    uint32_t offs[NV * (NV - 1)];
    double K[NV][NV];
    lv[0] = bits<0, IDB, RW>(w);
    lv[1] = bits<IDB, IDB, RW>(w);
    lv[2] = bits<2 * IDB, IDB, RW>(w);
    if constexpr (NV == 4) lv[3] = bits<3 * IDB, IDB, RW>(w);
    decode_fields<NV * IDB, OFB, RW>(w, offs, std::make_integer_sequence<int, NV * (NV - 1)>{});
    const double vol = element_matrix<D, FORM>(
        X, lv, FORM == FEMHIP_KAPPA_LAPLACE ? a.kappa[a.rec_elem[rec0 + tid]] : 1.0, K);
#pragma unroll
    for (int i = 0; i < NV; ++i) mk = max_abs(mk, K[i][i]);
    if constexpr (LOAD) mv = fmax(mv, vol * LOADC);
    if constexpr (SCATTER) {
      unsigned long long F[NV][NV];
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = i; j < NV; ++j) F[i][j] = F[j][i] = to_fixed(K[i][j], sck);
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if (lv[i] >= R) continue;  // halo vertex: its row belongs elsewhere
        // pruned entries (a pruned diagonal too) carry the offset of the
        // row's spare slot
        const uint32_t info = rinfo[lv[i]];
        unsigned long long* row = acc + (info & 0xffffu);
        lds_add(row + ((info >> 16) & 0xffu), F[i][i]);
        int q = 0;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          if (j == i) continue;
          lds_add(row + offs[i * (NV - 1) + q], F[i][j]);
          ++q;
        }
        if constexpr (LOAD) lds_add(bacc + lv[i], to_fixed(vol * LOADC, scv));
      }
    }
  };
  // chunk c uses ring slot c % kPrefetch, which is refilled with chunk
  // c + kPrefetch before the element work of chunk c
  auto step = [&](auto slot, int c) {
    constexpr int S = decltype(slot)::value;
    if (c >= nchunks) return;
    uint32_t w[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) w[k] = ring.w[S][k];
    const int32_t rec0 = ring.rec[S];
    const bool active = tid < ring.count[S];
    if (c + kPrefetch < nchunks) ring.template load<S>(a, chunk_start, c + kPrefetch, tid);
    if (active) element(w, rec0);
  };
  if constexpr (kPrefetch == 1) {
    // chunk c + 1 loads into one buffer while chunk c is processed from the
    // other, which then receives chunk c + 2
    for (int c = 0; c < nchunks; c += 2) {
      if (c + 1 < nchunks) ring.template load<1>(a, chunk_start, c + 1, tid);
      if (tid < ring.count[0]) element(ring.w[0], ring.rec[0]);
      if (c + 1 >= nchunks) break;
      if (c + 2 < nchunks) ring.template load<0>(a, chunk_start, c + 2, tid);
      if (tid < ring.count[1]) element(ring.w[1], ring.rec[1]);
    }
  } else {
    for (int c = 0; c < nchunks; c += kPrefetch) {
      step(std::integral_constant<int, 0>{}, c);
      step(std::integral_constant<int, (kPrefetch > 1 ? 1 : 0)>{}, c + 1);
      if constexpr (kPrefetch > 2) step(std::integral_constant<int, kPrefetch - 1>{}, c + 2);
    }
  }
}

// Fixed-point kernel: one workgroup per partition by default (a smaller
// grid walks the partitions persistently).
// NT = 512: three workgroups per CU without the load vector (at most 80
// VGPRs); NT = 1024: two workgroups per CU (at most 64 VGPRs)
template <int D, int RW, int IDB, int OFB, int FORM, bool LOAD, int NT>
__global__ void __launch_bounds__(NT)
    __attribute__((amdgpu_waves_per_eu(LOAD ? 4 : (NT == 1024 ? 8 : 6))))
p1_fused_fx_kernel(FusedArgs a) {
  constexpr int NW = NT / 64;
  constexpr int FSLOT = FORM == FEMHIP_LAPLACE ? 0 : (FORM == FEMHIP_MASS ? 1 : 2);
  extern __shared__ double smem[];
  // [max_slots accumulators | coordinates | load acc | 2*NW reduction |
  //  row info | row starts]
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(smem);
  double* X = smem + ((a.max_slots + 1) & ~1);
  unsigned long long* bacc = reinterpret_cast<unsigned long long*>(X + D * a.max_nodes);
  double* red = reinterpret_cast<double*>(bacc + (LOAD ? a.max_rows : 0));
  uint32_t* rinfo = reinterpret_cast<uint32_t*>(red + 2 * NW);
  int32_t* rstart = reinterpret_cast<int32_t*>(rinfo + a.max_rows);  // CSR row starts
  const int tid = threadIdx.x;

  for (int pi = blockIdx.x; pi < a.nparts; pi += gridDim.x) {
    if (pi != (int)blockIdx.x) __syncthreads();  // LDS reuse
    const int part = xcd_remap(pi, a.nparts);
    const int32_t* hdr = a.parts + 8 * (int64_t)part;
    const int node_start = hdr[0], L = hdr[1], row_start = hdr[2], R = hdr[3];
    const int nslots = hdr[4], chunk_start = hdr[5], nchunks = hdr[6];
    int32_t* sc = a.scale + 2 * ((int64_t)FSLOT * a.nparts + part);
    int s = sc[0], sv = sc[1];
    bool known = s != kFxUnknown && (!LOAD || sv != kFxUnknown);

    for (int l = tid; l < L; l += NT) {
      const int64_t g = a.nodemap[node_start + l];
#pragma unroll
      for (int d = 0; d < D; ++d) X[D * l + d] = a.p[d * a.nnodes + g];
    }
    for (int r = tid; r < R; r += NT) {
      const int2 rw = reinterpret_cast<const int2*>(a.rows)[row_start + r];
      rstart[r] = rw.x;
      rinfo[r] = (uint32_t)rw.y;
    }
    RecRing<RW> ring;
    ring.prime(a, chunk_start, nchunks, tid);  // overlaps the staging

    for (;;) {
      if (known) {
        for (int i = tid; i < nslots; i += NT) acc[i] = 0ull;
        if constexpr (LOAD)
          for (int r = tid; r < R; r += NT) bacc[r] = 0ull;
      }
      __syncthreads();  // staging / zeroing before the first contribution
      double mk = 0.0, mv = 0.0;
      if (known)
        fx_pass<D, RW, IDB, OFB, FORM, LOAD, true>(a, chunk_start, nchunks, R, X, rinfo, acc,
                                                   bacc, ldexp(1.0, s), ldexp(1.0, sv), ring,
                                                   mk, mv);
      else
        fx_pass<D, RW, IDB, OFB, FORM, LOAD, false>(a, chunk_start, nchunks, R, X, rinfo, acc,
                                                    bacc, 1.0, 1.0, ring, mk, mv);
      // partition maxima -> the exponents this pass should have used
      mk = wave_max(mk);
      if constexpr (LOAD) mv = wave_max(mv);
      if ((tid & 63) == 0) {
        red[tid >> 6] = mk;
        red[NW + (tid >> 6)] = mv;
      }
      __syncthreads();  // also: every contribution of the pass has landed
      double MK = red[0], MV = red[NW];
#pragma unroll
      for (int k = 1; k < NW; ++k) {
        MK = fmax(MK, red[k]);
        MV = fmax(MV, red[NW + k]);
      }
      const int s2 = fx_exponent(MK), sv2 = LOAD ? fx_exponent(MV) : sv;
      if (known && s2 == s && sv2 == sv) break;
      s = s2;
      sv = sv2;
      known = true;
      ring.prime(a, chunk_start, nchunks, tid);
      __syncthreads();  // reductions read before the next pass reuses red
    }
    // write back: 16 lanes per row
    const double inv = ldexp(1.0, -s);
    const int lane = tid & 15, grp = tid >> 4;
    for (int r = grp; r < R; r += NT / 16) {
      const int64_t gstart = rstart[r];
      const uint32_t info = rinfo[r];
      const int len = (int)(info >> 24), base = (int)(info & 0xffffu);
      for (int k = lane; k < len; k += 16)
        __builtin_nontemporal_store(from_fixed(acc[base + k], inv), a.out + gstart + k);
    }
    if constexpr (LOAD) {
      const double invv = ldexp(1.0, -sv);
      for (int r = tid; r < R; r += NT)
        a.load[a.nodemap[node_start + r]] = from_fixed(bacc[r], invv);
    }
    if (tid == 0) {
      sc[0] = s;
      sc[1] = sv;
    }
  }  // partitions
}


// groups == 0: one workgroup per partition; fewer: persistent workgroups
template <int D, int RW, int IDB, int OFB, int FORM, bool LOAD>
void launch_one(const FusedArgs& a, int nt, size_t lds, int groups, hipStream_t s) {
  if (a.scale == nullptr) {
    const int g = (groups == 0 || groups > a.nparts) ? a.nparts : groups;
    p1_fused_kernel<D, RW, IDB, OFB, FORM, LOAD><<<dim3(g), dim3(kFusedBlock), lds, s>>>(a);
    return;
  }
  auto launch = [&](auto kernel, int threads) {
    const int g = (groups == 0 || groups > a.nparts) ? a.nparts : groups;
    kernel<<<dim3(g), dim3(threads), lds, s>>>(a);
  };
  if (nt == 1024)
    launch(p1_fused_fx_kernel<D, RW, IDB, OFB, FORM, LOAD, 1024>, 1024);
  else
    launch(p1_fused_fx_kernel<D, RW, IDB, OFB, FORM, LOAD, 512>, 512);
}

template <int D, int RW, int IDB, int OFB, bool LOAD>
int launch_form(const FusedArgs& a, int form, int nt, size_t lds, int groups, hipStream_t s) {
  switch (form) {
    case FEMHIP_LAPLACE:
      launch_one<D, RW, IDB, OFB, FEMHIP_LAPLACE, LOAD>(a, nt, lds, groups, s);
      break;
    case FEMHIP_MASS:
      launch_one<D, RW, IDB, OFB, FEMHIP_MASS, LOAD>(a, nt, lds, groups, s);
      break;
    case FEMHIP_KAPPA_LAPLACE:
      launch_one<D, RW, IDB, OFB, FEMHIP_KAPPA_LAPLACE, LOAD>(a, nt, lds, groups, s);
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
    return launch_form<3, 3, 11, 4, LOAD>(a, form, I.block, lds, groups, st);
  if (dim == 3 && I.rec_words == 4 && I.idbits == 16 && I.offbits == 5)
    return launch_form<3, 4, 16, 5, LOAD>(a, form, I.block, lds, groups, st);
  if (dim == 2 && I.rec_words == 2 && I.idbits == 12 && I.offbits == 4)
    return launch_form<2, 2, 12, 4, LOAD>(a, form, I.block, lds, groups, st);
  if (dim == 2 && I.rec_words == 3 && I.idbits == 16 && I.offbits == 5)
    return launch_form<2, 3, 16, 5, LOAD>(a, form, I.block, lds, groups, st);
  set_error("fused: unsupported record format");
  return FEMHIP_EUNSUPPORTED;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" size_t femhip_fused_lds_bytes(const femhip_fused_plan_info_t* I, int32_t dim,
                                         int32_t with_load, int32_t fixed_point) {
  if (!I) return 0;
  const size_t coords = 8 * (size_t)dim * I->max_nodes;
  const size_t load_acc = with_load ? 8 * (size_t)I->max_rows : 0;
  const size_t rinfo = 4 * (size_t)I->max_rows;
  if (fixed_point)  // accumulators, coordinates, load, reduction, row info + starts
    return 8 * (((size_t)I->max_slots + 1) & ~(size_t)1) + coords + load_acc +
           8 * 2 * (I->block / 64) + 2 * rinfo;
  return 8 * (((size_t)I->max_slots + 3) & ~(size_t)1) + coords + load_acc + rinfo;
}

extern "C" int femhip_fused_assemble(const femhip_fused_dev_t* plan, int32_t form, int32_t dim,
                                     const double* p, int64_t nnodes, const double* kappa,
                                     double* data, double* load, int32_t groups,
                                     femhip_stream_t s) {
  FEMHIP_REQUIRE(plan != nullptr && p != nullptr && data != nullptr, FEMHIP_EINVAL,
                 "fused: NULL argument");
  const femhip_fused_plan_info_t& I = plan->info;
  FEMHIP_REQUIRE(I.block == kFusedBlock || (plan->scale != nullptr && I.block == 1024),
                 FEMHIP_EINVAL,
                 "fused: plan block must be 512 (or 1024 with fixed-point accumulation)");
  FEMHIP_REQUIRE(form != FEMHIP_KAPPA_LAPLACE || (kappa && plan->rec_elem), FEMHIP_EINVAL,
                 "fused: kappa_laplace needs kappa and rec_elem");
  FEMHIP_REQUIRE(groups >= 0, FEMHIP_EINVAL, "fused: negative group count");
  if (I.nparts == 0) return FEMHIP_OK;
  const int ngroups = groups;  // resolved per kernel in launch_one
  FusedArgs a{plan->parts, plan->nodemap, plan->rows, plan->chunks, plan->recs,
              plan->rec_elem, p, kappa, data, load, plan->scale, nnodes, I.nparts,
              I.max_slots, I.max_nodes, I.max_rows};
  const size_t lds = femhip_fused_lds_bytes(&I, dim, load != nullptr, plan->scale != nullptr);
  FEMHIP_REQUIRE(lds <= 160 * 1024, FEMHIP_EINVAL, "fused: plan exceeds LDS");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (load) return dispatch_fused<true>(a, I, form, dim, lds, ngroups, st);
  return dispatch_fused<false>(a, I, form, dim, lds, ngroups, st);
}
