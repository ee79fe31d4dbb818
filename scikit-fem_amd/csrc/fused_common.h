// Device helpers shared by the fused owner-computes kernels (fused_p1.hip:
// CSR values and load vector; fused_adjoint.hip: their adjoint): packed
// record fields, the XCD-aware partition order, fixed-point conversion and
// the launch-wide exponent protocol.
//
// Fixed point: a contribution v enters an LDS accumulator as the int64
// round(v * 2^s).  One exponent s serves a whole launch: s = 49 - ilogb(M)
// rounded down to a multiple of 4, M = the largest |contribution| of this
// input, so 2^46 <= M 2^s < 2^50.  A first pass (mode 0) uses the exponent
// the previous call measured and records the exponent field of its own M
// (atomicMax tagged with a pass epoch, so that no reset is needed); the
// repeat launch (mode 1) derives this input's exponent from it and repeats
// the pass when the two differ -- otherwise it returns at once.  No
// workgroup waits on another: the launch boundaries order the state
// updates.
#pragma once

#include <utility>

#include "common.h"

namespace femhip {
namespace {

constexpr double kFxMagic = 6755399441055744.0;  // 1.5 * 2^52
static_assert(sizeof(femhip_fused_state_t) == 256 &&
                  offsetof(femhip_fused_state_t, maxk) == 128,
              "femhip_fused_state_t layout: read line, then atomic line");
constexpr unsigned kRepeatGrid = 768;   // 3 workgroups on each of 256 CUs

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

// A field inside one word as one v_bfe_u32 (inline assembly, so that a
// following shift by 3 folds into the address add); fields that straddle
// two words use bits().
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

template <int BASE, int FB, int RW, int... F>
__device__ __forceinline__ void decode_fields(const uint32_t (&w)[RW], uint32_t* o,
                                              std::integer_sequence<int, F...>) {
  ((o[F] = field<BASE + F * FB, FB, RW>(w)), ...);
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

// round(v * scale) as int64 for |v * scale| < 2^51: the fma rounds the
// scaled value to an integer held in the low mantissa bits of the sum.
__device__ __forceinline__ unsigned long long to_fixed(double v, double scale) {
  const double y = __fma_rn(v, scale, kFxMagic);
  return (unsigned long long)(__double_as_longlong(y) - __double_as_longlong(kFxMagic));
}

__device__ __forceinline__ double from_fixed(long long v, double inv_scale) {
  return (double)v * inv_scale;
}

// The pass maxima travel as (epoch << 11) | the 11-bit exponent field of
// the double: atomicMax keeps the largest exponent of the newest epoch, so
// a new pass overrides the previous pass's maximum without a reset.
__device__ __forceinline__ unsigned long long exponent_field(double m) {
  return ((unsigned long long)__double_as_longlong(fabs(m)) >> 52) & 0x7ffull;
}

// exponent bucket of the largest |contribution| from its exponent field e:
// 2^(e-1023) <= m < 2^(e-1022), so s = 49 - (e - 1023) rounded down to a
// multiple of 4 gives 2^46 <= m 2^s < 2^50 (e = 0: zero or subnormal; e =
// 2047: not finite, flagged by the pass itself)
__device__ __forceinline__ int fx_bucket_tagged(unsigned long long tagged) {
  const int e = (int)(tagged & 0x7ffull);
  if (e == 0 || e == 0x7ff) return 0;
  int s = 49 - (e - 1023);
  s = s >= 0 ? (s & ~3) : -((-s + 3) & ~3);
  return s;
}

__device__ __forceinline__ void lds_add(unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// max(m, |v|) in one instruction (|v| as a source modifier)
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

// Partition of the k-th take of this workgroup: one per workgroup, or a
// strided walk of a persistent grid; consecutive workgroups (dealt to the
// eight XCDs round-robin) get spatially adjacent partitions on one XCD.
// Returns -1 when none is left.
__device__ __forceinline__ int next_partition(int nparts, int k) {
  const int pi = blockIdx.x + k * gridDim.x;
  return pi < nparts ? xcd_remap(pi, nparts) : -1;
}

// The exponents a pass runs with (mode 0: the previous call's; mode 1: this
// input's, from the first pass's maxima), whether it accumulates at all
// (mode 0 without a measured exponent only measures) and whether it runs
// (mode 1 only when the exponents changed).  Uniform across the launch.
struct FxPass {
  int s, sv;
  bool scatter;
  bool run;
};

template <bool LOAD>
__device__ __forceinline__ FxPass fx_begin(femhip_fused_state_t* st, int mode) {
  FxPass r;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (mode == 0) {
    // scalar registers: workgroup 0 stores to the state below, so the
    // compiler would otherwise keep these in vector registers and treat
    // every branch on them as divergent
    r.scatter = __builtin_amdgcn_readfirstlane(st->valid) != 0;
    r.s = r.scatter ? __builtin_amdgcn_readfirstlane(st->s) : 0;
    r.sv = r.scatter ? __builtin_amdgcn_readfirstlane(st->sv) : 0;
    if (lead) {
      st->s_first = r.scatter ? r.s : INT32_MIN;
      st->sv_first = r.scatter ? r.sv : INT32_MIN;
    }
    r.run = true;
  } else {
    // the first pass's maxima, final at the launch boundary; the fields
    // read here are not written by this launch
    const int s2 = fx_bucket_tagged(st->maxk), sv2 = LOAD ? fx_bucket_tagged(st->maxv) : 0;
    r.run = st->s_first == INT32_MIN || s2 != st->s_first || (LOAD && sv2 != st->sv_first);
    if (lead) {   // for the next first pass
      st->s = s2;
      if constexpr (LOAD) st->sv = sv2;
      st->valid = 1;
      st->epoch = (st->maxk >> 11) + 1;
    }
    r.scatter = true;
    r.s = s2;
    r.sv = sv2;
  }
  return r;
}

// End of a first pass: the workgroup's maxima into the launch's (read by
// the repeat launch; no fence: the kernel boundary orders them), and the
// flag of non-finite or vanishing element data.  red: 2 * NT / 64 doubles
// of LDS.
template <int NT, bool LOAD>
__device__ __forceinline__ void fx_end(femhip_fused_state_t* st, double mk, double mv, bool bad,
                                       bool flag_bad, double* red) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x;
  mk = wave_max(mk);
  if constexpr (LOAD) mv = wave_max(mv);
  const bool wbad = __any(bad);
  if ((tid & 63) == 0) {
    red[tid >> 6] = mk;
    red[NW + (tid >> 6)] = mv;
    if (wbad && flag_bad) atomicOr(&st->flags, FEMHIP_FUSED_BAD);
  }
  __syncthreads();
  if (tid == 0) {
    double MK = red[0], MV = red[NW];
#pragma unroll
    for (int q = 1; q < NW; ++q) {
      MK = fmax(MK, red[q]);
      MV = fmax(MV, red[NW + q]);
    }
    const unsigned long long ep = st->epoch << 11;
    atomicMax(reinterpret_cast<unsigned long long*>(&st->maxk), ep | exponent_field(MK));
    if constexpr (LOAD)
      atomicMax(reinterpret_cast<unsigned long long*>(&st->maxv), ep | exponent_field(MV));
  }
}

}  // namespace
}  // namespace femhip
