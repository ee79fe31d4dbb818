// Batched assembly of many per-element coefficient samples on one mesh.
//
// For a form linear in a per-element coefficient (kappa-Laplace), the CSR
// values of sample s are  out[s][slot] = sum_k kappa[s][elem[k]] * K1p[k]
// over the contributions k of the slot (seg[slot] <= k < seg[slot+1], the
// symbolic plan's fixed order), with K1p the local entries at kappa = 1
// listed in plan order and elem[k] their elements.  All samples share one
// pattern.  A workgroup owns a tile of 64 consecutive slots and kGroups
// sample groups: lanes run over slots (the contribution lists of
// neighbouring slots are neighbouring runs of K1p / elem, and the stores of
// out[s][slot..] are coalesced), each thread keeps TS sample accumulators in
// registers, and each slot is summed in plan order (deterministic).
#include "common.h"

namespace femhip {
namespace {

constexpr int kSlots = 64;   // slots per workgroup (one wave wide)
constexpr int kGroups = 1;   // sample groups per workgroup (2, 4, 16: slower)
constexpr int kTS = 16;      // samples per thread and pass

__global__ void __launch_bounds__(kSlots * kGroups)
batched_kappa_kernel(const double* __restrict__ K1p, const int32_t* __restrict__ elem,
                     const int32_t* __restrict__ seg, int64_t nseg, int64_t E,
                     const double* __restrict__ kappa, int64_t S,
                     double* __restrict__ out) {
  const int64_t slot = (int64_t)blockIdx.x * kSlots + (threadIdx.x % kSlots);
  const int group = threadIdx.x / kSlots;
  if (slot >= nseg) return;
  const int32_t b = seg[slot], en = seg[slot + 1];
  for (int64_t s0 = (int64_t)group * kTS; s0 < S; s0 += (int64_t)kGroups * kTS) {
    double acc[kTS];
    // the sample rows of this pass; past the last sample the last one is
    // read again (and not stored), so every lane issues all kTS loads
    // together instead of one conditional load at a time
    const double* ks[kTS];
#pragma unroll
    for (int j = 0; j < kTS; ++j) {
      acc[j] = 0.0;
      ks[j] = kappa + (s0 + j < S ? s0 + j : S - 1) * E;
    }
    for (int32_t k = b; k < en; ++k) {
      const double w = K1p[k];
      const int64_t e = elem[k];
      double kv[kTS];
#pragma unroll
      for (int j = 0; j < kTS; ++j) kv[j] = ks[j][e];
#pragma unroll
      for (int j = 0; j < kTS; ++j) acc[j] += kv[j] * w;
    }
#pragma unroll
    for (int j = 0; j < kTS; ++j)
      if (s0 + j < S) __builtin_nontemporal_store(acc[j], out + (s0 + j) * nseg + slot);
  }
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_batched_kappa_values(const double* K1p, const int32_t* elem,
                                           const int32_t* seg, int64_t nseg, int64_t nelems,
                                           const double* kappa, int64_t nsamples, double* out,
                                           femhip_stream_t s) {
  FEMHIP_REQUIRE(nseg >= 0 && nelems > 0 && nsamples >= 0, FEMHIP_EINVAL, "batched: bad sizes");
  if (nseg == 0 || nsamples == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(K1p && elem && seg && kappa && out, FEMHIP_EINVAL, "batched: NULL argument");
  batched_kappa_kernel<<<(unsigned)ceil_div(nseg, kSlots), kSlots * kGroups, 0,
                         reinterpret_cast<hipStream_t>(s)>>>(K1p, elem, seg, nseg, nelems,
                                                             kappa, nsamples, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
