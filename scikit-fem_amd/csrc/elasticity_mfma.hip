// Local matrices of linear elasticity on ElementVector(ElementHex2) with
// FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// With g^a_k(q) the global gradients of the 27 scalar Q2 functions and
// omega_q = w_q |det J_q|, the 81 x 81 local matrix of
// ddot(C(sym_grad u), sym_grad v) (models/elasticity.py:35-53) for test
// function (a, c) and trial function (b, d) is
//     K[(a,c),(b,d)] = lam S_cd[a,b] + mu S_dc[a,b] + mu delta_cd tr(S)[a,b]
// with S_kl = G_k^T diag(omega) G_l, G_k = [g^a_k(q)] (27 points x 27
// functions).  The six products S_kl (k <= l; S_lk = S_kl^T) are 27 x 27 x 27
// GEMMs, padded to 32 x 32 x 28 and computed as 16 x 16 tiles on MFMA; the
// gradient tables are built per element in LDS from the mapping (invDF,
// |detDF| per point, from femhip_geometry) and the reference gradients.
//
// One workgroup (4 waves) per element at a time (grid-stride over
// elements, 32,768 workgroups by default); each wave owns 6 of the 24
// output tiles.  A thread keeps the reference gradients of its table
// entries in registers for the whole launch and loads the mapping of the
// workgroup's next element while the current one is assembled; the kernel is
// held to 128 registers, so that the four workgroups LDS allows per CU stay
// resident.  The six 27 x 27
// products share LDS with the operand tables they are computed from, and the
// reference gradients are read through the cache, so a workgroup holds
// 39.2 KB and four fit on a CU: while some assemble their local matrices,
// others run their products.  The local matrix is written
// element-major and row-major, out[e][i*81 + j] (i test, j trial),
// contiguous per element.
#include "common.h"

namespace femhip {
namespace {

constexpr int kNB = 27;    // scalar Q2 functions
constexpr int kQ = 27;     // 3 x 3 x 3 Gauss points
constexpr int kQP = 28;    // points padded to a multiple of 4 (MFMA k)
constexpr int kBP = 32;    // functions padded to two 16-wide tiles
constexpr int kThreads = 256;

typedef double double4_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
hex2_elasticity_mfma_kernel(const double* __restrict__ G, const double* __restrict__ W,
                            const double* __restrict__ dphi, int64_t E, double lam,
                            double mu, double* __restrict__ out) {
  __shared__ double s_map[kQ * 10];                 // invDF (9) + omega per point
  // The operand tables, then (same memory) the products.  The tables hold
  // the 27 functions only: a tile's rows 27..31 of A read the next rows of
  // the buffer and its columns 27..31 of B the next row's first entries (or
  // the zeroed tail), which reach only the discarded rows / columns of the
  // product; the padded point q = 27 is an explicit zero in both tables.
  // The rows of G_k^T are kQA = 29 doubles apart: with an odd stride the
  // 16 rows that one group of 16 lanes fetches for an MFMA operand fall in
  // distinct LDS banks (with 28 they met four to a bank).
  constexpr int kQA = kQP + 1;
  constexpr int kTabA = 3 * kNB * kQA;
  constexpr int kTabB = 3 * kNB * kQP;
  constexpr int kTail = 8;
  constexpr int kBuf = kTabA + kTabB + kTail;
  static_assert(6 * kNB * kNB <= kBuf, "products fit the table memory");
  static_assert((2 * kNB + 31) * kQA + kQP <= kBuf, "padded A rows stay in the buffer");
  __shared__ double s_buf[kBuf];
  double(*s_ga)[kNB][kQA] = reinterpret_cast<double(*)[kNB][kQA]>(s_buf);  // G_k^T
  double(*s_gb)[kQP][kNB] =
      reinterpret_cast<double(*)[kQP][kNB]>(s_buf + kTabA);  // diag(omega) G_l
  double(*s_S)[kNB][kNB] = reinterpret_cast<double(*)[kNB][kNB]>(s_buf);  // S_kl, k <= l
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t EQ = E * kQ;
  if (tid < kTail) s_buf[kTabA + kTabB + tid] = 0.0;


  // The table build gives thread t the (function a, point q) pairs
  // i = t + 256 it, a = i / 28, q = i % 28, for every element: their
  // reference gradients stay in registers for the whole launch.
  constexpr int kPairs = kNB * kQP;
  constexpr int kPairIt = (kPairs + kThreads - 1) / kThreads;
  double dref[kPairIt][3];
#pragma unroll
  for (int it = 0; it < kPairIt; ++it) {
    const int i = tid + it * kThreads;
    const int a = i / kQP, q = i % kQP;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      dref[it][d] = (i < kPairs && q < kQ) ? dphi[(a * 3 + d) * kQ + q] : 0.0;
  }
  // the mapping of the workgroup's next element (invDF, then w_q |det J|),
  // loaded while the current one is assembled and written
  constexpr int kMapIt = (kQ * 10 + kThreads - 1) / kThreads;
  double mnext[kMapIt];
  auto load_map = [&](int64_t e) {
#pragma unroll
    for (int it = 0; it < kMapIt; ++it) {
      const int i = tid + it * kThreads;
      mnext[it] = 0.0;
      if (i < kQ * 10 && e < E) {
        const int q = i / 10, c = i % 10;
        double v = G[(int64_t)c * EQ + e * kQ + q];
        if (c == 9) v *= W[q];
        mnext[it] = v;
      }
    }
  };
  load_map(blockIdx.x);

  for (int64_t e = blockIdx.x; e < E; e += gridDim.x) {
    __syncthreads();   // the previous element's products have been written out
#pragma unroll
    for (int it = 0; it < kMapIt; ++it) {
      const int i = tid + it * kThreads;
      if (i < kQ * 10) s_map[i] = mnext[it];
    }
    __syncthreads();
    load_map(e + gridDim.x);
    // global gradients g^a_k(q) = sum_i invDF[i][k] dphi[a][i][q], and the
    // zeros of the padded point q = 27
#pragma unroll
    for (int it = 0; it < kPairIt; ++it) {
      const int i = tid + it * kThreads;
      if (i < kPairs) {
        const int a = i / kQP, q = i % kQP;
        const double* m = s_map + (q < kQ ? q : 0) * 10;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double g = 0.0, gw = 0.0;
          if (q < kQ) {
            g = m[0 * 3 + k] * dref[it][0] + m[1 * 3 + k] * dref[it][1] +
                m[2 * 3 + k] * dref[it][2];
            gw = g * m[9];
          }
          s_ga[k][a][q] = g;
          s_gb[k][q][a] = gw;
        }
      }
    }
    __syncthreads();
    // 6 products x 4 tiles = 24 tiles, 6 per wave, kept in registers until
    // every wave has read the operand tables
    double4_t acc[6];
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const int tile = wave + 4 * t;
      const int kl = tile >> 2, ti = (tile >> 1) & 1, tj = tile & 1;
      const int k = kl < 3 ? kl : (kl == 3 ? 0 : (kl == 4 ? 0 : 1));
      const int l = kl < 3 ? kl : (kl == 3 ? 1 : 2);
      acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q0 = 0; q0 < kQP; q0 += 4) {
        const double a = s_buf[(k * kNB + ti * 16 + r16) * kQA + q0 + k4];            // A[i][kk]
        const double b = s_buf[kTabA + (l * kQP + q0 + k4) * kNB + tj * 16 + r16];   // B[kk][j]
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
    // v_mfma_f64_16x16x4f64 output: register r of lane l is D[4r + l/16][l%16]
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const int tile = wave + 4 * t;
      const int kl = tile >> 2, ti = (tile >> 1) & 1, tj = tile & 1;
      const int col = tj * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = ti * 16 + 4 * r + k4;
        if (row < kNB && col < kNB) s_S[kl][row][col] = acc[t][r];
      }
    }
    __syncthreads();
    // 81 x 81 entries, element-major and row-major: index i*81 + j (i test,
    // j trial), so that what one element adds to one global row is adjacent.
    // Thread t < 243 writes entry t of rows 3a .. 3a+2 for a = 0..26: its
    // components c = t / 81 (test) and d = (t % 81) % 3 (trial) and trial
    // function b = (t % 81) / 3 stay fixed, the test function a advances.
    if (tid < 3 * 81) {
      const int c = tid / 81, j = tid - 81 * c, b = j / 3, d = j - 3 * b;
      // S_pr[a][b] for p <= r, S_rp[b][a] = S_pr[a][b] otherwise: the address
      // of S_pr[a][b] is base + a * step
      auto locate = [&](int p, int r, const double*& base, int& step) {
        const int lo = p < r ? p : r, hi = p < r ? r : p;
        const int kl = lo == hi ? lo : (lo == 0 ? (hi == 1 ? 3 : 4) : 5);
        if (p <= r) {
          base = &s_S[kl][0][b];
          step = kNB;
        } else {
          base = &s_S[kl][b][0];
          step = 1;
        }
      };
      const double *b1, *b2;
      int s1, s2;
      locate(c, d, b1, s1);
      locate(d, c, b2, s2);
      double* o = out + e * (int64_t)(81 * 81) + tid;
#pragma unroll 3
      for (int a = 0; a < kNB; ++a) {
        double v = lam * b1[a * s1] + mu * b2[a * s2];
        if (c == d) v += mu * (s_S[0][a][b] + s_S[1][a][b] + s_S[2][a][b]);
        __builtin_nontemporal_store(v, o + a * (3 * 81));
      }
    }
  }
}

// Division of n < 2^31 by a divisor d fixed for a launch, as a multiply and a
// shift (Granlund & Montgomery, "Division by invariant integers using
// multiplication", 1994): with 2^s >= d and m = ceil(2^(31+s) / d),
// floor(n m / 2^(31+s)) = floor(n / d), the error n (m d - 2^(31+s)) /
// (d 2^(31+s)) being below 2^-s <= 1/d; n m < 2^63.
struct Div31 {
  uint64_t m;
  uint32_t shift;
  static Div31 make(uint32_t d) {
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    return Div31{((1ull << (31 + s)) + d - 1) / d, 31 + s};
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    return (uint32_t)(((uint64_t)n * m) >> shift);
  }
};

// out[k] = sum of the plan's contributions c = perm[j], j in [seg[k], seg[k+1]),
// in that order; contribution c = (tj*nb + ti)*E + e is entry (ti test, tj
// trial) of element e, read from the row-major local matrices
// loc[e*nb*nb + ti*nb + tj].  The slots of one global row then gather from the
// same few cache lines of each element; up to four gathers are in flight.
__global__ void __launch_bounds__(256)
segment_sum_em_kernel(const double* __restrict__ loc, const int32_t* __restrict__ perm,
                      const int32_t* __restrict__ seg, int64_t nseg, uint32_t E, uint32_t nb,
                      Div31 byE, Div31 byNB, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= nseg) return;
  auto at = [&](int32_t c) -> double {
    const uint32_t cu = (uint32_t)c, ent = byE.div(cu), e = cu - ent * E;
    const uint32_t tj = byNB.div(ent), ti = ent - tj * nb;
    return loc[(int64_t)e * (nb * nb) + ti * nb + tj];
  };
  const int32_t b = seg[k], en = seg[k + 1];
  double acc = 0.0;
  int32_t j = b;
  for (; j + 4 <= en; j += 4) {
    const int32_t p0 = perm[j], p1 = perm[j + 1], p2 = perm[j + 2], p3 = perm[j + 3];
    const double v0 = at(p0), v1 = at(p1), v2 = at(p2), v3 = at(p3);
    acc += v0;
    acc += v1;
    acc += v2;
    acc += v3;
  }
  const int r = en - j;
  if (r > 0) {
    const int32_t p0 = perm[j];
    const int32_t p1 = r > 1 ? perm[j + 1] : p0;
    const int32_t p2 = r > 2 ? perm[j + 2] : p0;
    const double v0 = at(p0), v1 = at(p1), v2 = at(p2);
    acc += v0;
    if (r > 1) acc += v1;
    if (r > 2) acc += v2;
  }
  out[k] = acc;
}

// Row sums of element-major, row-major local matrices.  pos[lr*nb + j]
// (lr = e*nb + i, the local row of element e whose DOF is global row r) is
// the offset of entry (i, j) in row r of the CSR pattern, -1 for a
// contribution the pattern prunes.  One thread per row writes the offsets
// of its slots' contributions (c = (j*nb + i)*E + e in the plan).
__global__ void __launch_bounds__(256)
row_sum_pos_kernel(const int32_t* __restrict__ perm, const int32_t* __restrict__ seg,
                   const int32_t* __restrict__ indptr, int64_t nrows, uint32_t E, uint32_t nb,
                   Div31 byE, Div31 byNB, int16_t* __restrict__ pos) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= nrows) return;
  const int32_t k0 = indptr[r], k1 = indptr[r + 1];
  for (int32_t k = k0; k < k1; ++k)
    for (int32_t q = seg[k]; q < seg[k + 1]; ++q) {
      const uint32_t cu = (uint32_t)perm[q], ent = byE.div(cu), e = cu - ent * E;
      const uint32_t tj = byNB.div(ent), ti = ent - tj * nb;
      pos[((int64_t)e * nb + ti) * nb + tj] = (int16_t)(k - k0);
    }
}

// U local rows of one global row added into the row's LDS slots: their
// loads issued together, then the adds.  LDS adds without a return value:
// no wait between the local rows; one wave's adds reach a slot in
// instruction order (u), and no two lanes of one instruction share a slot,
// so every slot is summed in the same order on every call.
template <int U>
__device__ __forceinline__ void add_local_rows(const double* __restrict__ loc,
                                               const int16_t* __restrict__ pos,
                                               const int32_t* __restrict__ lrs, int nb,
                                               int lane, double* a) {
  const int j0 = min(lane, nb - 1), j1 = min(lane + 64, nb - 1);
  double v0[U], v1[U];
  int p0[U], p1[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t base = (int64_t)lrs[u] * nb;
    v0[u] = loc[base + j0];
    p0[u] = pos[base + j0];
    v1[u] = loc[base + j1];
    p1[u] = pos[base + j1];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (lane < nb && p0[u] >= 0)
      __hip_atomic_fetch_add(a + p0[u], v0[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (lane + 64 < nb && p1[u] >= 0)
      __hip_atomic_fetch_add(a + p1[u], v1[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// One wave per global row r: the row's slots accumulate in LDS while the
// wave walks the local rows of r (rowe[rowe_ptr[r] .. rowe_ptr[r+1]), in
// increasing element order), each read as nb contiguous values with their
// offsets; the row is then written once, coalesced.  The row index is
// uniform in the wave, so its bounds come from scalar loads and the local
// rows go eight at a time through a body unrolled for exactly their count
// (a row of the vector Q2 pattern has 1, 2, 4 or 8).
constexpr int kRowWaves = 4;
constexpr int kMaxRowLen = 384;
__global__ void __launch_bounds__(64 * kRowWaves)
row_sum_em_kernel(const double* __restrict__ loc, const int16_t* __restrict__ pos,
                  const int32_t* __restrict__ rowe_ptr, const int32_t* __restrict__ rowe,
                  const int32_t* __restrict__ indptr, int64_t nrows, int nb,
                  double* __restrict__ out) {
  __shared__ double acc[kRowWaves][kMaxRowLen];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * kRowWaves + w;
  if (r >= nrows) return;   // a whole wave: no barrier below
  double* a = acc[w];
  const int32_t k0 = indptr[r], len = indptr[r + 1] - k0;
  for (int q = lane; q < len; q += 64) a[q] = 0.0;
  const int32_t t0 = rowe_ptr[r], nt = rowe_ptr[r + 1] - t0;
  for (int c = 0; c < nt; c += 8) {
    const int32_t* lrs = rowe + t0 + c;
    switch (min(nt - c, 8)) {
      case 1: add_local_rows<1>(loc, pos, lrs, nb, lane, a); break;
      case 2: add_local_rows<2>(loc, pos, lrs, nb, lane, a); break;
      case 3: add_local_rows<3>(loc, pos, lrs, nb, lane, a); break;
      case 4: add_local_rows<4>(loc, pos, lrs, nb, lane, a); break;
      case 5: add_local_rows<5>(loc, pos, lrs, nb, lane, a); break;
      case 6: add_local_rows<6>(loc, pos, lrs, nb, lane, a); break;
      case 7: add_local_rows<7>(loc, pos, lrs, nb, lane, a); break;
      default: add_local_rows<8>(loc, pos, lrs, nb, lane, a); break;
    }
  }
  for (int q = lane; q < len; q += 64) out[k0 + q] = a[q];
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_row_sum_em_plan(const int32_t* perm, const int32_t* seg,
                                      const int32_t* indptr, int64_t nrows, int64_t nelems,
                                      int64_t nb, int16_t* pos, femhip_stream_t s) {
  FEMHIP_REQUIRE(nrows >= 0 && nelems > 0 && nb > 0, FEMHIP_EINVAL, "row sums: bad sizes");
  FEMHIP_REQUIRE(nelems * nb * nb <= INT32_MAX, FEMHIP_EOVERFLOW,
                 "row sums: contributions exceed int32 indices");
  if (nrows == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(perm && seg && indptr && pos, FEMHIP_EINVAL, "row sums: NULL argument");
  row_sum_pos_kernel<<<(unsigned)ceil_div(nrows, 256), 256, 0,
                       reinterpret_cast<hipStream_t>(s)>>>(
      perm, seg, indptr, nrows, (uint32_t)nelems, (uint32_t)nb, Div31::make((uint32_t)nelems),
      Div31::make((uint32_t)nb), pos);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_row_sum_em(const double* loc, const int16_t* pos, const int32_t* rowe_ptr,
                                 const int32_t* rowe, const int32_t* indptr, int64_t nrows,
                                 int64_t nb, int32_t max_row_len, double* out, femhip_stream_t s) {
  FEMHIP_REQUIRE(nrows >= 0 && nb > 0, FEMHIP_EINVAL, "row sums: bad sizes");
  FEMHIP_REQUIRE(max_row_len <= kMaxRowLen, FEMHIP_EUNSUPPORTED,
                 "row sums: a CSR row longer than 384 slots");
  if (nrows == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(loc && pos && rowe_ptr && rowe && indptr && out, FEMHIP_EINVAL,
                 "row sums: NULL argument");
  row_sum_em_kernel<<<(unsigned)ceil_div(nrows, kRowWaves), 64 * kRowWaves, 0,
                      reinterpret_cast<hipStream_t>(s)>>>(loc, pos, rowe_ptr, rowe, indptr,
                                                          nrows, (int)nb, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_hex2_elasticity_mfma(const femhip_geom_t* g, const femhip_space_t* sp,
                                           double lam, double mu, const double* G,
                                           double* out, int32_t groups, femhip_stream_t s) {
  FEMHIP_REQUIRE(g && sp && G && out, FEMHIP_EINVAL, "hex2 mfma: NULL argument");
  FEMHIP_REQUIRE(g->kind == FEMHIP_HEX && g->nq == kQ && sp->nbs == kNB && sp->vector,
                 FEMHIP_EUNSUPPORTED,
                 "hex2 mfma: ElementVector(ElementHex2) with 27 quadrature points only");
  if (g->nelems == 0) return FEMHIP_OK;
  const int64_t grid = groups > 0 ? groups : (g->nelems < 32768 ? g->nelems : 32768);
  hex2_elasticity_mfma_kernel<<<(unsigned)grid, kThreads, 0,
                                reinterpret_cast<hipStream_t>(s)>>>(
      G, g->W, sp->dphi, g->nelems, lam, mu, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_segment_sum_em(const double* loc, const int32_t* perm, const int32_t* seg,
                                     int64_t nseg, int64_t nelems, int64_t nb, double* out,
                                     femhip_stream_t s) {
  FEMHIP_REQUIRE(nseg >= 0 && nelems > 0 && nb > 0, FEMHIP_EINVAL, "bad sizes");
  FEMHIP_REQUIRE(nelems * nb * nb <= INT32_MAX, FEMHIP_EOVERFLOW,
                 "segment sum: contributions exceed int32 indices");
  if (nseg == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(loc && perm && seg && out, FEMHIP_EINVAL, "NULL argument");
  segment_sum_em_kernel<<<(unsigned)ceil_div(nseg, 256), 256, 0,
                          reinterpret_cast<hipStream_t>(s)>>>(loc, perm, seg, nseg, (uint32_t)nelems,
                                                              (uint32_t)nb,
                                                              Div31::make((uint32_t)nelems),
                                                              Div31::make((uint32_t)nb), out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
