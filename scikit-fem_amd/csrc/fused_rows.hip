// Fused P2 assembly on simplices: the local entries summed straight into
// LDS rows and written to the CSR values once -- no local-matrix array, no
// separate segmented sums.
//
// The plan (fused_plan.cpp, rows mode) cuts the DOFs into partitions whose
// CSR rows fit LDS and lists every element touching an owned DOF (a visit:
// the element's local DOF ids) and every (visit, owned DOF) pair (a row
// record: the visit, the DOF, the offsets of the element's columns in the
// DOF's row and of the DOF's column in the rows of the element's other owned
// DOFs).  One workgroup per partition:
//   1. forms the element-independent tables of the form in LDS and stages
//      the owned rows' table and zeroed int64 accumulators;
//   2. one lane per row record (element e, owned DOF i) forms the entries
//      (i, j) from the element metric: on an affine simplex the quadrature
//      sum of the reference (sum_q w_q |det| (A^-T dphi_i) . (A^-T dphi_j))
//      equals sum_ab M_ab T_ab,ij with M = |det| A^-1 A^-T (6 numbers per
//      element in 3-D) and T_ab,ij = sum_q w_q dphi_i,a dphi_j,b (mass:
//      |det| sum_q w_q phi_i phi_j), so an entry is D(D+1)/2 fused
//      multiply-adds.  A pair of two owned DOFs is computed once, by the
//      lower local index, and added to both rows; a pair with a DOF owned
//      elsewhere is computed here for this partition's row;
//   3. writes the rows once, 16 lanes per row.
//
// Validation (as the fused P1 kernel): the entries differ from the
// reference's by rounding; an entry the reference computes as exactly 0.0
// comes out within tau_ij = 2^-44 D max|A^-1|^2 |det| B_ij of zero (mass:
// 2^-44 |det| B_ij), B_ij = sum_q |w_q| |dphi_i|_1 |dphi_j|_1 (mass:
// sum_q |w_q| |phi_i| |phi_j|), a bound ~9x above the rounding of both
// evaluations.  Entries within tau are queued and recomputed in the
// reference's arithmetic afterwards (fused_recheck_kernel), and the call
// is valid iff the zero set found is the plan's; an entry above tau whose
// pair is not in the pattern sets FEMHIP_FUSED_BAD.  Sums are int64 fixed
// point with one exponent per launch (fused_common.h).
#include "exact_local.h"
#include "fused_check.h"
#include "fused_common.h"

namespace femhip {
namespace {

struct RowsArgs {
  const int32_t* parts;
  const int32_t* nodemap;
  const int32_t* rows;
  const uint32_t* recs;
  const int32_t* rec_elem;
  const uint32_t* rowrecs;  // kRowRecWords words per row record
  const double* metric;     // [E][kMetric] (femhip_rows_metric)
  const double* kappa;      // kappa-Laplace: per element
  const double* W;
  const double* phi;        // [nb][Q]
  const double* dphi;       // [nb][D][Q]
  double* out;
  femhip_fused_state_t* st;
  uint64_t* susp;           // suspect queue (checked calls) or NULL
  int64_t susp_cap;
  int64_t E;
  int32_t nparts;
  int32_t max_slots;
  int32_t max_nodes;
  int32_t max_rows;
  int32_t mode;
};

// per element: the metric |det| A^-1 A^-T (pairs (a, a), then (0,1), (0,2),
// (1,2)) in [0, D(D+1)/2), the Laplace suspect scale 2^-44 D max|A^-1|^2
// |det| in [6] and |det| in [7]
constexpr int kMetric = 8;
constexpr int kRowRecWords = 6;   // visit << 4 | DOF, 2 x 10 seven-bit offsets

// LDS: [form tables (TAB doubles: T then B) | accumulators (max_slots
//  int64) | row info, row starts (max_rows each) | reduction (2 NW doubles)]
template <int NT, int TAB>
struct RowsLds {
  double* tab;
  unsigned long long* acc;
  uint32_t* rinfo;
  int32_t* rstart;
  double* red;
  __device__ RowsLds(double* smem, const RowsArgs& a) {
    tab = smem;
    acc = reinterpret_cast<unsigned long long*>(smem + TAB);
    rinfo = reinterpret_cast<uint32_t*>(acc + a.max_slots);
    rstart = reinterpret_cast<int32_t*>(rinfo + a.max_rows);
    red = reinterpret_cast<double*>(rstart + a.max_rows);
  }
};

// without the form tables (rows_tab_doubles() more per launch); the
// workgroup size is the plan's block (512 or 1024)
size_t rows_lds_bytes(const femhip_fused_plan_info_t& I) {
  return 8 * (size_t)I.max_slots + 8 * (size_t)I.max_rows + 8 * 2 * (size_t)(I.block / 64);
}

// doubles of the form tables: T (NB^2 D(D+1)/2 for Laplace, NB^2 for mass)
// and the suspect bound B (NB^2)
constexpr int rows_tab_doubles(int D, int NB, bool grad) {
  return NB * NB * (grad ? D * (D + 1) / 2 : 1) + NB * NB;
}

// x selected from five words by a per-lane index, as a chain of selects
__device__ __forceinline__ uint32_t sel5(int k, uint32_t a0, uint32_t a1, uint32_t a2,
                                         uint32_t a3, uint32_t a4) {
  uint32_t x = a0;
  x = k == 1 ? a1 : x;
  x = k == 2 ? a2 : x;
  x = k == 3 ? a3 : x;
  x = k == 4 ? a4 : x;
  return x;
}

// Suspects of one row record into the global queue: bits j of m1 are the
// entries (row i, column j), bits j of m2 the entries (row j, column i);
// one queue reservation per lane that has any (rare).
__device__ __forceinline__ void push_row_suspects(const RowsArgs& a, int64_t e, int i,
                                                  uint32_t m1, uint32_t m2) {
  const int cnt = __popc(m1) + __popc(m2);
  long long pos = (long long)atomicAdd(reinterpret_cast<unsigned long long*>(&a.st->nsusp),
                                       (unsigned long long)cnt);
  if (pos + cnt > a.susp_cap) atomicOr(&a.st->flags, FEMHIP_FUSED_OVERFLOW);
  const uint64_t base = (uint64_t)(uint32_t)e;
  while (m1) {
    const int j = __ffs(m1) - 1;
    m1 &= m1 - 1;
    if (pos < a.susp_cap) a.susp[pos] = base | ((uint64_t)i << 32) | ((uint64_t)j << 40);
    ++pos;
  }
  while (m2) {
    const int j = __ffs(m2) - 1;
    m2 &= m2 - 1;
    if (pos < a.susp_cap) a.susp[pos] = base | ((uint64_t)j << 32) | ((uint64_t)i << 40);
    ++pos;
  }
}

template <int D, int NB, int Q, int FORM, int NT>
__global__ void __launch_bounds__(NT) p2_fused_rows_kernel(RowsArgs a) {
  constexpr bool GRAD = FORM != FEMHIP_MASS;   // laplace, kappa_laplace
  constexpr bool KAPPA = FORM == FEMHIP_KAPPA_LAPLACE;
  constexpr int NM = GRAD ? D * (D + 1) / 2 : 1;   // metric entries per element
  constexpr int RW = (NB * 16 + 31) / 32;
  constexpr int RV = RW + 1;   // visit record: local ids, then the element id
  static_assert((RV == 4 || RV == 6) && kRowRecWords == 6, "record words");
  constexpr int TAB = rows_tab_doubles(D, NB, GRAD);
  extern __shared__ double smem[];
  RowsLds<NT, TAB> L(smem, a);
  double* const T = L.tab;                 // [NB][NB][NM]
  double* const B = L.tab + NB * NB * NM;  // [NB][NB]
  const int tid = threadIdx.x;
  femhip_fused_state_t* st = a.st;
  const bool first = a.mode == 0;
  const FxPass pass = fx_begin<false>(st, a.mode);
  if (!pass.run) return;  // uniform
  const bool scatter = pass.scatter;
  const bool suspects = first && a.susp != nullptr;
  const double sc = ldexp(1.0, pass.s);
  double mk = 0.0;
  bool bad = false;
  // the form tables from the reference tables and weights: metric pair m =
  // (a, a) for m < D, then (0,1), (0,2), (1,2) (the two orders summed)
  for (int k = tid; k < NB * NB * NM + NB * NB; k += NT) {
    double acc = 0.0;
    if (k < NB * NB * NM) {
      const int pr = k / NM, m = k - pr * NM;
      const int i = pr / NB, j = pr - i * NB;
      if constexpr (GRAD) {
        const int ca = m < D ? m : (m == D ? 0 : (m == D + 1 ? 0 : 1));
        const int cb = m < D ? m : (m == D ? 1 : 2);
        for (int q = 0; q < Q; ++q) {
          const double wi = a.W[q] * a.dphi[(i * D + ca) * Q + q];
          acc += wi * a.dphi[(j * D + cb) * Q + q];
          if (m >= D) acc += a.W[q] * a.dphi[(i * D + cb) * Q + q] * a.dphi[(j * D + ca) * Q + q];
        }
      } else {
        for (int q = 0; q < Q; ++q) acc += a.W[q] * a.phi[i * Q + q] * a.phi[j * Q + q];
      }
      T[k] = acc;
    } else {
      const int pr = k - NB * NB * NM;
      const int i = pr / NB, j = pr - i * NB;
      for (int q = 0; q < Q; ++q) {
        double ni = 0.0, nj = 0.0;
        if constexpr (GRAD) {
          for (int d = 0; d < D; ++d) {
            ni += fabs(a.dphi[(i * D + d) * Q + q]);
            nj += fabs(a.dphi[(j * D + d) * Q + q]);
          }
        } else {
          ni = fabs(a.phi[i * Q + q]);
          nj = fabs(a.phi[j * Q + q]);
        }
        acc += fabs(a.W[q]) * ni * nj;
      }
      B[pr] = acc;
    }
  }

  for (int k = 0;; ++k) {
    const int part = next_partition(a.nparts, k);
    if (part < 0) break;
    __syncthreads();  // the tables (k = 0), LDS reuse (k > 0)
    const int32_t* hdr = a.parts + 8 * (int64_t)part;
    const int row_start = hdr[2], R = hdr[3];
    const int nslots = hdr[4], rr_start = hdr[5], nrr = hdr[6];

    for (int r = tid; r < R; r += NT) {
      const int2 rw = reinterpret_cast<const int2*>(a.rows)[row_start + r];
      L.rstart[r] = rw.x;
      L.rinfo[r] = (uint32_t)rw.y;
    }
    for (int i = tid; i < nslots; i += NT) L.acc[i] = 0ull;
    __syncthreads();

    // row records in chunks of NT, software-pipelined: the next chunk's
    // records and their visits load while this chunk computes (the index
    // of a lane past the end is clamped to the last record, so every lane
    // issues the same loads)
    struct RecIn {
      uint32_t rr, o0, o1, o2, o3, o4;   // row record
      uint32_t w0, w1, w2, w3, w4;       // its visit's local DOF ids
      int64_t e;                         // and element
    };
    auto load_rec = [&](int c0, RecIn& x) {
      const int r = min(c0 + tid, nrr - 1);
      const uint2* rp =
          reinterpret_cast<const uint2*>(a.rowrecs + (int64_t)(rr_start + r) * kRowRecWords);
      const uint2 p0 = rp[0], p1 = rp[1], p2 = rp[2];
      x.rr = p0.x;
      x.o0 = p0.y;
      x.o1 = p1.x;
      x.o2 = p1.y;
      x.o3 = p2.x;
      x.o4 = p2.y;
    };
    auto load_visit = [&](RecIn& x) {
      const int64_t v = x.rr >> 4;
      if constexpr (RV == 4) {
        const uint4 q = reinterpret_cast<const uint4*>(a.recs)[v];
        x.w0 = q.x;
        x.w1 = q.y;
        x.w2 = q.z;
        x.w3 = 0u;
        x.w4 = 0u;
        x.e = q.w;
      } else {
        const uint2* vp = reinterpret_cast<const uint2*>(a.recs + v * RV);
        const uint2 q0 = vp[0], q1 = vp[1], q2 = vp[2];
        x.w0 = q0.x;
        x.w1 = q0.y;
        x.w2 = q1.x;
        x.w3 = q1.y;
        x.w4 = q2.x;
        x.e = q2.y;
      }
    };
    RecIn cur{}, nxt{};
    if (nrr > 0) {
      load_rec(0, cur);
      load_visit(cur);
    }
    for (int c0 = 0; c0 < nrr; c0 += NT) {
      // this chunk's element metrics (4 x 16 B per lane)
      const double2* mp = reinterpret_cast<const double2*>(a.metric + kMetric * cur.e);
      double2 q0 = make_double2(0.0, 0.0), q1 = q0, q2 = q0;
      if constexpr (GRAD) {
        q0 = mp[0];
        q1 = mp[1];
        if constexpr (D == 3) q2 = mp[2];
      }
      const double2 q3 = mp[3];
      const double kap = KAPPA ? a.kappa[cur.e] : 1.0;
      // the next chunk's loads are issued on every path (past the last
      // chunk they reload the last record)
      load_rec(c0 + NT, nxt);
      load_visit(nxt);
      if (c0 + tid < nrr) {
        const uint32_t rr = cur.rr;
        const uint32_t o0 = cur.o0, o1 = cur.o1, o2 = cur.o2, o3 = cur.o3, o4 = cur.o4;
        const int i = (int)(rr & 15u);
        const uint32_t w0 = cur.w0, w1 = cur.w1, w2 = cur.w2, w3 = cur.w3, w4 = cur.w4;
        // the element's DOF j: local id (16-bit fields) and the 7-bit offset
        // fields of the row record (four per word)
        auto lid_of = [&](int j) -> int {
          return (int)((sel5(j >> 1, w0, w1, w2, w3, w4) >> ((j & 1) * 16)) & 0xffffu);
        };
        auto field = [&](int kk) -> int {
          return (int)((sel5(kk >> 2, o0, o1, o2, o3, o4) >> (7 * (kk & 3))) & 127u);
        };
        const int64_t e = cur.e;
        // the metric and the suspect scale of the element
        double M[NM];
        double tsc;
        if constexpr (GRAD) {
          M[0] = q0.x;
          M[1] = q0.y;
          M[2] = q1.x;
          if constexpr (D == 3) {
            M[3] = q1.y;
            M[4] = q2.x;
            M[5] = q2.y;
          }
          tsc = q3.x;
          if constexpr (KAPPA) {   // kappa |det| A^-1 A^-T, the bound times |kappa|
#pragma unroll
            for (int mm = 0; mm < NM; ++mm) M[mm] *= kap;
            tsc *= fabs(kap);
          }
        } else {
          M[0] = q3.y;
          tsc = 0x1p-44 * q3.y;
        }
        bad = bad || !(tsc < INFINITY);
        // the local row of DOF i, from the top nibbles of offset words 0-3
        const int ri = (int)((o0 >> 28) | ((o1 >> 28) << 4) | ((o2 >> 28) << 8) |
                             ((o3 >> 28) << 12));
        const int basei = (int)(L.rinfo[ri] & 0xffffu);   // an owned row
        // the trial functions this lane computes: j >= i when both DOFs are
        // owned here (the pair is added to both rows), every j whose DOF is
        // owned elsewhere
        uint32_t need = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (lid_of(j) >= R || j >= i) need |= 1u << j;
        uint32_t s1 = 0, s2 = 0;   // suspects (row i, col j) and (row j, col i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          if (!((need >> j) & 1u)) continue;
          const int rj = lid_of(j);
          const bool ownj = rj < R;
          const double* tij = T + (i * NB + j) * NM;
          double val = M[0] * tij[0];
#pragma unroll
          for (int mm = 1; mm < NM; ++mm) val = __fma_rn(M[mm], tij[mm], val);
          mk = max_abs(mk, val);
          const bool sus = !(fabs(val) > tsc * B[i * NB + j]);
          const bool two = ownj && j != i;
          if (sus) {
            s1 |= 1u << j;
            if (two) s2 |= 1u << j;
          }
          const unsigned long long fx = to_fixed(val, sc);
          const int oij = field(j);
          if (oij == 127) {
            bad = bad || !sus;   // a nonzero entry outside the pattern
          } else if (scatter) {
            lds_add(L.acc + basei + oij, fx);
          }
          if (two) {
            const int oji = field(NB + j);
            if (oji == 127) {
              bad = bad || !sus;
            } else if (scatter) {
              lds_add(L.acc + (int)(L.rinfo[rj] & 0xffffu) + oji, fx);
            }
          }
        }
        if (suspects && (s1 | s2)) push_row_suspects(a, e, i, s1, s2);
      }
      cur = nxt;
    }
    if (!scatter) continue;
    __syncthreads();  // every contribution has landed
    const double inv = ldexp(1.0, -pass.s);
    const int lane = tid & 15, grp = tid >> 4;
    for (int r = grp; r < R; r += NT / 16) {
      const uint32_t info = L.rinfo[r];
      const int base = (int)(info & 0xffffu), len = (int)((info >> 24) & 0x7fu);
      const int64_t gs = L.rstart[r];
      // plain stores: the L2 merges the partial lines at row ends (P2 rows
      // are not contiguous in the CSR; streaming stores wrote 1.5x the bytes)
      for (int q = lane; q < len; q += 16)
        a.out[gs + q] = from_fixed((long long)L.acc[base + q], inv);
    }
  }  // partitions

  if (__any(bad) && (tid & 63) == 0) atomicOr(&st->flags, FEMHIP_FUSED_BAD);
  if (first) fx_end<NT, false>(st, mk, 0.0, false, false, L.red);
}

// Per element e of an affine simplex basis: kMetric doubles (see kMetric)
// from the mapping G (femhip_geometry: A^-1 row-major, then |det|).
template <int D>
__global__ void __launch_bounds__(256) rows_metric_kernel(const double* __restrict__ G,
                                                          int64_t E, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= E) return;
  exact::MapAt<D, true, 1> m;
  m.load(G, E, e, 0);
  double o[kMetric] = {0, 0, 0, 0, 0, 0, 0, 0};
  double mi = 0.0;
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) mi = max_abs(mi, m.inv[r][c]);
  constexpr int NM = D * (D + 1) / 2;
#pragma unroll
  for (int mm = 0; mm < NM; ++mm) {
    const int ca = mm < D ? mm : (mm == D ? 0 : (mm == D + 1 ? 0 : 1));
    const int cb = mm < D ? mm : (mm == D ? 1 : 2);
    double s = m.inv[ca][0] * m.inv[cb][0];
#pragma unroll
    for (int c = 1; c < D; ++c) s = __fma_rn(m.inv[ca][c], m.inv[cb][c], s);
    o[mm] = m.adet * s;
  }
  o[6] = (0x1p-44 * D) * (mi * mi) * m.adet;
  o[7] = m.adet;
  double2* dst = reinterpret_cast<double2*>(out + kMetric * e);
#pragma unroll
  for (int k = 0; k < kMetric / 2; ++k) dst[k] = make_double2(o[2 * k], o[2 * k + 1]);
}

template <int D, int NB, int Q, int NT>
int launch_rows_nt(const RowsArgs& a, int form, size_t lds, unsigned grid, hipStream_t s) {
  if (form == FEMHIP_KAPPA_LAPLACE)
    p2_fused_rows_kernel<D, NB, Q, FEMHIP_KAPPA_LAPLACE, NT>
        <<<grid, NT, lds + 8 * rows_tab_doubles(D, NB, true), s>>>(a);
  else if (form == FEMHIP_LAPLACE)
    p2_fused_rows_kernel<D, NB, Q, FEMHIP_LAPLACE, NT>
        <<<grid, NT, lds + 8 * rows_tab_doubles(D, NB, true), s>>>(a);
  else
    p2_fused_rows_kernel<D, NB, Q, FEMHIP_MASS, NT>
        <<<grid, NT, lds + 8 * rows_tab_doubles(D, NB, false), s>>>(a);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

template <int D, int NB, int Q>
int launch_rows(const RowsArgs& a, int form, int block, size_t lds, unsigned grid,
                hipStream_t s) {
  return block == 1024 ? launch_rows_nt<D, NB, Q, 1024>(a, form, lds, grid, s)
                       : launch_rows_nt<D, NB, Q, 512>(a, form, lds, grid, s);
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" size_t femhip_rows_fused_lds_bytes(const femhip_fused_plan_info_t* I) {
  return I ? rows_lds_bytes(*I) : 0;
}

extern "C" int femhip_rows_metric(const femhip_geom_t* g, const double* G, double* metric,
                                  femhip_stream_t s) {
  FEMHIP_REQUIRE(g && (g->kind == FEMHIP_TRI || g->kind == FEMHIP_TET) && g->nelems >= 0,
                 FEMHIP_EINVAL, "rows metric: affine simplex geometry required");
  if (g->nelems == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(G && metric && reinterpret_cast<uintptr_t>(metric) % 16 == 0, FEMHIP_EINVAL,
                 "rows metric: NULL or misaligned buffer");
  const unsigned grid = (unsigned)ceil_div(g->nelems, 256);
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (g->dim == 3)
    rows_metric_kernel<3><<<grid, 256, 0, st>>>(G, g->nelems, metric);
  else
    rows_metric_kernel<2><<<grid, 256, 0, st>>>(G, g->nelems, metric);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_rows_fused_assemble(const femhip_fused_dev_t* plan, const uint32_t* rowrecs,
                                          femhip_fused_state_t* state, int32_t form,
                                          const femhip_geom_t* g, const femhip_space_t* sp,
                                          const double* metric, const double* kappa,
                                          double* data, const femhip_fused_check_t* check,
                                          femhip_stream_t s) {
  FEMHIP_REQUIRE(plan && rowrecs && state && g && sp && metric && data, FEMHIP_EINVAL,
                 "rows fused: NULL argument");
  FEMHIP_REQUIRE(reinterpret_cast<uintptr_t>(metric) % 16 == 0, FEMHIP_EINVAL,
                 "rows fused: metric must be 16-byte aligned");
  FEMHIP_REQUIRE(reinterpret_cast<uintptr_t>(state) % 128 == 0, FEMHIP_EINVAL,
                 "rows fused: state must be 128-byte aligned");
  FEMHIP_REQUIRE(form == FEMHIP_LAPLACE || form == FEMHIP_MASS ||
                     form == FEMHIP_KAPPA_LAPLACE,
                 FEMHIP_EUNSUPPORTED, "rows fused: laplace, mass or kappa_laplace only");
  FEMHIP_REQUIRE((form == FEMHIP_KAPPA_LAPLACE) == (kappa != nullptr), FEMHIP_EINVAL,
                 "rows fused: kappa_laplace takes kappa (per element), the others none");
  FEMHIP_REQUIRE(check == nullptr || (check->verdict != nullptr && check->zero_mask != nullptr &&
                                        check->queue != nullptr && check->capacity > 0),
                 FEMHIP_EINVAL, "rows fused: the check needs a verdict, a zero mask and a queue");
  const femhip_fused_plan_info_t& I = plan->info;
  FEMHIP_REQUIRE(I.block == 512 || I.block == 1024, FEMHIP_EINVAL,
                 "rows fused: plan block must be 512 or 1024");
  FEMHIP_REQUIRE(I.idbits == 16 && I.offbits == 0 &&
                     I.rec_words == (sp->nbs * 16 + 31) / 32 + 1,
                 FEMHIP_EINVAL, "rows fused: not a rows plan of this element");
  const bool tet = g->kind == FEMHIP_TET && g->dim == 3 && sp->nbs == 10 && !sp->vector;
  const bool tri = g->kind == FEMHIP_TRI && g->dim == 2 && sp->nbs == 6 && !sp->vector;
  FEMHIP_REQUIRE(tet || tri, FEMHIP_EUNSUPPORTED, "rows fused: P2 triangles or tetrahedra");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (I.nparts > 0) {
    RowsArgs a{plan->parts,  plan->nodemap, plan->rows,  plan->recs,  plan->rec_elem,
               rowrecs,      metric,        kappa,       g->W,        sp->phi,
               sp->dphi,
               data,         state,
               check ? check->queue : nullptr, check ? check->capacity : 0,
               g->nelems,    I.nparts,      I.max_slots, I.max_nodes, I.max_rows,    0};
    const size_t lds = rows_lds_bytes(I);   // + the form tables (<= 5,600 B)
    FEMHIP_REQUIRE(lds + 8 * rows_tab_doubles(3, 10, true) <= 160 * 1024, FEMHIP_EINVAL,
                   "rows fused: plan exceeds LDS");
    for (int mode = 0; mode < 2; ++mode) {
      a.mode = mode;
      const unsigned grid =
          (unsigned)(mode == 0 || I.nparts < (int)kRepeatGrid ? I.nparts : kRepeatGrid);
      int rc = FEMHIP_EUNSUPPORTED;
      if (tet && g->nq == 11) rc = launch_rows<3, 10, 11>(a, form, I.block, lds, grid, st);
      else if (tet && g->nq == 4) rc = launch_rows<3, 10, 4>(a, form, I.block, lds, grid, st);
      else if (tri && g->nq == 6) rc = launch_rows<2, 6, 6>(a, form, I.block, lds, grid, st);
      else if (tri && g->nq == 3) rc = launch_rows<2, 6, 3>(a, form, I.block, lds, grid, st);
      else set_error("rows fused: unsupported number of quadrature points");
      if (rc != FEMHIP_OK) return rc;
    }
  }
  if (check) {
    if (I.nparts == 0) return fused_check_empty(check, state, st);
    return rows_check_run(check, state, form, g, sp, kappa, st);
  }
  return FEMHIP_OK;
}
