// Fused P2 assembly on simplices: the local entries in the reference's
// arithmetic, summed straight into LDS rows and written to the CSR values
// once -- no local-matrix array, no separate segmented sums.
//
// The plan (fused_plan.cpp, rows mode) cuts the DOFs into partitions whose
// CSR rows fit LDS and lists every element touching an owned DOF (a visit:
// the element's local DOF ids) and every (visit, owned DOF) pair (a row
// record: the visit, the DOF, the offsets of the element's columns in the
// DOF's row and of the DOF's column in the rows of the element's other owned
// DOFs).  One workgroup per partition:
//   1. stages the owned rows' table and zeroed int64 accumulators and
//      "nonzero seen" flags in LDS;
//   2. one lane per row record (element e, owned DOF i) computes the
//      entries (i, j) exactly as the exact-order kernel (local_row_kernel,
//      exact_local.h: the affine inverse of mapping_affine.py:55-131,
//      einsum-ordered gradients, NumPy's pairwise sum over the points).
//      Laplace and mass integrands are symmetric bit for bit, so a pair of
//      two owned DOFs is computed once, by the lower local index, and added
//      to both rows; a pair with a DOF owned elsewhere is computed here for
//      this partition's row;
//   3. writes the rows once, 16 lanes per row.
//
// The pattern is checked exactly in the same pass: an entry is zero in the
// reference iff it is 0.0 here (same operations), so a slot of the plan's
// pattern must receive at least one nonzero contribution (else
// eliminate_zeros would have dropped it), and a nonzero contribution must
// have a slot (offset 127: the pair is not in the pattern).  Either failure
// sets FEMHIP_FUSED_BAD.  Sums
// are int64 fixed point with one exponent per launch (fused_common.h).
// This file is compiled with -ffp-contract=off (reference arithmetic).
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
  const double* G;          // affine mapping per element (exact order)
  const double* W;
  const double* phi;        // [nb][Q]
  const double* dphi;       // [nb][D][Q]
  double* out;
  femhip_fused_state_t* st;
  int64_t E;
  int32_t nparts;
  int32_t max_slots;
  int32_t max_nodes;
  int32_t max_rows;
  int32_t mode;
};

constexpr int kRowsThreads = 512;
constexpr int kRowRecWords = 6;   // visit << 4 | DOF, 2 x 10 seven-bit offsets

// LDS: [reference tables (TAB doubles) | accumulators (max_slots int64) |
//  row info, row starts (max_rows each) | reduction (2 NW doubles) |
//  nonzero flags (max_slots bytes)]
template <int NT, int TAB>
struct RowsLds {
  double* tab;
  unsigned long long* acc;
  uint32_t* rinfo;
  int32_t* rstart;
  double* red;
  uint8_t* nz;
  __device__ RowsLds(double* smem, const RowsArgs& a) {
    tab = smem;
    acc = reinterpret_cast<unsigned long long*>(smem + TAB);
    rinfo = reinterpret_cast<uint32_t*>(acc + a.max_slots);
    rstart = reinterpret_cast<int32_t*>(rinfo + a.max_rows);
    red = reinterpret_cast<double*>(rstart + a.max_rows);
    nz = reinterpret_cast<uint8_t*>(red + 2 * (NT / 64));
  }
};

// without the reference tables (8 * TAB bytes more per instantiation)
size_t rows_lds_bytes(const femhip_fused_plan_info_t& I) {
  return 8 * (size_t)I.max_slots + 8 * (size_t)I.max_rows + 8 * 2 * (kRowsThreads / 64) +
         (size_t)I.max_slots;
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

template <int D, int NB, int Q, int FORM, int NT>
__global__ void __launch_bounds__(NT) p2_fused_rows_kernel(RowsArgs a) {
#pragma clang fp contract(off)
  constexpr bool GRAD = FORM == FEMHIP_LAPLACE;
  constexpr int RW = (NB * 16 + 31) / 32;
  static_assert(RW <= 5 && kRowRecWords == 6, "record words");
  constexpr int TAB = GRAD ? NB * D * Q : NB * Q;
  extern __shared__ double smem[];
  RowsLds<NT, TAB> L(smem, a);
  const int tid = threadIdx.x;
  femhip_fused_state_t* st = a.st;
  const bool first = a.mode == 0;
  const FxPass pass = fx_begin<false>(st, a.mode);
  if (!pass.run) return;  // uniform
  const bool scatter = pass.scatter;
  const double sc = ldexp(1.0, pass.s);
  double mk = 0.0;
  bool bad = false;
  // reference gradients [nb][D][Q] (values [nb][Q] for mass) in LDS: each
  // lane reads the table rows of its own trial functions
  for (int k = threadIdx.x; k < TAB; k += NT) L.tab[k] = GRAD ? a.dphi[k] : a.phi[k];

  for (int k = 0;; ++k) {
    const int part = next_partition(a.nparts, k);
    if (part < 0) break;
    if (k > 0) __syncthreads();  // LDS reuse
    const int32_t* hdr = a.parts + 8 * (int64_t)part;
    const int node_start = hdr[0], Ln = hdr[1], row_start = hdr[2], R = hdr[3];
    const int nslots = hdr[4], rr_start = hdr[5], nrr = hdr[6];

    for (int r = tid; r < R; r += NT) {
      const int2 rw = reinterpret_cast<const int2*>(a.rows)[row_start + r];
      L.rstart[r] = rw.x;
      L.rinfo[r] = (uint32_t)rw.y;
    }
    for (int i = tid; i < nslots; i += NT) {
      L.acc[i] = 0ull;
      L.nz[i] = 0;
    }
    __syncthreads();

    for (int c0 = 0; c0 < nrr; c0 += NT) {
      if (c0 + tid < nrr) {
        const uint32_t* rp = a.rowrecs + (int64_t)(rr_start + c0 + tid) * kRowRecWords;
        const uint32_t rr = rp[0];
        const uint32_t o0 = rp[1], o1 = rp[2], o2 = rp[3], o3 = rp[4], o4 = rp[5];
        const int64_t v = rr >> 4;
        const int i = (int)(rr & 15u);
        const uint32_t* vp = a.recs + v * RW;
        const uint32_t w0 = vp[0], w1 = vp[1], w2 = RW > 2 ? vp[2] : 0u,
                       w3 = RW > 3 ? vp[3] : 0u, w4 = RW > 4 ? vp[4] : 0u;
        // the element's DOF j: local id (16-bit fields) and the 7-bit offset
        // fields of the row record (four per word)
        auto lid_of = [&](int j) -> int {
          return (int)((sel5(j >> 1, w0, w1, w2, w3, w4) >> ((j & 1) * 16)) & 0xffffu);
        };
        auto field = [&](int k) -> int {
          return (int)((sel5(k >> 2, o0, o1, o2, o3, o4) >> (7 * (k & 3))) & 127u);
        };
        const int64_t e = a.rec_elem[v];
        exact::MapAt<D, true, Q> m;
        m.load(a.G, a.E, e, 0);
        double vi[Q][GRAD ? D : 1];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if constexpr (GRAD)
            exact::global_grad<D>(m.inv, L.tab, i, q, Q, vi[q]);
          else
            vi[q][0] = L.tab[i * Q + q];
        }
        // the local row of DOF i, from the top nibbles of offset words 0-3
        const int ri = (int)((o0 >> 28) | ((o1 >> 28) << 4) | ((o2 >> 28) << 8) |
                             ((o3 >> 28) << 12));
        const int basei = (int)(L.rinfo[ri] & 0xffffu);   // an owned row
        // the trial functions this lane computes: j >= i when both DOFs are
        // owned here (the pair is added to both rows), every j whose DOF is
        // owned elsewhere.  The records are sorted by (i, this set), so a
        // wave's lanes mostly share it and the loop index stays uniform
        // (the reference gradients of j are read as scalars).
        uint32_t need = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (lid_of(j) >= R || j >= i) need |= 1u << j;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          if (!((need >> j) & 1u)) continue;
          const int rj = lid_of(j);
          const bool ownj = rj < R;
          auto term = [&](int q) -> double {
#pragma clang fp contract(off)
            const double dx = m.adet * a.W[q];
            double val;
            if constexpr (GRAD) {
              double gj[D];
              exact::global_grad<D>(m.inv, a.dphi, j, q, Q, gj);
              double s = 0.0;
#pragma unroll
              for (int d = 0; d < D; ++d) s = s + gj[d] * vi[q][d];
              val = s;
            } else {
              val = a.phi[j * Q + q] * vi[q][0];
            }
            return val * dx;
          };
          const double r = 0.0 + np_pairwise<0, Q>(term);
          mk = max_abs(mk, r);
          if (r == 0.0) continue;   // dropped by eliminate_zeros
          const unsigned long long fx = to_fixed(r, sc);
          const int oij = field(j);
          if (oij == 127) {
            bad = true;
          } else if (scatter) {
            lds_add(L.acc + basei + oij, fx);
            L.nz[basei + oij] = 1;
          }
          if (ownj && j != i) {
            const int oji = field(NB + j);
            if (oji == 127) {
              bad = true;
            } else if (scatter) {
              const int sji = (int)(L.rinfo[rj] & 0xffffu) + oji;
              lds_add(L.acc + sji, fx);
              L.nz[sji] = 1;
            }
          }
        }
      }
    }
    if (!scatter) continue;
    __syncthreads();  // every contribution has landed
    const double inv = ldexp(1.0, -pass.s);
    const int lane = tid & 15, grp = tid >> 4;
    for (int r = grp; r < R; r += NT / 16) {
      const uint32_t info = L.rinfo[r];
      const int base = (int)(info & 0xffffu), len = (int)((info >> 24) & 0x7fu);
      const int64_t gs = L.rstart[r];
      for (int q = lane; q < len; q += 16) {
        bad = bad || L.nz[base + q] == 0;   // a slot the reference would prune
        __builtin_nontemporal_store(from_fixed((long long)L.acc[base + q], inv), a.out + gs + q);
      }
    }
  }  // partitions

  if (__any(bad) && (tid & 63) == 0) atomicOr(&st->flags, FEMHIP_FUSED_BAD);
  if (first) fx_end<NT, false>(st, mk, 0.0, false, false, L.red);
}

template <int D, int NB, int Q>
int launch_rows(const RowsArgs& a, int form, size_t lds, unsigned grid, hipStream_t s) {
  if (form == FEMHIP_LAPLACE)
    p2_fused_rows_kernel<D, NB, Q, FEMHIP_LAPLACE, kRowsThreads>
        <<<grid, kRowsThreads, lds + 8 * NB * D * Q, s>>>(a);
  else
    p2_fused_rows_kernel<D, NB, Q, FEMHIP_MASS, kRowsThreads>
        <<<grid, kRowsThreads, lds + 8 * NB * Q, s>>>(a);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" size_t femhip_rows_fused_lds_bytes(const femhip_fused_plan_info_t* I) {
  return I ? rows_lds_bytes(*I) : 0;
}

extern "C" int femhip_rows_fused_assemble(const femhip_fused_dev_t* plan, const uint32_t* rowrecs,
                                          femhip_fused_state_t* state, int32_t form,
                                          const femhip_geom_t* g, const femhip_space_t* sp,
                                          const double* G, double* data,
                                          const femhip_fused_check_t* check, femhip_stream_t s) {
  FEMHIP_REQUIRE(plan && rowrecs && state && g && sp && G && data, FEMHIP_EINVAL,
                 "rows fused: NULL argument");
  FEMHIP_REQUIRE(reinterpret_cast<uintptr_t>(state) % 128 == 0, FEMHIP_EINVAL,
                 "rows fused: state must be 128-byte aligned");
  FEMHIP_REQUIRE(form == FEMHIP_LAPLACE || form == FEMHIP_MASS, FEMHIP_EUNSUPPORTED,
                 "rows fused: laplace or mass only");
  FEMHIP_REQUIRE(check == nullptr || check->verdict != nullptr, FEMHIP_EINVAL,
                 "rows fused: the check needs a verdict output");
  const femhip_fused_plan_info_t& I = plan->info;
  FEMHIP_REQUIRE(I.idbits == 16 && I.offbits == 0, FEMHIP_EINVAL,
                 "rows fused: not a rows plan");
  const bool tet = g->kind == FEMHIP_TET && g->dim == 3 && sp->nbs == 10 && !sp->vector;
  const bool tri = g->kind == FEMHIP_TRI && g->dim == 2 && sp->nbs == 6 && !sp->vector;
  FEMHIP_REQUIRE(tet || tri, FEMHIP_EUNSUPPORTED, "rows fused: P2 triangles or tetrahedra");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (I.nparts > 0) {
    RowsArgs a{plan->parts, plan->nodemap, plan->rows,  plan->recs,  plan->rec_elem,
               rowrecs,     G,             g->W,        sp->phi,     sp->dphi,
               data,        state,         g->nelems,   I.nparts,    I.max_slots,
               I.max_nodes, I.max_rows,    0};
    const size_t lds = rows_lds_bytes(I);   // + the reference tables (<= 2,640 B)
    FEMHIP_REQUIRE(lds + 8 * 10 * 3 * 11 <= 160 * 1024, FEMHIP_EINVAL,
                   "rows fused: plan exceeds LDS");
    for (int mode = 0; mode < 2; ++mode) {
      a.mode = mode;
      const unsigned grid =
          (unsigned)(mode == 0 || I.nparts < (int)kRepeatGrid ? I.nparts : kRepeatGrid);
      int rc = FEMHIP_EUNSUPPORTED;
      if (tet && g->nq == 11) rc = launch_rows<3, 10, 11>(a, form, lds, grid, st);
      else if (tet && g->nq == 4) rc = launch_rows<3, 10, 4>(a, form, lds, grid, st);
      else if (tri && g->nq == 6) rc = launch_rows<2, 6, 6>(a, form, lds, grid, st);
      else if (tri && g->nq == 3) rc = launch_rows<2, 6, 3>(a, form, lds, grid, st);
      else set_error("rows fused: unsupported number of quadrature points");
      if (rc != FEMHIP_OK) return rc;
    }
  }
  if (check) return fused_check_empty(check, state, st);
  return FEMHIP_OK;
}
