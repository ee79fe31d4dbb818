// Host-side builder of the fused P1 assembly plan (one-time, per mesh and
// CSR pattern).
//
// The fused kernel assembles P1 forms owner-computes style: rows are cut
// into partitions of spatially close nodes (recursive coordinate bisection
// of the node coordinates); one workgroup per partition accumulates the
// partition's CSR rows in LDS and writes them once.  Every element with at
// least one owned vertex is visited by that partition (halo elements are
// recomputed by each partition they touch instead of exchanging partial
// sums), so every off-diagonal contribution (e, i, j) of the reference's
// COO list is formed by exactly one partition: the owner of row
// element_dofs[i, e].
//
// Per element visit the plan stores one packed record: the local ids of the
// element's vertices (partition-local numbering, owned nodes first, each
// group in global order) and, for every vertex a and other vertex b, the
// offset of column b inside row a of the CSR pattern (a halo vertex's row
// too: the values kernel never reads those, the adjoint reads its upstream
// gradient there).  A partition's records are contiguous; the kernel takes
// them in chunks of `block`.  Records are ordered by a greedy colouring
// (elements of one colour share no owned vertex, so the lanes of a chunk
// of one colour add into distinct rows) or by element id.
//
// Rows are laid out back to back in LDS: with the usual odd row lengths (15
// in a tetrahedral interior) consecutive rows start on different LDS banks.
// A row that lost entries to eliminate_zeros has one spare slot after its
// CSR slots; its pruned entries (a pruned diagonal too) carry the offset
// len = the spare slot, which is never written out.
//
// The blocks are built independently by a pool of host threads
// (FEMHIP_PLAN_THREADS, default: the CPUs available) and concatenated in
// block order: the plan is the same for any number of threads.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/femhip.h"

namespace femhip {
void set_error(const std::string& msg);
}

namespace {

// rows plans: words per row record (visit << 4 | DOF, then 2 x 10 7-bit
// column offsets)
constexpr int kRowRecWords = 6;

struct Plan {
  femhip_fused_plan_info_t info;
  std::vector<int32_t> parts;    // 8 int32 per partition
  std::vector<int32_t> nodemap;  // global node id per local node
  std::vector<int32_t> rows;     // 2 int32 per owned row: CSR row start, packed info
  std::vector<uint32_t> recs;    // rec_words uint32 per record
  std::vector<int32_t> rec_elem; // global element id per record
  std::vector<uint32_t> rowrecs; // rows plans: kRowRecWords words per owned row visit
};

// Writes `nbits` (<= 32) of `v` at bit position `pos` of a little-endian
// word array (the field may straddle two words).
void put_bits(uint32_t* w, int pos, int nbits, uint32_t v) {
  const uint64_t x = ((uint64_t)v & ((1ull << nbits) - 1ull)) << (pos & 31);
  w[pos >> 5] |= (uint32_t)x;
  if ((pos & 31) + nbits > 32) w[(pos >> 5) + 1] |= (uint32_t)(x >> 32);
}

// Threads for the per-block phase: FEMHIP_PLAN_THREADS, else the CPUs this
// process may run on (at most 32), at most one per block.
int plan_threads(int64_t nblocks) {
  int n = 0;
  if (const char* e = std::getenv("FEMHIP_PLAN_THREADS")) n = std::atoi(e);
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  if (n <= 0) n = 1;
  n = std::min(n, 32);
  return (int)std::max<int64_t>(1, std::min<int64_t>(n, nblocks));
}

}  // namespace

namespace {

// mode 0: P1 plans (records with the column offsets of every vertex pair);
// mode 1: rows plans for P2 simplices (records with the local DOF ids only,
// plus one row record per (visit, owned DOF); columns are found in LDS)
int build_impl(int mode, int32_t nverts, int32_t dim, int64_t E, int64_t Nv, const int32_t* t,
               const double* p, const int32_t* indptr, const int32_t* indices,
               int32_t slot_cap, int32_t node_cap, int32_t block, int32_t order_mode,
               void** handle_out, femhip_fused_plan_info_t* info_out) {
  using femhip::set_error;
  if (mode == 0 && !(nverts == 3 && dim == 2) && !(nverts == 4 && dim == 3)) {
    set_error("fused plan: P1 triangles or tetrahedra only");
    return FEMHIP_EUNSUPPORTED;
  }
  if (mode == 1 && !(nverts == 6 && dim == 2) && !(nverts == 10 && dim == 3)) {
    set_error("rows plan: P2 triangles or tetrahedra only");
    return FEMHIP_EUNSUPPORTED;
  }
  if (!t || !p || !indptr || !indices || !handle_out || !info_out || E < 0 || Nv <= 0 ||
      slot_cap <= 0 || node_cap <= 0 || block <= 0 || block > 1024 || block % 64 != 0 ||
      order_mode < 0 || order_mode > 1) {
    set_error("fused plan: bad arguments");
    return FEMHIP_EINVAL;
  }
  if (E >= INT32_MAX / 2 || Nv >= INT32_MAX / 2) {
    set_error("fused plan: mesh exceeds int32 ids");
    return FEMHIP_EOVERFLOW;
  }
  const int NV = nverts;
  const int noff = mode == 0 ? NV * (NV - 1) : 0;
  int max_row = 0;
  for (int64_t n = 0; n < Nv; ++n) max_row = std::max(max_row, indptr[n + 1] - indptr[n]);
  // offsets 0..len (len: the spare slot of a row with pruned entries)
  const int offbits = mode == 1 ? 0 : (max_row <= 15 ? 4 : (max_row <= 31 ? 5 : 0));
  if (mode == 0 && offbits == 0) {
    set_error("fused plan: CSR rows longer than 31 entries");
    return FEMHIP_EUNSUPPORTED;
  }
  if (mode == 1 && max_row > 126) {
    set_error("rows plan: CSR rows longer than 126 entries");
    return FEMHIP_EUNSUPPORTED;
  }
  const int idbits = offbits == 4 ? 12 : 16;
  const int total_bits = NV * idbits + noff * offbits;
  const int rec_words = (total_bits + 31) / 32;
  node_cap = std::min(node_cap, 1 << idbits);

  // host threads of every phase below, and fn(k, thread) for k in [0, n)
  // on up to that many threads
  const int nth = plan_threads(std::max<int64_t>(1, std::max<int64_t>(Nv, E) / 4096));
  auto parallel_for = [&](size_t n, auto&& fn) {
    std::atomic<size_t> next{0};
    auto work = [&](int th) {
      for (;;) {
        const size_t k = next.fetch_add(1);
        if (k >= n) break;
        fn(k, th);
      }
    };
    std::vector<std::thread> pool;
    const int m = (int)std::min<size_t>((size_t)nth, n);
    for (int th = 1; th < m; ++th) pool.emplace_back(work, th);
    work(0);
    for (auto& th : pool) th.join();
  };
  // element ranges of the parallel passes over the connectivity
  const int64_t echunk = 1 << 16;
  const size_t nechunks = (size_t)((E + echunk - 1) / echunk);

  // node -> element incidence (counted and filled with atomic increments;
  // the order of a node's elements is not used: every consumer sorts or
  // only counts them)
  std::vector<int64_t> inc_ptr(Nv + 1, 0);
  std::atomic<bool> out_of_range{false};
  parallel_for(nechunks, [&](size_t c, int) {
    const int64_t e1 = std::min<int64_t>(E, (int64_t)(c + 1) * echunk);
    for (int64_t e = (int64_t)c * echunk; e < e1; ++e)
      for (int k = 0; k < NV; ++k) {
        const int32_t v = t[(int64_t)k * E + e];
        if (v < 0 || v >= Nv) {
          out_of_range = true;
          continue;
        }
        __atomic_fetch_add(&inc_ptr[v + 1], (int64_t)1, __ATOMIC_RELAXED);
      }
  });
  if (out_of_range) {
    set_error("fused plan: connectivity index out of range");
    return FEMHIP_EINVAL;
  }
  for (int64_t n = 0; n < Nv; ++n) inc_ptr[n + 1] += inc_ptr[n];
  std::vector<int32_t> inc(inc_ptr[Nv]);
  {
    std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    parallel_for(nechunks, [&](size_t c, int) {
      const int64_t e1 = std::min<int64_t>(E, (int64_t)(c + 1) * echunk);
      for (int64_t e = (int64_t)c * echunk; e < e1; ++e)
        for (int k = 0; k < NV; ++k) {
          const int64_t pos =
              __atomic_fetch_add(&fill[t[(int64_t)k * E + e]], (int64_t)1, __ATOMIC_RELAXED);
          inc[pos] = (int32_t)e;
        }
    });
  }

  // 1. recursive coordinate bisection of the nodes into blocks whose CSR
  // rows fit `slot_cap` LDS slots and whose owned + halo nodes fit
  // `node_cap`; every split is weighted so the leaves come out close to the
  // caps (fewer, larger blocks = less halo work).
  std::vector<int32_t> order(Nv);
  for (int64_t n = 0; n < Nv; ++n) order[n] = (int32_t)n;
  std::vector<std::pair<int64_t, int64_t>> leaves;  // [begin, end) in `order`
  // one split step of the range [b, e): a leaf when its rows fit the caps,
  // else partitioned in `order` at the returned midpoint (-1: a leaf; -2: a
  // single row that does not fit).  `stamp` / `check_id` are the caller's.
  auto split = [&](int64_t b, int64_t e, std::vector<int64_t>& stamp,
                   int64_t& check_id) -> int64_t {
    auto local_nodes = [&]() {
      const int64_t id = check_id++;
      int64_t L = 0;
      for (int64_t k = b; k < e; ++k) {
        const int32_t n = order[k];
        if (stamp[n] != id) {
          stamp[n] = id;
          ++L;
        }
        for (int64_t q = inc_ptr[n]; q < inc_ptr[n + 1]; ++q)
          for (int j = 0; j < NV; ++j) {
            const int32_t v = t[(int64_t)j * E + inc[q]];
            if (stamp[v] != id) {
              stamp[v] = id;
              ++L;
            }
          }
      }
      return L;
    };
    int64_t slots = 0;
    for (int64_t k = b; k < e; ++k) slots += indptr[order[k] + 1] - indptr[order[k]] + 1;
    int64_t L = -1;
    if (slots <= slot_cap) {
      L = local_nodes();
      if (L <= node_cap) return -1;
    }
    if (e - b <= 1) return -2;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t k = b; k < e; ++k)
      for (int d = 0; d < dim; ++d) {
        const double v = p[d * Nv + order[k]];
        lo[d] = std::min(lo[d], v);
        hi[d] = std::max(hi[d], v);
      }
    int ax = 0;
    for (int d = 1; d < dim; ++d)
      if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    const int64_t nleaf = std::max<int64_t>(
        2, std::max<int64_t>((slots + slot_cap - 1) / slot_cap,
                             L > 0 ? (L + node_cap - 1) / node_cap : 2));
    const int64_t mid = b + (int64_t)((double)(e - b) * (double)(nleaf / 2) / (double)nleaf);
    std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                     [&](int32_t x, int32_t y) {
                       const double px = p[ax * Nv + x], py = p[ax * Nv + y];
                       return px < py || (px == py && x < y);
                     });
    return mid;
  };
  // depth-first, left part first, within [b, e): leaves come out spatially
  // ordered
  auto bisect = [&](int64_t b0, int64_t e0, std::vector<int64_t>& stamp, int64_t& check_id,
                    std::vector<std::pair<int64_t, int64_t>>& out) -> bool {
    std::vector<std::pair<int64_t, int64_t>> stack{{b0, e0}};
    while (!stack.empty()) {
      auto [b, e] = stack.back();
      stack.pop_back();
      const int64_t mid = split(b, e, stamp, check_id);
      if (mid == -1) {
        out.push_back({b, e});
      } else if (mid == -2) {
        return false;
      } else {
        stack.push_back({mid, e});
        stack.push_back({b, mid});
      }
    }
    return true;
  };
  {
    // the top levels breadth first (the ranges of a level split in
    // parallel) until there are enough independent ranges, then each range
    // bisected depth first on a pool thread; the leaves are concatenated in
    // range order, which is the depth-first order of one thread, so the
    // plan does not depend on the threads
    // per-thread stamp arrays (allocated on first use)
    std::vector<std::vector<int64_t>> stamps(nth);
    std::vector<int64_t> ids(nth, 0);
    auto stamp_of = [&](int th) -> std::vector<int64_t>& {
      if (stamps[th].empty()) stamps[th].assign(Nv, -1);
      return stamps[th];
    };
    // each entry: a range and whether it is already a leaf
    std::vector<std::tuple<int64_t, int64_t, bool>> ranges{{0, Nv, false}};
    while (nth > 1 && (int64_t)ranges.size() < 8 * (int64_t)nth) {
      std::vector<int64_t> mids(ranges.size(), -1);
      parallel_for(ranges.size(), [&](size_t k, int th) {
        const auto& [b, e, leaf] = ranges[k];
        if (!leaf) mids[k] = split(b, e, stamp_of(th), ids[th]);
      });
      std::vector<std::tuple<int64_t, int64_t, bool>> next;
      bool split_any = false;
      for (size_t k = 0; k < ranges.size(); ++k) {
        const auto& [b, e, leaf] = ranges[k];
        if (mids[k] == -2) {
          set_error("fused plan: one row does not fit the LDS budget");
          return FEMHIP_EUNSUPPORTED;
        }
        if (leaf || mids[k] == -1) {
          next.push_back({b, e, true});
        } else {
          next.push_back({b, mids[k], false});
          next.push_back({mids[k], e, false});
          split_any = true;
        }
      }
      ranges.swap(next);
      if (!split_any) break;
    }
    std::vector<std::vector<std::pair<int64_t, int64_t>>> parts_out(ranges.size());
    std::vector<char> ok(ranges.size(), 1);
    parallel_for(ranges.size(), [&](size_t k, int th) {
      const auto& [b, e, leaf] = ranges[k];
      if (leaf)
        parts_out[k].push_back({b, e});
      else
        ok[k] = bisect(b, e, stamp_of(th), ids[th], parts_out[k]) ? 1 : 0;
    });
    for (size_t k = 0; k < ranges.size(); ++k) {
      if (!ok[k]) {
        set_error("fused plan: one row does not fit the LDS budget");
        return FEMHIP_EUNSUPPORTED;
      }
      leaves.insert(leaves.end(), parts_out[k].begin(), parts_out[k].end());
    }
  }

  // 2. every leaf block on its own (per-thread scratch), in parallel; the
  // blocks are then concatenated in leaf order, so the plan does not depend
  // on the number of threads
  struct LeafOut {
    std::vector<int32_t> nodemap, rows, rec_elem;
    std::vector<uint32_t> recs, rowrecs, rowkeys;
    int L = 0, R = 0, lds = 0, nelems = 0, err = 0;
    bool pruned = false;   // some owned row lacks a pair of one of its elements
    std::string msg;
  };
  struct Scratch {
    std::vector<int32_t> nstamp, lid, estamp, owned, elems, color, locals;
    std::vector<uint64_t> used;
    Scratch(int64_t nv, int64_t ne) : nstamp(nv, -1), lid(nv, -1), estamp(ne, -1) {}
  };
  std::vector<LeafOut> outs(leaves.size());
  auto build_leaf = [&](size_t leaf, Scratch& S, LeafOut& O) {
    auto& owned = S.owned;
    auto& elems = S.elems;
    auto& lid = S.lid;
    auto& locals = S.locals;
    const int32_t pid = (int32_t)leaf;
    owned.clear();
    elems.clear();
    for (int64_t k = leaves[leaf].first; k < leaves[leaf].second; ++k)
      owned.push_back(order[k]);
    for (int32_t n : owned)
      for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
        const int32_t e = inc[k];
        if (S.estamp[e] == pid) continue;
        S.estamp[e] = pid;
        elems.push_back(e);
      }
    // local numbering: owned rows first, then halo nodes, each in global
    // order -- consecutive owned rows are then mostly consecutive CSR rows
    // (the write-back of neighbouring lanes lands in the same cache lines)
    // and the coordinate gathers of the staging step are coalesced
    std::sort(owned.begin(), owned.end());
    std::sort(elems.begin(), elems.end());
    const int R = (int)owned.size();
    for (int r = 0; r < R; ++r) lid[owned[r]] = r;
    locals.assign(owned.begin(), owned.end());
    for (int32_t e : elems)
      for (int j = 0; j < NV; ++j) {
        const int32_t v = t[(int64_t)j * E + e];
        if (lid[v] == -1) {
          lid[v] = -2;  // halo, numbered below
          locals.push_back(v);
        }
      }
    std::sort(locals.begin() + R, locals.end());
    const int L = (int)locals.size();
    auto release = [&]() {
      for (int32_t v : locals) lid[v] = -1;
    };
    if (L > node_cap || R > 65535) {
      release();
      O.err = FEMHIP_EUNSUPPORTED;
      O.msg = "fused plan: a block exceeds the local node budget";
      return;
    }
    for (int l = R; l < L; ++l) lid[locals[l]] = l;
    // record order: colour classes (no two elements of a colour share an
    // owned vertex), each in element order; element order when more than
    // 64 colours would be needed or requested
    std::vector<int32_t> ord(elems.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int32_t)i;
    if (order_mode == 0) {
      S.used.assign(R, 0);
      S.color.assign(elems.size(), 0);
      bool ok = true;
      for (size_t i = 0; i < elems.size() && ok; ++i) {
        const int32_t e = elems[i];
        uint64_t forbid = 0;
        for (int j = 0; j < NV; ++j) {
          const int l = lid[t[(int64_t)j * E + e]];
          if (l < R) forbid |= S.used[l];
        }
        if (~forbid == 0) {
          ok = false;
          break;
        }
        const int c = __builtin_ctzll(~forbid);
        S.color[i] = c;
        for (int j = 0; j < NV; ++j) {
          const int l = lid[t[(int64_t)j * E + e]];
          if (l < R) S.used[l] |= (1ull << c);
        }
      }
      if (ok)
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int32_t a, int32_t b) { return S.color[a] < S.color[b]; });
    }
    // rows of the partition; a row that loses an entry to eliminate_zeros
    // (a pruned off-diagonal pair of one of its elements, or its diagonal)
    // gets one spare LDS slot after its CSR slots
    auto in_row = [&](int32_t g, int32_t col) {
      const int32_t* s0 = indices + indptr[g];
      const int32_t* f0 = indices + indptr[g + 1];
      const int32_t* it = std::lower_bound(s0, f0, col);
      return it != f0 && *it == col;
    };
    std::vector<char> spare(R, 0);
    for (int32_t e : elems)
      for (int a = 0; a < NV; ++a) {
        const int32_t ga = t[(int64_t)a * E + e];
        const int la = lid[ga];
        if (la >= R || spare[la]) continue;
        for (int bb = 0; bb < NV; ++bb)
          if (!in_row(ga, t[(int64_t)bb * E + e])) spare[la] = 1;
      }
    int lds = 0;
    for (int r = 0; r < R; ++r) {
      const int32_t g = owned[r];
      const int b = indptr[g], en = indptr[g + 1];
      const int len = en - b;
      int diag = len;  // pruned diagonal: the spare slot
      for (int k = b; k < en; ++k)
        if (indices[k] == g) diag = k - b;
      O.rows.push_back(b);
      // base (16 bits) | diagonal offset (8) | length (7) | spare slot (1)
      // rows plans: no spare slot (the kernel finds the columns by search
      // and flags a nonzero entry outside the pattern)
      const uint32_t sp = mode == 0 ? (uint32_t)spare[r] : 0u;
      O.rows.push_back((int32_t)((uint32_t)lds | ((uint32_t)diag << 16) |
                                 ((uint32_t)len << 24) | (sp << 31)));
      lds += len + (int)sp;
      O.pruned = O.pruned || spare[r];
    }
    if (lds > 65535) {
      release();
      O.err = FEMHIP_EUNSUPPORTED;
      O.msg = "fused plan: a block exceeds 65535 LDS slots";
      return;
    }
    int32_t visit = 0;
    for (int32_t q : ord) {
      const int32_t e = elems[q];
      uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int32_t gv[10];
      int lv[10];
      for (int a = 0; a < NV; ++a) {
        gv[a] = t[(int64_t)a * E + e];
        lv[a] = lid[gv[a]];
        put_bits(w, a * idbits, idbits, (uint32_t)lv[a]);
      }
      if (mode == 1) {
        // one row record per owned DOF a: (visit << 4) | a, then 7-bit
        // offsets, four per word: column b in row a (field b < NV), and
        // column a in row b for every owned b != a (field NV + b; 127: the
        // pair is not in the pattern)
        auto offset_in = [&](int32_t g, int32_t col) -> uint32_t {
          const int32_t* s0 = indices + indptr[g];
          const int32_t* f0 = indices + indptr[g + 1];
          const int32_t* it = std::lower_bound(s0, f0, col);
          return (it != f0 && *it == col) ? (uint32_t)(it - s0) : 127u;
        };
        for (int a = 0; a < NV; ++a) {
          if (lv[a] >= R) continue;
          uint32_t rw[kRowRecWords] = {};
          rw[0] = ((uint32_t)visit << 4) | (uint32_t)a;
          // field k: word 1 + k / 4, bits 7 (k % 4): no field straddles a word
          auto put_field = [&](int k, uint32_t v) {
            put_bits(rw + 1, 32 * (k >> 2) + 7 * (k & 3), 7, v);
          };
          for (int b = 0; b < NV; ++b) {
            put_field(b, offset_in(gv[a], gv[b]));
            put_field(NV + b, (b != a && lv[b] < R) ? offset_in(gv[b], gv[a]) : 127u);
          }
          // the local row of DOF a in the top nibbles of offset words 1-4
          for (int q = 0; q < 4; ++q) rw[1 + q] |= (((uint32_t)lv[a] >> (4 * q)) & 15u) << 28;
          // sort key (not stored): the trial functions the record computes
          uint32_t need = 0;
          for (int b = 0; b < NV; ++b)
            if (lv[b] >= R || b >= a) need |= 1u << b;
          O.rowkeys.push_back(((uint32_t)a << 16) | need);
          O.rowrecs.insert(O.rowrecs.end(), rw, rw + kRowRecWords);
        }
        // rows plans: the element id follows the local ids in the record
        for (int k = 0; k < rec_words; ++k) O.recs.push_back(w[k]);
        O.recs.push_back((uint32_t)e);
        O.rec_elem.push_back(e);
        ++visit;
        continue;
      }
      int bitpos = NV * idbits;
      for (int a = 0; a < NV; ++a)
        for (int b = 0; b < NV; ++b) {
          if (b == a) continue;
          // the pair's offset in row a's CSR slots, for halo rows too (the
          // adjoint reads their upstream values from global memory); a
          // pruned pair: the row's spare slot (offset = row length)
          const int32_t g = gv[a];
          const int32_t* s = indices + indptr[g];
          const int32_t* f = indices + indptr[g + 1];
          const int32_t* it = std::lower_bound(s, f, gv[b]);
          const uint32_t off =
              (it != f && *it == gv[b]) ? (uint32_t)(it - s) : (uint32_t)(f - s);
          put_bits(w, bitpos, offbits, off);
          bitpos += offbits;
        }
      for (int k = 0; k < rec_words; ++k) O.recs.push_back(w[k]);
      O.rec_elem.push_back(e);
    }
    if (mode == 1) {
      // row records ordered by (local DOF index, trial functions computed):
      // a wave's lanes then mostly run the same entries
      const size_t n = O.rowrecs.size() / kRowRecWords;
      std::vector<uint32_t> idx(n);
      for (size_t k = 0; k < n; ++k) idx[k] = (uint32_t)k;
      std::stable_sort(idx.begin(), idx.end(),
                       [&](uint32_t x, uint32_t y) { return O.rowkeys[x] < O.rowkeys[y]; });
      std::vector<uint32_t> sorted(O.rowrecs.size());
      for (size_t k = 0; k < n; ++k)
        std::copy_n(O.rowrecs.begin() + (size_t)idx[k] * kRowRecWords, kRowRecWords,
                    sorted.begin() + k * kRowRecWords);
      O.rowrecs.swap(sorted);
      std::vector<uint32_t>().swap(O.rowkeys);
    }
    O.nodemap.assign(locals.begin(), locals.end());
    release();
    O.L = L;
    O.R = R;
    O.lds = lds;
    O.nelems = (int32_t)elems.size();
  };
  {
    std::vector<std::unique_ptr<Scratch>> scratch(nth);
    parallel_for(leaves.size(), [&](size_t k, int th) {
      if (!scratch[th]) scratch[th] = std::make_unique<Scratch>(Nv, E);
      build_leaf(k, *scratch[th], outs[k]);
    });
  }
  for (const LeafOut& O : outs)
    if (O.err) {
      set_error(O.msg);
      return O.err;
    }

  Plan* P = new Plan();
  std::memset(&P->info, 0, sizeof(P->info));
  int max_slots = 0, max_nodes = 0, max_rows = 0, max_recs = 0;
  int64_t visits = 0;
  bool pruned = false;
  // every block's offsets in the concatenated arrays, then the copies in
  // parallel
  const size_t nl = outs.size();
  std::vector<int64_t> o_node(nl + 1, 0), o_rows(nl + 1, 0), o_recs(nl + 1, 0),
      o_elem(nl + 1, 0), o_rr(nl + 1, 0);
  for (size_t leaf = 0; leaf < nl; ++leaf) {
    const LeafOut& O = outs[leaf];
    o_node[leaf + 1] = o_node[leaf] + (int64_t)O.nodemap.size();
    o_rows[leaf + 1] = o_rows[leaf] + (int64_t)O.rows.size();
    o_recs[leaf + 1] = o_recs[leaf] + (int64_t)O.recs.size();
    o_elem[leaf + 1] = o_elem[leaf] + (int64_t)O.rec_elem.size();
    o_rr[leaf + 1] = o_rr[leaf] + (int64_t)O.rowrecs.size();
    pruned = pruned || O.pruned;
    max_slots = std::max(max_slots, O.lds);
    max_nodes = std::max(max_nodes, O.L);
    max_rows = std::max(max_rows, O.R);
    max_recs = std::max(max_recs, O.nelems);
    visits += O.nelems;
  }
  if (o_node[nl] >= INT32_MAX || o_rows[nl] / 2 >= INT32_MAX ||
      o_rr[nl] / kRowRecWords >= INT32_MAX) {
    delete P;
    set_error("fused plan: more than 2^31-1 local nodes, rows or row records");
    return FEMHIP_EOVERFLOW;
  }
  P->nodemap.resize(o_node[nl]);
  P->rows.resize(o_rows[nl]);
  P->recs.resize(o_recs[nl]);
  P->rec_elem.resize(o_elem[nl]);
  P->rowrecs.resize(o_rr[nl]);
  P->parts.resize(8 * nl);
  parallel_for(nl, [&](size_t leaf, int) {
    LeafOut& O = outs[leaf];
    const int32_t node_start = (int32_t)o_node[leaf];
    const int32_t row_start = (int32_t)(o_rows[leaf] / 2);
    const int32_t rec_start = (int32_t)o_elem[leaf];
    std::copy(O.nodemap.begin(), O.nodemap.end(), P->nodemap.begin() + o_node[leaf]);
    std::copy(O.rows.begin(), O.rows.end(), P->rows.begin() + o_rows[leaf]);
    std::copy(O.recs.begin(), O.recs.end(), P->recs.begin() + o_recs[leaf]);
    std::copy(O.rec_elem.begin(), O.rec_elem.end(), P->rec_elem.begin() + o_elem[leaf]);
    int32_t* hdr = P->parts.data() + 8 * leaf;
    if (mode == 0) {
      const int32_t h[8] = {node_start, O.L, row_start, O.R, O.lds, rec_start, O.nelems, 0};
      std::copy(h, h + 8, hdr);
    } else {
      // rows plans: [5] first row record, [6] row records, [7] first visit;
      // the records' visit indices rebased to the concatenated visits
      const int32_t rr_start = (int32_t)(o_rr[leaf] / kRowRecWords);
      const size_t n = O.rowrecs.size() / kRowRecWords;
      uint32_t* dst = P->rowrecs.data() + o_rr[leaf];
      for (size_t k = 0; k < n; ++k) {
        const uint32_t rr = O.rowrecs[k * kRowRecWords];
        dst[k * kRowRecWords] = (((rr >> 4) + (uint32_t)rec_start) << 4) | (rr & 15u);
        std::copy(O.rowrecs.begin() + k * kRowRecWords + 1,
                  O.rowrecs.begin() + (k + 1) * kRowRecWords, dst + k * kRowRecWords + 1);
      }
      const int32_t h[8] = {node_start, O.L, row_start, O.R, O.lds, rr_start, (int32_t)n,
                            rec_start};
      std::copy(h, h + 8, hdr);
    }
    // release as we go
    std::vector<int32_t>().swap(O.nodemap);
    std::vector<int32_t>().swap(O.rows);
    std::vector<uint32_t>().swap(O.recs);
    std::vector<int32_t>().swap(O.rec_elem);
    std::vector<uint32_t>().swap(O.rowrecs);
  });
  if (visits >= INT32_MAX || (mode == 1 && visits >= (1ll << 28))) {
    delete P;
    set_error("fused plan: more than 2^31-1 element visits");
    return FEMHIP_EOVERFLOW;
  }
  femhip_fused_plan_info_t& I = P->info;
  I.nparts = (int32_t)outs.size();
  I.nrecs = (int64_t)P->rec_elem.size();
  I.max_recs = max_recs;
  I.nlocal = (int64_t)P->nodemap.size();
  I.nrows = (int32_t)(P->rows.size() / 2);
  I.max_slots = max_slots;
  I.max_nodes = max_nodes;
  I.max_rows = max_rows;
  I.rec_words = mode == 1 ? rec_words + 1 : rec_words;
  I.idbits = idbits;
  I.offbits = offbits;
  I.block = block;
  I.pruned = pruned ? 1 : 0;
  I.nrowrecs = (int32_t)(P->rowrecs.size() / kRowRecWords);
  I.dup = E > 0 ? (double)visits / (double)E : 0.0;
  *info_out = I;
  *handle_out = P;
  return FEMHIP_OK;
}

}  // namespace

extern "C" int femhip_fused_plan_build(int32_t nverts, int32_t dim, int64_t E, int64_t Nv,
                                       const int32_t* t, const double* p,
                                       const int32_t* indptr, const int32_t* indices,
                                       int32_t slot_cap, int32_t node_cap, int32_t block,
                                       int32_t order_mode, void** handle_out,
                                       femhip_fused_plan_info_t* info_out) {
  return build_impl(0, nverts, dim, E, Nv, t, p, indptr, indices, slot_cap, node_cap, block,
                    order_mode, handle_out, info_out);
}

extern "C" int femhip_rows_fused_plan_build(int32_t nverts, int32_t dim, int64_t E, int64_t N,
                                            const int32_t* t, const double* p,
                                            const int32_t* indptr, const int32_t* indices,
                                            int32_t slot_cap, int32_t node_cap, int32_t block,
                                            void** handle_out,
                                            femhip_fused_plan_info_t* info_out) {
  return build_impl(1, nverts, dim, E, N, t, p, indptr, indices, slot_cap, node_cap, block, 1,
                    handle_out, info_out);
}

extern "C" int femhip_rows_fused_plan_export(void* handle, int32_t* rowrecs) {
  if (!handle) {
    femhip::set_error("rows plan: NULL handle");
    return FEMHIP_EINVAL;
  }
  Plan* P = static_cast<Plan*>(handle);
  if (rowrecs && !P->rowrecs.empty())  // kRowRecWords words per record
    std::memcpy(rowrecs, P->rowrecs.data(), P->rowrecs.size() * 4);
  return FEMHIP_OK;
}

extern "C" int femhip_fused_plan_export(void* handle, int32_t* parts, int32_t* nodemap,
                                        int32_t* rows, uint32_t* recs, int32_t* rec_elem) {
  if (!handle) {
    femhip::set_error("fused plan: NULL handle");
    return FEMHIP_EINVAL;
  }
  Plan* P = static_cast<Plan*>(handle);
  auto cp = [](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  cp(parts, P->parts.data(), P->parts.size() * 4);
  cp(nodemap, P->nodemap.data(), P->nodemap.size() * 4);
  cp(rows, P->rows.data(), P->rows.size() * 4);
  cp(recs, P->recs.data(), P->recs.size() * 4);
  cp(rec_elem, P->rec_elem.data(), P->rec_elem.size() * 4);
  return FEMHIP_OK;
}

extern "C" void femhip_fused_plan_free(void* handle) { delete static_cast<Plan*>(handle); }
