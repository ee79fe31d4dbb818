// Host-side builder of the layer-sweep plan of the fused P1 kernel.
//
// Rows are owned by blocks (weighted recursive coordinate bisection, about
// `block_rows` rows each, spatially compact); a workgroup assembles its
// block by sweeping the block's nodes level by level, where the level of a
// node is its breadth-first distance (over element adjacency) from the
// lowest nodes of the mesh.  The vertices of an element are pairwise
// adjacent, so an element spans at most two consecutive levels; it is
// processed at step k = the lowest level among its vertices and touches
// rows of levels k and k+1 only.  Rows of level k are complete after step
// k.  The workgroup therefore keeps two level buffers of row accumulators
// and of node coordinates in LDS (a ring of two), whatever the block size.
//
// Determinism: within a step, elements are greedily coloured so that two
// elements of a colour never share an owned vertex, sorted by colour and cut
// into chunks of at most `chunk` elements spanning at most 16 colours; the
// kernel applies a chunk's colours in barrier-separated rounds.
//
// Outputs (host vectors, exported to caller buffers):
//   blocks : 4 int32 per block  {level_start, nlevels, step_chunk_start, 0}
//   levels : 6 int32 per (block, level) {node_off, nnodes, row_off, nrows,
//            nslots, chunk_off}  -- nodes of the level in this block's window
//            (owned rows first), its owned rows, and the chunks of the step
//            that starts at this level; chunk_off of level i+1 closes step i
//   nodemap: global node id per window node
//   rows   : 2 int32 per owned row {CSR start, lds_base | diag << 16 | len << 24}
//   chunks : 2 int32 per chunk {record start, count | nrounds << 16}
//   recs   : packed records: 4 local ids (parity bit + index), 12 column
//            offsets (4 or 5 bits, all-ones = skip), 4-bit relative colour
//   rec_elem: global element id per record
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/femhip.h"

namespace femhip {
void set_error(const std::string& msg);
}

namespace {

struct SweepPlan {
  femhip_sweep_plan_info_t info;
  std::vector<int32_t> blocks, levels, nodemap, rows, chunks, rec_elem;
  std::vector<uint32_t> recs;
};

void put_bits(uint32_t* w, int pos, int nbits, uint32_t v) {
  for (int i = 0; i < nbits; ++i)
    if ((v >> i) & 1u) w[(pos + i) >> 5] |= (1u << ((pos + i) & 31));
}

}  // namespace

extern "C" int femhip_sweep_plan_build(int32_t nverts, int32_t dim, int64_t E, int64_t Nv,
                                       const int32_t* t, const double* p,
                                       const int32_t* indptr, const int32_t* indices,
                                       int32_t block_rows, int32_t chunk, void** handle_out,
                                       femhip_sweep_plan_info_t* info_out) {
  using femhip::set_error;
  if (!(nverts == 3 && dim == 2) && !(nverts == 4 && dim == 3)) {
    set_error("sweep plan: P1 triangles or tetrahedra only");
    return FEMHIP_EUNSUPPORTED;
  }
  if (!t || !p || !indptr || !indices || !handle_out || !info_out || E < 0 || Nv <= 0 ||
      block_rows <= 0 || chunk <= 0 || chunk > 1024) {
    set_error("sweep plan: bad arguments");
    return FEMHIP_EINVAL;
  }
  const int NV = nverts;
  int max_row = 0;
  for (int64_t n = 0; n < Nv; ++n) max_row = std::max(max_row, indptr[n + 1] - indptr[n]);
  const int offbits = max_row <= 15 ? 4 : (max_row <= 31 ? 5 : 0);
  if (offbits == 0) {
    set_error("sweep plan: CSR rows longer than 31 entries");
    return FEMHIP_EUNSUPPORTED;
  }
  const uint32_t sentinel = (1u << offbits) - 1;
  // node -> element incidence
  std::vector<int64_t> inc_ptr(Nv + 1, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int k = 0; k < NV; ++k) {
      const int32_t v = t[(int64_t)k * E + e];
      if (v < 0 || v >= Nv) {
        set_error("sweep plan: connectivity index out of range");
        return FEMHIP_EINVAL;
      }
      inc_ptr[v + 1]++;
    }
  for (int64_t n = 0; n < Nv; ++n) inc_ptr[n + 1] += inc_ptr[n];
  std::vector<int32_t> inc(inc_ptr[Nv]);
  {
    std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int k = 0; k < NV; ++k) inc[fill[t[(int64_t)k * E + e]]++] = (int32_t)e;
  }
  // 1. BFS levels from the lowest nodes (last coordinate), per component
  std::vector<int32_t> level(Nv, -1);
  {
    const int zc = dim - 1;
    double zmin = 1e300, zmax = -1e300;
    for (int64_t n = 0; n < Nv; ++n) {
      zmin = std::min(zmin, p[zc * Nv + n]);
      zmax = std::max(zmax, p[zc * Nv + n]);
    }
    const double tol = 1e-9 * std::max(1.0, zmax - zmin);
    std::vector<int32_t> front, next;
    for (int64_t n = 0; n < Nv; ++n)
      if (p[zc * Nv + n] <= zmin + tol && inc_ptr[n + 1] > inc_ptr[n]) {
        level[n] = 0;
        front.push_back((int32_t)n);
      }
    std::vector<int32_t> by_z(Nv);
    for (int64_t n = 0; n < Nv; ++n) by_z[n] = (int32_t)n;
    std::sort(by_z.begin(), by_z.end(), [&](int32_t a, int32_t b) {
      return p[zc * Nv + a] < p[zc * Nv + b];
    });
    size_t seed_pos = 0;
    int32_t base = 0;
    while (true) {
      int32_t lv = base;
      while (!front.empty()) {
        next.clear();
        for (int32_t n : front)
          for (int64_t q = inc_ptr[n]; q < inc_ptr[n + 1]; ++q)
            for (int j = 0; j < NV; ++j) {
              const int32_t v = t[(int64_t)j * E + inc[q]];
              if (level[v] < 0) {
                level[v] = lv + 1;
                next.push_back(v);
              }
            }
        front.swap(next);
        ++lv;
      }
      // next unvisited component (or isolated node)
      while (seed_pos < by_z.size() && level[by_z[seed_pos]] >= 0) ++seed_pos;
      if (seed_pos >= by_z.size()) break;
      base = lv + 2;
      level[by_z[seed_pos]] = base;
      front.assign(1, by_z[seed_pos]);
    }
  }
  // 2. blocks: weighted recursive coordinate bisection into ~block_rows rows
  std::vector<int32_t> order(Nv);
  for (int64_t n = 0; n < Nv; ++n) order[n] = (int32_t)n;
  std::vector<std::pair<int64_t, int64_t>> leaves;
  {
    std::vector<std::pair<int64_t, int64_t>> stack{{0, Nv}};
    while (!stack.empty()) {
      auto [b, e] = stack.back();
      stack.pop_back();
      if (e - b <= block_rows) {
        leaves.push_back({b, e});
        continue;
      }
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int64_t k = b; k < e; ++k)
        for (int d = 0; d < dim; ++d) {
          const double v = p[d * Nv + order[k]];
          lo[d] = std::min(lo[d], v);
          hi[d] = std::max(hi[d], v);
        }
      // split across the sweep direction first: keep blocks tall in levels
      int ax = 0;
      for (int d = 1; d < dim; ++d)
        if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
      const int64_t nleaf = (e - b + block_rows - 1) / block_rows;
      const int64_t mid = b + (int64_t)((double)(e - b) * (double)(nleaf / 2) / (double)nleaf);
      std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                       [&](int32_t x, int32_t y) {
                         const double px = p[ax * Nv + x], py = p[ax * Nv + y];
                         return px < py || (px == py && x < y);
                       });
      stack.push_back({mid, e});
      stack.push_back({b, mid});
    }
  }
  // 3. per block: steps, window nodes, rows, colouring, records
  SweepPlan* P = new SweepPlan();
  std::memset(&P->info, 0, sizeof(P->info));
  std::vector<int32_t> owner(Nv, -1);
  for (size_t bl = 0; bl < leaves.size(); ++bl)
    for (int64_t k = leaves[bl].first; k < leaves[bl].second; ++k) owner[order[k]] = (int32_t)bl;
  std::vector<int32_t> estamp(E, -1), nstamp(Nv, -1), widx(Nv, -1);
  int max_lnodes = 0, max_lrows = 0, max_lslots = 0, max_colors = 0;
  int64_t visits = 0;
  const int idx_cap = 1024;  // window index (10 bits) + parity bit
  for (size_t bl = 0; bl < leaves.size(); ++bl) {
    const int32_t bid = (int32_t)bl;
    // elements touching owned rows, grouped by step = min level
    std::vector<int32_t> elems;
    int32_t kmin = INT32_MAX, kmax = INT32_MIN;
    for (int64_t k = leaves[bl].first; k < leaves[bl].second; ++k) {
      const int32_t n = order[k];
      for (int64_t q = inc_ptr[n]; q < inc_ptr[n + 1]; ++q) {
        const int32_t e = inc[q];
        if (estamp[e] == bid) continue;
        estamp[e] = bid;
        elems.push_back(e);
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for (int j = 0; j < NV; ++j) {
          const int32_t l = level[t[(int64_t)j * E + e]];
          lo = std::min(lo, l);
          hi = std::max(hi, l);
        }
        if (hi - lo > 1) {
          delete P;
          set_error("sweep plan: element spans more than two levels");
          return FEMHIP_EUNSUPPORTED;
        }
        kmin = std::min(kmin, lo);
        kmax = std::max(kmax, hi);
      }
      kmin = std::min(kmin, level[n]);
      kmax = std::max(kmax, level[n]);
    }
    auto step_of = [&](int32_t e) {
      int32_t lo = INT32_MAX;
      for (int j = 0; j < NV; ++j) lo = std::min(lo, level[t[(int64_t)j * E + e]]);
      return lo;
    };
    const int nl = kmax - kmin + 1;
    std::vector<std::vector<int32_t>> step_elems(nl);
    for (int32_t e : elems) step_elems[step_of(e) - kmin].push_back(e);
    // window node lists per level: owned rows first, then halo nodes
    std::vector<std::vector<int32_t>> lnodes(nl);
    for (int64_t k = leaves[bl].first; k < leaves[bl].second; ++k) {
      const int32_t n = order[k];
      lnodes[level[n] - kmin].push_back(n);
      nstamp[n] = bid;
    }
    std::vector<int> lrows(nl);
    for (int l = 0; l < nl; ++l) lrows[l] = (int)lnodes[l].size();
    for (int32_t e : elems)
      for (int j = 0; j < NV; ++j) {
        const int32_t v = t[(int64_t)j * E + e];
        if (nstamp[v] != bid) {
          nstamp[v] = bid;
          lnodes[level[v] - kmin].push_back(v);
        }
      }
    const int32_t level_start = (int32_t)(P->levels.size() / 6);
    for (int l = 0; l < nl; ++l) {
      const int nn = (int)lnodes[l].size();
      if (nn > idx_cap) {
        delete P;
        set_error("sweep plan: a level window exceeds 1024 nodes; use smaller blocks");
        return FEMHIP_EUNSUPPORTED;
      }
      for (int i = 0; i < nn; ++i) widx[lnodes[l][i]] = i;
      const int32_t node_off = (int32_t)P->nodemap.size();
      P->nodemap.insert(P->nodemap.end(), lnodes[l].begin(), lnodes[l].end());
      const int32_t row_off = (int32_t)(P->rows.size() / 2);
      int lds = 0;
      for (int r = 0; r < lrows[l]; ++r) {
        const int32_t g = lnodes[l][r];
        const int b = indptr[g], en = indptr[g + 1];
        int dg = 255;
        for (int k = b; k < en; ++k)
          if (indices[k] == g) dg = k - b;
        P->rows.push_back(b);
        P->rows.push_back((int32_t)((uint32_t)lds | ((uint32_t)dg << 16) |
                                    ((uint32_t)(en - b) << 24)));
        lds += en - b;
      }
      if (lds > 65535) {
        delete P;
        set_error("sweep plan: a level holds more than 65535 CSR slots");
        return FEMHIP_EUNSUPPORTED;
      }
      P->levels.insert(P->levels.end(), {node_off, nn, row_off, lrows[l], lds, 0});
      max_lnodes = std::max(max_lnodes, nn);
      max_lrows = std::max(max_lrows, lrows[l]);
      max_lslots = std::max(max_lslots, lds);
    }
    // steps: colour, sort, chunk, records
    for (int l = 0; l < nl; ++l) {
      P->levels[(size_t)(level_start + l) * 6 + 5] = (int32_t)(P->chunks.size() / 2);
      std::vector<int32_t>& se = step_elems[l];
      if (se.empty()) continue;
      // colours: owned vertices are identified by (level, window index)
      std::vector<int32_t> color(se.size());
      std::vector<uint64_t> used_a(lrows[l], 0), used_b(l + 1 < nl ? lrows[l + 1] : 0, 0);
      int ncol = 0;
      for (size_t i = 0; i < se.size(); ++i) {
        const int32_t e = se[i];
        uint64_t forbid = 0;
        for (int j = 0; j < NV; ++j) {
          const int32_t v = t[(int64_t)j * E + e];
          if (owner[v] != bid) continue;
          forbid |= (level[v] - kmin == l) ? used_a[widx[v]] : used_b[widx[v]];
        }
        if (~forbid == 0) {
          delete P;
          set_error("sweep plan: more than 64 colours needed");
          return FEMHIP_EUNSUPPORTED;
        }
        const int c = __builtin_ctzll(~forbid);
        color[i] = c;
        ncol = std::max(ncol, c + 1);
        for (int j = 0; j < NV; ++j) {
          const int32_t v = t[(int64_t)j * E + e];
          if (owner[v] != bid) continue;
          ((level[v] - kmin == l) ? used_a[widx[v]] : used_b[widx[v]]) |= (1ull << c);
        }
      }
      max_colors = std::max(max_colors, ncol);
      std::vector<int32_t> ord(se.size());
      for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int32_t)i;
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int32_t a, int32_t b) { return color[a] < color[b]; });
      size_t i = 0;
      while (i < ord.size()) {
        const int c0 = color[ord[i]];
        size_t j = i;
        while (j < ord.size() && (int)(j - i) < chunk && color[ord[j]] - c0 < 16) ++j;
        P->chunks.push_back((int32_t)P->rec_elem.size());
        P->chunks.push_back((int32_t)((j - i) | ((size_t)(color[ord[j - 1]] - c0 + 1) << 16)));
        for (size_t q = i; q < j; ++q) {
          const int32_t e = se[ord[q]];
          uint32_t w[4] = {0, 0, 0, 0};
          int32_t gv[4];
          for (int a = 0; a < NV; ++a) {
            gv[a] = t[(int64_t)a * E + e];
            const uint32_t par = (uint32_t)((level[gv[a]] - kmin) & 1);
            put_bits(w, a * 11, 11, (par << 10) | (uint32_t)widx[gv[a]]);
          }
          int bitpos = NV * 11;
          for (int a = 0; a < NV; ++a)
            for (int b = 0; b < NV; ++b) {
              if (b == a) continue;
              uint32_t off = sentinel;
              if (owner[gv[a]] == bid) {
                const int32_t* s = indices + indptr[gv[a]];
                const int32_t* f = indices + indptr[gv[a] + 1];
                const int32_t* it = std::lower_bound(s, f, gv[b]);
                if (it != f && *it == gv[b]) off = (uint32_t)(it - s);
              }
              put_bits(w, bitpos, offbits, off);
              bitpos += offbits;
            }
          put_bits(w, bitpos, 4, (uint32_t)(color[ord[q]] - c0));
          const int words = (bitpos + 4 + 31) / 32;
          for (int k = 0; k < words; ++k) P->recs.push_back(w[k]);
          P->rec_elem.push_back(e);
        }
        visits += (int64_t)(j - i);
        i = j;
      }
    }
    P->blocks.insert(P->blocks.end(), {level_start, nl, 0, 0});
  }
  // sentinel level entry closing the last step of the last block
  P->levels.insert(P->levels.end(), {(int32_t)P->nodemap.size(), 0, (int32_t)(P->rows.size() / 2),
                                     0, 0, (int32_t)(P->chunks.size() / 2)});
  femhip_sweep_plan_info_t& I = P->info;
  I.nblocks = (int32_t)leaves.size();
  I.nlevels = (int32_t)(P->levels.size() / 6);
  I.nchunks = (int32_t)(P->chunks.size() / 2);
  I.nrows = (int32_t)(P->rows.size() / 2);
  I.max_level_nodes = max_lnodes;
  I.max_level_rows = max_lrows;
  I.max_level_slots = max_lslots;
  I.max_colors = max_colors;
  I.rec_words = (NV * 11 + NV * (NV - 1) * offbits + 4 + 31) / 32;
  I.offbits = offbits;
  I.chunk = chunk;
  I.nrecs = (int64_t)P->rec_elem.size();
  I.nlocal = (int64_t)P->nodemap.size();
  I.dup = E > 0 ? (double)visits / (double)E : 0.0;
  *info_out = I;
  *handle_out = P;
  return FEMHIP_OK;
}

extern "C" int femhip_sweep_plan_export(void* handle, int32_t* blocks, int32_t* levels,
                                        int32_t* nodemap, int32_t* rows, int32_t* chunks,
                                        uint32_t* recs, int32_t* rec_elem) {
  if (!handle) {
    femhip::set_error("sweep plan: NULL handle");
    return FEMHIP_EINVAL;
  }
  SweepPlan* P = static_cast<SweepPlan*>(handle);
  auto cp = [](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  cp(blocks, P->blocks.data(), P->blocks.size() * 4);
  cp(levels, P->levels.data(), P->levels.size() * 4);
  cp(nodemap, P->nodemap.data(), P->nodemap.size() * 4);
  cp(rows, P->rows.data(), P->rows.size() * 4);
  cp(chunks, P->chunks.data(), P->chunks.size() * 4);
  cp(recs, P->recs.data(), P->recs.size() * 4);
  cp(rec_elem, P->rec_elem.data(), P->rec_elem.size() * 4);
  return FEMHIP_OK;
}

extern "C" void femhip_sweep_plan_free(void* handle) { delete static_cast<SweepPlan*>(handle); }
