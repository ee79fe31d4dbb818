// Symbolic phase: the canonical CSR pattern of the assembled matrix and
// the reduction plan that maps every CSR slot to its contributions.
//
// SciPy's coo_matrix -> eliminate_zeros -> tocsr (coo_data.py:28-36)
// drops explicit zeros per contribution, then sorts by (row, col) and sums
// duplicates; the result is canonical (sorted unique int32 column indices
// per row).  Here the same pattern is built from integer keys
// key = row * ncols + col (int64): contributions whose exact local value is
// 0.0 get a sentinel key, a stable LSD radix sort orders the rest, head
// flags + an integer scan number the unique keys, and one pass writes
// indices, indptr and the per-slot contribution ranges.  The sort is
// stable, so within a slot contributions stay in the reference's COO order
// and the numeric reduction (femhip_segment_sum) is deterministic.
#include <cstring>
#include <string>

#include <rocprim/rocprim.hpp>

#include "common.h"

namespace femhip {
namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock)
make_keys_kernel(const int32_t* __restrict__ row_dofs, const int32_t* __restrict__ col_dofs,
                 int nb_row, int64_t E, int64_t ncols, int64_t C, uint64_t sentinel,
                 const double* __restrict__ contrib, uint64_t* __restrict__ keys,
                 int32_t* __restrict__ vals) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= C) return;
  const int64_t entry = c / E;
  const int64_t e = c - entry * E;
  const int64_t j = entry / nb_row;  // trial (column) local index
  const int64_t i = entry - j * nb_row;  // test (row) local index
  const int64_t r = row_dofs[i * E + e];
  const int64_t col = col_dofs[j * E + e];
  const bool keep = (contrib == nullptr) || (contrib[c] != 0.0);
  keys[c] = keep ? (uint64_t)(r * ncols + col) : sentinel;
  vals[c] = (int32_t)c;
}

__global__ void __launch_bounds__(kBlock)
make_row_keys_kernel(const int32_t* __restrict__ row_dofs, int64_t E, int64_t C,
                     uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= C) return;
  // contribution c = i * E + e reads element_dofs[i][e] = row_dofs[c]
  keys[c] = (uint64_t)row_dofs[c];
  vals[c] = (int32_t)c;
}

__global__ void __launch_bounds__(kBlock)
head_flags_kernel(const uint64_t* __restrict__ keys, int64_t C, uint64_t sentinel,
                  int32_t* __restrict__ flags) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint64_t key = keys[k];
  flags[k] = (key != sentinel && (k == 0 || keys[k - 1] != key)) ? 1 : 0;
}

// For every head: the slot's column index and the start of its range; the
// end of the last range is the first sentinel position.
__global__ void __launch_bounds__(kBlock)
fill_slots_kernel(const uint64_t* __restrict__ keys, const int32_t* __restrict__ scan,
                  int64_t C, uint64_t sentinel, int64_t ncols,
                  int32_t* __restrict__ indices, int32_t* __restrict__ seg,
                  int32_t* __restrict__ slot_row) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint64_t key = keys[k];
  if (key == sentinel) return;
  const bool head = (k == 0 || keys[k - 1] != key);
  if (head) {
    const int32_t s = scan[k] - 1;
    indices[s] = (int32_t)(key % (uint64_t)ncols);
    slot_row[s] = (int32_t)(key / (uint64_t)ncols);
    seg[s] = (int32_t)k;
  }
  if (k == C - 1 || keys[k + 1] == sentinel) seg[scan[k]] = (int32_t)(k + 1);
}

// indptr[r] = first slot whose row is >= r.
__global__ void __launch_bounds__(kBlock)
fill_indptr_kernel(const int32_t* __restrict__ slot_row, int64_t nnz, int64_t nrows,
                   int32_t* __restrict__ indptr) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= nnz) return;
  const int64_t r = slot_row[s];
  const int64_t rp = (s == 0) ? -1 : slot_row[s - 1];
  for (int64_t rr = rp + 1; rr <= r; ++rr) indptr[rr] = (int32_t)s;
  if (s == nnz - 1)
    for (int64_t rr = r + 1; rr <= nrows; ++rr) indptr[rr] = (int32_t)nnz;
}

// seg[r] = first sorted position whose row key is >= r (rows plan).
__global__ void __launch_bounds__(kBlock)
fill_row_seg_kernel(const uint64_t* __restrict__ keys, int64_t C, int64_t nrows,
                    int32_t* __restrict__ seg) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const int64_t r = (int64_t)keys[k];
  const int64_t rp = (k == 0) ? -1 : (int64_t)keys[k - 1];
  for (int64_t rr = rp + 1; rr <= r; ++rr) seg[rr] = (int32_t)k;
  if (k == C - 1)
    for (int64_t rr = r + 1; rr <= nrows; ++rr) seg[rr] = (int32_t)C;
}

__global__ void __launch_bounds__(kBlock)
segment_sum_kernel(const double* __restrict__ contrib, const int32_t* __restrict__ perm,
                   const int32_t* __restrict__ seg, int64_t nseg, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= nseg) return;
  const int32_t b = seg[k], en = seg[k + 1];
  // the gathers of up to four contributions are issued together; the sum
  // keeps its order
  double acc = 0.0;
  int32_t j = b;
  for (; j + 4 <= en; j += 4) {
    const int32_t p0 = perm[j], p1 = perm[j + 1], p2 = perm[j + 2], p3 = perm[j + 3];
    const double v0 = contrib[p0], v1 = contrib[p1], v2 = contrib[p2], v3 = contrib[p3];
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
    const double v0 = contrib[p0], v1 = contrib[p1], v2 = contrib[p2];
    acc += v0;
    if (r > 1) acc += v1;
    if (r > 2) acc += v2;
  }
  out[k] = acc;
}

// Exact-zero mask of contrib: bit (c & 63) of word c >> 6 is set iff
// contrib[c] == 0.0 -- one wave ballot per 64 contributions.  A plan keeps
// the mask of the contributions it was built from; another set of
// contributions has the same pattern when its mask is equal (compared word
// by word, exactly).
__global__ void __launch_bounds__(kBlock)
zero_mask_kernel(const double* __restrict__ contrib, int64_t n, uint64_t* __restrict__ mask) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t b = __ballot(c < n && contrib[c] == 0.0);
  if ((threadIdx.x & 63) == 0 && c < n) mask[c >> 6] = b;
}

__global__ void zero_count_kernel(unsigned long long* __restrict__ count) { *count = 0; }

// number of contributions whose zero bit differs from mask
__global__ void __launch_bounds__(kBlock)
zero_mask_mismatch_kernel(const double* __restrict__ contrib, int64_t n,
                          const uint64_t* __restrict__ mask,
                          unsigned long long* __restrict__ count) {
  unsigned long long diff = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += stride) {
    const int64_t c = base + threadIdx.x;
    const uint64_t b = __ballot(c < n && contrib[c] == 0.0);
    if ((threadIdx.x & 63) == 0 && c < n) diff += (unsigned long long)__popcll(b ^ mask[c >> 6]);
  }
  if ((threadIdx.x & 63) == 0 && diff != 0) atomicAdd(count, diff);
}

// out[0] = min, out[1] = max of idx[n] (initialised by range_init_kernel)
__global__ void range_init_kernel(int32_t* __restrict__ out) {
  out[0] = INT32_MAX;
  out[1] = INT32_MIN;
}

__global__ void __launch_bounds__(kBlock)
index_range_kernel(const int32_t* __restrict__ idx, int64_t n, int32_t* __restrict__ out) {
  int32_t lo = INT32_MAX, hi = INT32_MIN;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kBlock) {
    const int32_t v = idx[k];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(out, lo);
    atomicMax(out + 1, hi);
  }
}

// Checks 0 <= idx[k] < bound for idx[n] on the device (scratch: 2 int32 of
// device memory); synchronises the stream.
int check_index_range(const int32_t* idx, int64_t n, int64_t bound, int32_t* scratch,
                      hipStream_t st, const char* what) {
  if (n == 0) return FEMHIP_OK;
  range_init_kernel<<<1, 1, 0, st>>>(scratch);
  FEMHIP_CHECK_LAUNCH();
  int64_t blocks = ceil_div(n, kBlock);
  if (blocks > 1024) blocks = 1024;
  index_range_kernel<<<(unsigned)blocks, kBlock, 0, st>>>(idx, n, scratch);
  FEMHIP_CHECK_LAUNCH();
  int32_t h[2];
  FEMHIP_TRY_HIP(hipMemcpyAsync(h, scratch, sizeof(h), hipMemcpyDeviceToHost, st));
  FEMHIP_TRY_HIP(hipStreamSynchronize(st));
  if (h[0] < 0 || (int64_t)h[1] >= bound) {
    set_error(std::string(what) + ": index out of range [0, " + std::to_string(bound) +
              "): min " + std::to_string(h[0]) + ", max " + std::to_string(h[1]));
    return FEMHIP_EINVAL;
  }
  return FEMHIP_OK;
}

unsigned bits_for(uint64_t v) {
  unsigned b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

struct Workspace {
  uint64_t* keys_in;
  uint64_t* keys_out;
  int32_t* vals_in;
  int32_t* scan;
  int32_t* slot_row;
  void* temp;
  size_t temp_bytes;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t rocprim_temp_bytes(int64_t C) {
  size_t sort_bytes = 0, scan_bytes = 0;
  (void)rocprim::radix_sort_pairs<rocprim::default_config>(
      nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr, (int32_t*)nullptr,
      (int32_t*)nullptr, (size_t)C, 0, 64);
  (void)rocprim::inclusive_scan<rocprim::default_config>(
      nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, (size_t)C,
      rocprim::plus<int32_t>());
  return sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
}

size_t workspace_bytes(int64_t C) {
  return align256(8 * C) * 2 + align256(4 * C) * 3 + align256(rocprim_temp_bytes(C));
}

Workspace carve(void* ws, int64_t C) {
  char* b = static_cast<char*>(ws);
  Workspace w;
  w.keys_in = reinterpret_cast<uint64_t*>(b);
  b += align256(8 * C);
  w.keys_out = reinterpret_cast<uint64_t*>(b);
  b += align256(8 * C);
  w.vals_in = reinterpret_cast<int32_t*>(b);
  b += align256(4 * C);
  w.scan = reinterpret_cast<int32_t*>(b);
  b += align256(4 * C);
  w.slot_row = reinterpret_cast<int32_t*>(b);
  b += align256(4 * C);
  w.temp = b;
  w.temp_bytes = rocprim_temp_bytes(C);
  return w;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_pattern_workspace(int64_t ncontrib, size_t* bytes) {
  FEMHIP_REQUIRE(bytes != nullptr && ncontrib >= 0, FEMHIP_EINVAL, "bad arguments");
  FEMHIP_REQUIRE(ncontrib < INT32_MAX, FEMHIP_EOVERFLOW,
                 "more than 2^31-1 contributions");
  *bytes = workspace_bytes(ncontrib > 0 ? ncontrib : 1);
  return FEMHIP_OK;
}

extern "C" int femhip_pattern_build(const int32_t* row_dofs, const int32_t* col_dofs,
                                    int32_t nb_row, int32_t nb_col, int64_t nelems,
                                    int64_t nrows, int64_t ncols, const double* contrib,
                                    int32_t* indptr, int32_t* indices, int32_t* perm,
                                    int32_t* seg, int64_t* nnz_out, void* workspace,
                                    size_t ws_bytes, femhip_stream_t s) {
  FEMHIP_REQUIRE(nb_row > 0 && nb_col > 0 && nelems >= 0 && nrows >= 0 && ncols > 0,
                 FEMHIP_EINVAL, "bad sizes");
  FEMHIP_REQUIRE(nnz_out != nullptr && indptr != nullptr, FEMHIP_EINVAL, "NULL output");
  FEMHIP_REQUIRE(nrows < INT32_MAX && ncols < INT32_MAX, FEMHIP_EOVERFLOW,
                 "matrix dimension exceeds int32");
  const int64_t C = (int64_t)nb_row * nb_col * nelems;
  FEMHIP_REQUIRE(C < INT32_MAX, FEMHIP_EOVERFLOW, "more than 2^31-1 contributions");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (C == 0) {
    FEMHIP_TRY_HIP(hipMemsetAsync(indptr, 0, sizeof(int32_t) * (nrows + 1), st));
    *nnz_out = 0;
    return FEMHIP_OK;
  }
  FEMHIP_REQUIRE(workspace != nullptr && ws_bytes >= workspace_bytes(C), FEMHIP_ESIZE,
                 "workspace too small");
  FEMHIP_REQUIRE(row_dofs && col_dofs && indices && perm && seg, FEMHIP_EINVAL,
                 "NULL input/output");
  Workspace w = carve(workspace, C);
  // a DOF index outside the matrix would corrupt the sort keys and the
  // indptr fill: refuse it before any indexing kernel runs
  if (int rc = check_index_range(row_dofs, (int64_t)nb_row * nelems, nrows, w.scan, st,
                                 "pattern: row DOF");
      rc != FEMHIP_OK)
    return rc;
  if (int rc = check_index_range(col_dofs, (int64_t)nb_col * nelems, ncols, w.scan, st,
                                 "pattern: column DOF");
      rc != FEMHIP_OK)
    return rc;
  const uint64_t sentinel = (uint64_t)nrows * (uint64_t)ncols;
  const unsigned end_bit = bits_for(sentinel);
  const unsigned grid = (unsigned)ceil_div(C, kBlock);
  make_keys_kernel<<<grid, kBlock, 0, st>>>(row_dofs, col_dofs, nb_row, nelems, ncols, C,
                                            sentinel, contrib, w.keys_in, w.vals_in);
  FEMHIP_CHECK_LAUNCH();
  size_t tb = w.temp_bytes;
  FEMHIP_TRY_HIP(rocprim::radix_sort_pairs<rocprim::default_config>(
      w.temp, tb, w.keys_in, w.keys_out, w.vals_in, perm, (size_t)C, 0, end_bit, st));
  head_flags_kernel<<<grid, kBlock, 0, st>>>(w.keys_out, C, sentinel, w.vals_in);
  FEMHIP_CHECK_LAUNCH();
  tb = w.temp_bytes;
  FEMHIP_TRY_HIP(rocprim::inclusive_scan<rocprim::default_config>(
      w.temp, tb, w.vals_in, w.scan, (size_t)C, rocprim::plus<int32_t>(), st));
  int32_t nnz32 = 0;
  FEMHIP_TRY_HIP(hipMemcpyAsync(&nnz32, w.scan + (C - 1), sizeof(int32_t),
                                hipMemcpyDeviceToHost, st));
  FEMHIP_TRY_HIP(hipStreamSynchronize(st));
  const int64_t nnz = nnz32;
  FEMHIP_TRY_HIP(hipMemsetAsync(seg, 0, sizeof(int32_t), st));
  fill_slots_kernel<<<grid, kBlock, 0, st>>>(w.keys_out, w.scan, C, sentinel, ncols,
                                             indices, seg, w.slot_row);
  FEMHIP_CHECK_LAUNCH();
  if (nnz == 0) {
    FEMHIP_TRY_HIP(hipMemsetAsync(indptr, 0, sizeof(int32_t) * (nrows + 1), st));
  } else {
    fill_indptr_kernel<<<(unsigned)ceil_div(nnz, kBlock), kBlock, 0, st>>>(
        w.slot_row, nnz, nrows, indptr);
    FEMHIP_CHECK_LAUNCH();
  }
  FEMHIP_TRY_HIP(hipStreamSynchronize(st));
  *nnz_out = nnz;
  return FEMHIP_OK;
}

extern "C" int femhip_rows_plan_build(const int32_t* row_dofs, int32_t nb_row,
                                      int64_t nelems, int64_t nrows, int32_t* perm,
                                      int32_t* seg, void* workspace, size_t ws_bytes,
                                      femhip_stream_t s) {
  FEMHIP_REQUIRE(nb_row > 0 && nelems >= 0 && nrows >= 0, FEMHIP_EINVAL, "bad sizes");
  FEMHIP_REQUIRE(seg != nullptr, FEMHIP_EINVAL, "NULL seg");
  const int64_t C = (int64_t)nb_row * nelems;
  FEMHIP_REQUIRE(C < INT32_MAX, FEMHIP_EOVERFLOW, "more than 2^31-1 contributions");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  if (C == 0) {
    FEMHIP_TRY_HIP(hipMemsetAsync(seg, 0, sizeof(int32_t) * (nrows + 1), st));
    return FEMHIP_OK;
  }
  FEMHIP_REQUIRE(workspace != nullptr && ws_bytes >= workspace_bytes(C), FEMHIP_ESIZE,
                 "workspace too small");
  FEMHIP_REQUIRE(row_dofs && perm, FEMHIP_EINVAL, "NULL input/output");
  Workspace w = carve(workspace, C);
  if (int rc = check_index_range(row_dofs, C, nrows, w.scan, st, "rows plan: DOF");
      rc != FEMHIP_OK)
    return rc;
  const unsigned grid = (unsigned)ceil_div(C, kBlock);
  make_row_keys_kernel<<<grid, kBlock, 0, st>>>(row_dofs, nelems, C, w.keys_in, w.vals_in);
  FEMHIP_CHECK_LAUNCH();
  size_t tb = w.temp_bytes;
  FEMHIP_TRY_HIP(rocprim::radix_sort_pairs<rocprim::default_config>(
      w.temp, tb, w.keys_in, w.keys_out, w.vals_in, perm, (size_t)C, 0,
      bits_for((uint64_t)nrows), st));
  fill_row_seg_kernel<<<grid, kBlock, 0, st>>>(w.keys_out, C, nrows, seg);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_segment_sum(const double* contrib, const int32_t* perm,
                                  const int32_t* seg, int64_t nseg, double* out,
                                  femhip_stream_t s) {
  FEMHIP_REQUIRE(nseg >= 0, FEMHIP_EINVAL, "negative segment count");
  if (nseg == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(contrib && perm && seg && out, FEMHIP_EINVAL, "NULL argument");
  segment_sum_kernel<<<(unsigned)ceil_div(nseg, kBlock), kBlock, 0,
                       reinterpret_cast<hipStream_t>(s)>>>(contrib, perm, seg, nseg, out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_zero_mask(const double* contrib, int64_t n, uint64_t* mask,
                                femhip_stream_t s) {
  FEMHIP_REQUIRE(n >= 0, FEMHIP_EINVAL, "zero mask: negative size");
  if (n == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(contrib != nullptr && mask != nullptr, FEMHIP_EINVAL, "zero mask: NULL argument");
  zero_mask_kernel<<<(unsigned)ceil_div(n, kBlock), kBlock, 0,
                     reinterpret_cast<hipStream_t>(s)>>>(contrib, n, mask);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_zero_mask_mismatch(const double* contrib, int64_t n, const uint64_t* mask,
                                         int64_t* count, femhip_stream_t s) {
  FEMHIP_REQUIRE(n >= 0 && count != nullptr, FEMHIP_EINVAL, "zero mask: bad arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(count);
  zero_count_kernel<<<1, 1, 0, st>>>(cnt);
  FEMHIP_CHECK_LAUNCH();
  if (n == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(contrib != nullptr && mask != nullptr, FEMHIP_EINVAL, "zero mask: NULL argument");
  int64_t blocks = ceil_div(n, kBlock);
  if (blocks > 4096) blocks = 4096;
  zero_mask_mismatch_kernel<<<(unsigned)blocks, kBlock, 0, st>>>(contrib, n, mask, cnt);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}
