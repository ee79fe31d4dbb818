// Keeping the assembled system on the device for the solve:
//
//   * row/column selection of a CSR matrix (utils.condense: A[I][:, I] and
//     b[I] - A[I][:, D] @ x[D], utils.py:540-603), canonical output;
//   * y = A x, and a Jacobi-preconditioned conjugate gradient iteration
//     (the algorithm of utils.solver_iter_pcg, utils.py:121-169 with
//     build_pc_diag, utils.py:56-58).
//
// Every reduction runs in a fixed order (16 lanes per CSR row summed by a
// fixed shuffle tree, per-block partial sums, one final block), so the
// iterates are bitwise repeatable.
#include <cmath>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "common.h"

namespace femhip {
namespace {

constexpr int kBlock = 256;           // threads per block
constexpr int kRowLanes = 16;         // lanes per CSR row
constexpr int kRowsPerBlock = kBlock / kRowLanes;
constexpr int kPcgGrid = 1024;        // blocks of the PCG kernels

// sum over the 16 lanes of a row group, fixed order (xor tree)
__device__ __forceinline__ double row_sum(double v) {
#pragma unroll
  for (int o = kRowLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kRowLanes);
  return v;
}

// block sum in a fixed order: lanes of a wave by an xor tree, waves in order
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kBlock / 64; ++k) s += red[k];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------- select --
__global__ void __launch_bounds__(kBlock)
select_count_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                    const int32_t* __restrict__ rows, int64_t nsel,
                    const int32_t* __restrict__ colmap, int32_t* __restrict__ counts) {
  const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + (threadIdx.x / kRowLanes);
  const int lane = threadIdx.x % kRowLanes;
  if (r >= nsel) return;
  const int32_t g = rows[r];
  int c = 0;
  for (int32_t k = indptr[g] + lane; k < indptr[g + 1]; k += kRowLanes) c += colmap[indices[k]] >= 0;
#pragma unroll
  for (int o = kRowLanes / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, kRowLanes);
  if (lane == 0) counts[r] = c;
}

__global__ void zero_first_kernel(int32_t* p) { p[0] = 0; }

// one lane per selected row: copies the kept entries in their order (the
// selected columns are numbered increasingly, so the row stays sorted) and
// forms b[g] - sum over dropped columns of a_gj x_j
__global__ void __launch_bounds__(kBlock)
select_fill_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                   const double* __restrict__ data, const int32_t* __restrict__ rows,
                   int64_t nsel, const int32_t* __restrict__ colmap,
                   const int32_t* __restrict__ out_indptr, int32_t* __restrict__ out_indices,
                   double* __restrict__ out_data, const double* __restrict__ b,
                   const double* __restrict__ x, double* __restrict__ b_out) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nsel) return;
  const int32_t g = rows[r];
  int32_t o = out_indptr[r];
  double s = 0.0;
  for (int32_t k = indptr[g]; k < indptr[g + 1]; ++k) {
    const int32_t j = indices[k];
    const int32_t c = colmap[j];
    if (c >= 0) {
      out_indices[o] = c;
      if (out_data) out_data[o] = data[k];
      ++o;
    } else if (x != nullptr) {
      s += data[k] * x[j];
    }
  }
  if (b_out) b_out[r] = b[g] - s;
}

// ------------------------------------------------------------------- spmv --
__global__ void __launch_bounds__(kBlock)
spmv_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
            const double* __restrict__ data, int64_t n, const double* __restrict__ x,
            double* __restrict__ y) {
  const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + (threadIdx.x / kRowLanes);
  const int lane = threadIdx.x % kRowLanes;
  double s = 0.0;
  if (r < n)
    for (int32_t k = indptr[r] + lane; k < indptr[r + 1]; k += kRowLanes)
      s += data[k] * x[indices[k]];
  s = row_sum(s);
  if (r < n && lane == 0) y[r] = s;
}

// ------------------------------------------------------------------- pcg --
// scalars: [0] rz  [1] alpha  [2] beta  [3] rr  [4] bb
__global__ void __launch_bounds__(kBlock)
pcg_init_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                const double* __restrict__ data, int64_t n, const double* __restrict__ b,
                const double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                double* __restrict__ dinv, double* __restrict__ p, double* __restrict__ part) {
  // r = b - A x, dinv = 1 / diag(A), z = dinv r, p = 0; partials of r.z, r.r, b.b
  __shared__ double red[kBlock / 64];
  const int lane = threadIdx.x % kRowLanes;
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t base = (int64_t)blockIdx.x * kRowsPerBlock; base < n;
       base += (int64_t)gridDim.x * kRowsPerBlock) {
    const int64_t r0 = base + threadIdx.x / kRowLanes;
    double s = 0.0, dg = 0.0;
    if (r0 < n)
      for (int32_t k = indptr[r0] + lane; k < indptr[r0 + 1]; k += kRowLanes) {
        s += data[k] * x[indices[k]];
        if (indices[k] == r0) dg = data[k];
      }
    s = row_sum(s);
    dg = row_sum(dg);  // one lane holds the diagonal, the others 0
    if (r0 < n && lane == 0) {
      const double ri = b[r0] - s;
      const double di = 1.0 / dg;
      r[r0] = ri;
      dinv[r0] = di;
      z[r0] = di * ri;
      p[r0] = 0.0;
      rz += ri * di * ri;
      rr += ri * ri;
      bb += b[r0] * b[r0];
    }
  }
  rz = block_sum(rz, red);
  rr = block_sum(rr, red);
  bb = block_sum(bb, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = rz;
    part[3 * blockIdx.x + 1] = rr;
    part[3 * blockIdx.x + 2] = bb;
  }
}

// final sums of per-block partials (NC per block), fixed order; MODE 0:
// init (rz, rr, bb, beta = 0), MODE 1: alpha = rz / pq (0 when pq = 0),
// MODE 2: beta = rz_new / rz (0 when rz = 0), rr
template <int NC, int MODE>
__global__ void __launch_bounds__(kBlock)
pcg_reduce_kernel(const double* __restrict__ part, int64_t nblocks, double* __restrict__ sc) {
  __shared__ double red[kBlock / 64];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  for (int64_t k = threadIdx.x; k < nblocks; k += kBlock) {
    v0 += part[NC * k];
    if constexpr (NC > 1) v1 += part[NC * k + 1];
    if constexpr (NC > 2) v2 += part[NC * k + 2];
  }
  v0 = block_sum(v0, red);
  if constexpr (NC > 1) v1 = block_sum(v1, red);
  if constexpr (NC > 2) v2 = block_sum(v2, red);
  if (threadIdx.x != 0) return;
  if constexpr (MODE == 0) {
    sc[0] = v0;
    sc[3] = v1;
    sc[4] = v2;
    sc[2] = 0.0;
  } else if constexpr (MODE == 1) {
    // a converged residual (r = 0 exactly, so z = p = 0 and p.q = 0) makes
    // the step a no-op instead of 0/0
    sc[1] = v0 != 0.0 ? sc[0] / v0 : 0.0;
  } else {
    sc[2] = sc[0] != 0.0 ? v0 / sc[0] : 0.0;
    sc[0] = v0;
    sc[3] = v1;
  }
}

// p_new = z + beta p_old, q = A p_new (neighbour values formed on the fly
// from z and p_old), partials of p_new . q
__global__ void __launch_bounds__(kBlock)
pcg_spmv_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                const double* __restrict__ data, int64_t n, const double* __restrict__ z,
                const double* __restrict__ p_old, double* __restrict__ p_new,
                double* __restrict__ q, const double* __restrict__ sc,
                double* __restrict__ part) {
  __shared__ double red[kBlock / 64];
  const double beta = sc[2];
  const int lane = threadIdx.x % kRowLanes;
  double pq = 0.0;
  for (int64_t base = (int64_t)blockIdx.x * kRowsPerBlock; base < n;
       base += (int64_t)gridDim.x * kRowsPerBlock) {
    const int64_t r0 = base + threadIdx.x / kRowLanes;
    double s = 0.0;
    if (r0 < n)
      for (int32_t k = indptr[r0] + lane; k < indptr[r0 + 1]; k += kRowLanes) {
        const int32_t j = indices[k];
        s += data[k] * (z[j] + beta * p_old[j]);
      }
    s = row_sum(s);
    if (r0 < n && lane == 0) {
      const double pi = z[r0] + beta * p_old[r0];
      p_new[r0] = pi;
      q[r0] = s;
      pq += pi * s;
    }
  }
  pq = block_sum(pq, red);
  if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

// x += alpha p, r -= alpha q, z = dinv r; partials of r.z and r.r
__global__ void __launch_bounds__(kBlock)
pcg_update_kernel(int64_t n, const double* __restrict__ p, const double* __restrict__ q,
                  const double* __restrict__ dinv, double* __restrict__ x,
                  double* __restrict__ r, double* __restrict__ z,
                  const double* __restrict__ sc, double* __restrict__ part) {
  __shared__ double red[kBlock / 64];
  const double alpha = sc[1];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    const double zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = block_sum(rz, red);
  rr = block_sum(rr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rz;
    part[2 * blockIdx.x + 1] = rr;
  }
}

struct PcgWork {
  double *r, *z, *dinv, *p0, *p1, *q, *part, *sc;
};

size_t pcg_bytes(int64_t n) {
  const int64_t nb = ceil_div(n, kRowsPerBlock) + 1;
  return sizeof(double) * (size_t)(7 * n + 3 * nb + 8) + 256;
}

// the five PCG scalars to the host once the stream has drained (a blocking
// copy after the synchronization: the values are final)
hipError_t read_scalars(double* hs, const double* sc, hipStream_t st) {
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  return hipMemcpy(hs, sc, 5 * sizeof(double), hipMemcpyDeviceToHost);
}

PcgWork carve(void* ws, int64_t n) {
  const int64_t nb = ceil_div(n, kRowsPerBlock) + 1;
  double* b = reinterpret_cast<double*>(ws);
  PcgWork w;
  w.r = b;
  w.z = b + n;
  w.dinv = b + 2 * n;
  w.p0 = b + 3 * n;
  w.p1 = b + 4 * n;
  w.q = b + 5 * n;
  w.part = b + 6 * n;
  w.sc = b + 6 * n + 3 * nb;
  return w;
}

}  // namespace
}  // namespace femhip

using namespace femhip;

extern "C" int femhip_csr_select_workspace(int64_t nsel, size_t* bytes) {
  FEMHIP_REQUIRE(bytes != nullptr && nsel >= 0, FEMHIP_EINVAL, "select workspace: bad arguments");
  size_t tb = 0;
  FEMHIP_TRY_HIP(rocprim::inclusive_scan<rocprim::default_config>(
      nullptr, tb, (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)(nsel > 0 ? nsel : 1),
      rocprim::plus<int32_t>()));
  *bytes = tb + 256;
  return FEMHIP_OK;
}

extern "C" int femhip_csr_select_pattern(const int32_t* indptr, const int32_t* indices,
                                         const int32_t* rows, int64_t nsel,
                                         const int32_t* colmap, int32_t* out_indptr,
                                         int64_t* nnz_out, void* workspace, size_t ws_bytes,
                                         femhip_stream_t s) {
  FEMHIP_REQUIRE(indptr && indices && colmap && out_indptr && nnz_out, FEMHIP_EINVAL,
                 "select: NULL argument");
  FEMHIP_REQUIRE(nsel >= 0 && nsel < (1ll << 31), FEMHIP_EINVAL, "select: bad row count");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  zero_first_kernel<<<1, 1, 0, st>>>(out_indptr);
  FEMHIP_CHECK_LAUNCH();
  if (nsel == 0) {
    *nnz_out = 0;
    return FEMHIP_OK;
  }
  FEMHIP_REQUIRE(rows != nullptr, FEMHIP_EINVAL, "select: NULL rows");
  size_t need = 0;
  if (int rc = femhip_csr_select_workspace(nsel, &need); rc != FEMHIP_OK) return rc;
  FEMHIP_REQUIRE(workspace != nullptr && ws_bytes >= need, FEMHIP_ESIZE,
                 "select: workspace too small");
  // counts into out_indptr[1..], then an in-place inclusive scan
  select_count_kernel<<<(unsigned)ceil_div(nsel, kRowsPerBlock), kBlock, 0, st>>>(
      indptr, indices, rows, nsel, colmap, out_indptr + 1);
  FEMHIP_CHECK_LAUNCH();
  size_t tb = need - 256;
  FEMHIP_TRY_HIP(rocprim::inclusive_scan<rocprim::default_config>(
      workspace, tb, out_indptr + 1, out_indptr + 1, (size_t)nsel, rocprim::plus<int32_t>(), st));
  int32_t nnz = 0;
  FEMHIP_TRY_HIP(hipStreamSynchronize(st));
  FEMHIP_TRY_HIP(hipMemcpy(&nnz, out_indptr + nsel, sizeof(int32_t), hipMemcpyDeviceToHost));
  *nnz_out = nnz;
  return FEMHIP_OK;
}

extern "C" int femhip_csr_select_fill(const int32_t* indptr, const int32_t* indices,
                                      const double* data, const int32_t* rows, int64_t nsel,
                                      const int32_t* colmap, const int32_t* out_indptr,
                                      int32_t* out_indices, double* out_data, const double* b,
                                      const double* x, double* b_out, femhip_stream_t s) {
  FEMHIP_REQUIRE(indptr && indices && colmap && out_indptr && out_indices, FEMHIP_EINVAL,
                 "select fill: NULL argument");
  FEMHIP_REQUIRE(out_data == nullptr || data != nullptr, FEMHIP_EINVAL,
                 "select fill: values need data");
  FEMHIP_REQUIRE(b_out == nullptr || b != nullptr, FEMHIP_EINVAL, "select fill: b_out needs b");
  if (nsel == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(rows != nullptr, FEMHIP_EINVAL, "select fill: NULL rows");
  select_fill_kernel<<<(unsigned)ceil_div(nsel, kBlock), kBlock, 0,
                       reinterpret_cast<hipStream_t>(s)>>>(indptr, indices, data, rows, nsel,
                                                           colmap, out_indptr, out_indices,
                                                           out_data, b, x, b_out);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" int femhip_spmv(const int32_t* indptr, const int32_t* indices, const double* data,
                           int64_t n, const double* x, double* y, femhip_stream_t s) {
  FEMHIP_REQUIRE(n >= 0, FEMHIP_EINVAL, "spmv: negative size");
  if (n == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(indptr && indices && data && x && y, FEMHIP_EINVAL, "spmv: NULL argument");
  spmv_kernel<<<(unsigned)ceil_div(n, kRowsPerBlock), kBlock, 0,
                reinterpret_cast<hipStream_t>(s)>>>(indptr, indices, data, n, x, y);
  FEMHIP_CHECK_LAUNCH();
  return FEMHIP_OK;
}

extern "C" size_t femhip_pcg_workspace_size(int64_t n) { return n > 0 ? pcg_bytes(n) : 0; }

extern "C" int femhip_pcg(const int32_t* indptr, const int32_t* indices, const double* data,
                          int64_t n, const double* b, double* x, double rtol, double atol,
                          int64_t maxiter, int32_t check_every, void* workspace,
                          size_t ws_bytes, int64_t* iters, double* resid, femhip_stream_t s) {
  FEMHIP_REQUIRE(n >= 0 && maxiter >= 0 && check_every > 0 && rtol >= 0.0 && atol >= 0.0,
                 FEMHIP_EINVAL, "pcg: bad arguments");
  FEMHIP_REQUIRE(iters != nullptr && resid != nullptr, FEMHIP_EINVAL, "pcg: NULL outputs");
  *iters = 0;
  *resid = 0.0;
  if (n == 0) return FEMHIP_OK;
  FEMHIP_REQUIRE(indptr && indices && data && b && x, FEMHIP_EINVAL, "pcg: NULL argument");
  FEMHIP_REQUIRE(workspace != nullptr && ws_bytes >= pcg_bytes(n), FEMHIP_ESIZE,
                 "pcg: workspace too small");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  PcgWork w = carve(workspace, n);
  // fixed grids with grid-stride loops: at most kPcgGrid partial sums per
  // reduction (the grid depends on n only, so the summation order does too)
  const unsigned gr = (unsigned)std::min<int64_t>(ceil_div(n, kRowsPerBlock), kPcgGrid);
  const unsigned gv = (unsigned)std::min<int64_t>(ceil_div(n, kBlock), kPcgGrid);
  pcg_init_kernel<<<gr, kBlock, 0, st>>>(indptr, indices, data, n, b, x, w.r, w.z, w.dinv, w.p0,
                                         w.part);
  FEMHIP_CHECK_LAUNCH();
  pcg_reduce_kernel<3, 0><<<1, kBlock, 0, st>>>(w.part, gr, w.sc);
  FEMHIP_CHECK_LAUNCH();
  double hs[5];
  FEMHIP_TRY_HIP(read_scalars(hs, w.sc, st));
  const double bnorm = sqrt(hs[4]);
  const double target = fmax(rtol * bnorm, atol);
  double rnorm = sqrt(hs[3]);
  int64_t it = 0;
  double* pold = w.p0;
  double* pnew = w.p1;
  while (rnorm > target && it < maxiter) {
    const int64_t todo = std::min<int64_t>(check_every, maxiter - it);
    for (int64_t k = 0; k < todo; ++k) {
      pcg_spmv_kernel<<<gr, kBlock, 0, st>>>(indptr, indices, data, n, w.z, pold, pnew, w.q, w.sc,
                                             w.part);
      pcg_reduce_kernel<1, 1><<<1, kBlock, 0, st>>>(w.part, gr, w.sc);
      pcg_update_kernel<<<gv, kBlock, 0, st>>>(n, pnew, w.q, w.dinv, x, w.r, w.z, w.sc, w.part);
      pcg_reduce_kernel<2, 2><<<1, kBlock, 0, st>>>(w.part, gv, w.sc);
      std::swap(pold, pnew);
    }
    FEMHIP_CHECK_LAUNCH();
    it += todo;
    FEMHIP_TRY_HIP(read_scalars(hs, w.sc, st));
    rnorm = sqrt(hs[3]);
  }
  *iters = it;
  *resid = rnorm;
  FEMHIP_REQUIRE(std::isfinite(rnorm), FEMHIP_EINVAL,
                 "pcg: the residual is not finite (matrix not symmetric positive definite?)");
  return FEMHIP_OK;
}
