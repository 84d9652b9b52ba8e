// csr_kernels.h -- the device Jacobi-PCG behind slot A (the linear-solver
// policy, Policy_Solve_Linear_Umfpack.h:7-71): kernels, the device state of
// one solve and the host loop that queues them.  A is negative
// (semi)definite, so CG runs on (-A) x = -b; every CG scalar stays on the
// device.
//   n <= CSR_SINGLE_MAX  one workgroup runs the whole solve (setup, every
//                        iteration and the convergence test) in one launch;
//                        the vectors stay in that CU's caches, and no launch
//                        or host round trip separates the iterations.
//   larger n             three grid-stride launches per iteration, scalars
//                        reduced on the device in fixed block order; a batch
//                        of CSR_BATCH iterations is queued before the host
//                        reads the convergence flag once (launches after
//                        convergence return at once).
// The work arrays live wherever the caller puts them (carve()).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>

#include "../../include/mac_amd.h"
#include "brick.h"
#include "mac_internal.h"

namespace {

constexpr int CSR_SINGLE_MAX = 32768;  // rows solved by the one-workgroup kernel
constexpr int CSR_SINGLE_THREADS = 1024;
constexpr int CSR_THREADS = 256;
constexpr int CSR_GRID_MAX = 2048;
constexpr int CSR_BATCH = 16;

// Device state of one solve.  status: 0 running, 1 converged, -1 breakdown
// (p.(-A)p <= 0: the matrix is not negative definite), -2 max_iter reached,
// -3 malformed CSR (device-side check), -4 singular small system.
struct CsrState {
  double rz, rr, bb, pq, tol2;
  int iters, status, done, maxit;
  unsigned counter;
  unsigned pad[3];
};
static_assert(sizeof(CsrState) <= 256, "state block");

// Device arrays of one solve.
struct CsrWork {
  CsrState* st;
  double *partials, *r, *z, *p, *q, *dinv;
};

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the work arrays for n rows (state, partials, five vectors).
constexpr size_t work_bytes(size_t n) {
  return 256 + align256(CSR_GRID_MAX * sizeof(double)) + 5 * align256(n * sizeof(double));
}

CsrWork carve(void* base, size_t n) {
  char* b = (char*)base;
  CsrWork w;
  w.st = (CsrState*)b;
  b += 256;
  w.partials = (double*)b;
  b += align256(CSR_GRID_MAX * sizeof(double));
  double** v[5] = {&w.r, &w.z, &w.p, &w.q, &w.dinv};
  for (auto* pv : v) {
    *pv = (double*)b;
    b += align256(n * sizeof(double));
  }
  return w;
}

// ---------------------------------------------------------------- kernels --

// Sum of v over the block, returned in every thread (fixed order).
__device__ __forceinline__ double block_sum_all(double v, double* sh, double* bc) {
  double s = block_sum(v, sh);
  if (threadIdx.x == 0) *bc = s;
  __syncthreads();
  s = *bc;
  __syncthreads();
  return s;
}

// Rows 0..n-1 well formed: off[0] = 0, off[n] = nnz, non-decreasing offsets,
// column indices in [0, n).  Returns this thread's verdict (true = bad).
__device__ __forceinline__ bool bad_rows(int n, int nnz, const int* __restrict__ off, const int* __restrict__ idx, int i0,
                                         int step) {
  bool bad = false;
  if (i0 == 0 && (off[0] != 0 || off[n] != nnz)) bad = true;
  for (int i = i0; i < n && !bad; i += step) {
    const int a = off[i], e = off[i + 1];
    if (e < a || a < 0 || e > nnz) {
      bad = true;
      break;
    }
    for (int k = a; k < e; ++k)
      if (idx[k] < 0 || idx[k] >= n) {
        bad = true;
        break;
      }
  }
  return bad;
}

// Jacobi factor of row i of -A: 1 / (-sum of the row's diagonal entries), or
// 1 for a zero diagonal.
__device__ __forceinline__ double inv_diag(const int* __restrict__ off, const int* __restrict__ idx,
                                           const double* __restrict__ val, int i) {
  double dg = 0.0;
  for (int k = off[i]; k < off[i + 1]; ++k)
    if (idx[k] == i) dg -= val[k];
  return dg != 0.0 ? 1.0 / dg : 1.0;
}

// The whole Jacobi-PCG in one workgroup.
__global__ __launch_bounds__(CSR_SINGLE_THREADS) void k_csr_cg_single(int n, int nnz, const int* __restrict__ off,
                                                                       const int* __restrict__ idx,
                                                                       const double* __restrict__ val,
                                                                       const double* __restrict__ b, double* __restrict__ x,
                                                                       CsrWork w, double rtol, int maxit) {
  __shared__ double sh[16];
  __shared__ double bc;
  const int t = threadIdx.x, S = blockDim.x;
  const bool bad = bad_rows(n, nnz, off, idx, t, S);
  if (block_sum_all(bad ? 1.0 : 0.0, sh, &bc) != 0.0) {
    if (t == 0) {
      w.st->status = -3;
      w.st->iters = 0;
    }
    return;
  }
  double arz = 0.0, arr = 0.0;
  for (int i = t; i < n; i += S) {
    const double di = inv_diag(off, idx, val, i);
    const double ri = -b[i];
    const double zi = di * ri;
    w.dinv[i] = di;
    x[i] = 0.0;
    w.r[i] = ri;
    w.p[i] = zi;
    arz += ri * zi;
    arr += ri * ri;
  }
  double rz = block_sum_all(arz, sh, &bc);
  const double bb = block_sum_all(arr, sh, &bc);
  const double tol2 = rtol * rtol * bb;
  int status = 0, it = 0;
  if (bb == 0.0) status = 1;
  while (status == 0) {
    if (it >= maxit) {
      status = -2;
      break;
    }
    // q = (-A) p, p.q   (the barrier inside the sum orders the p writes of
    // the previous pass before these reads)
    double apq = 0.0;
    for (int i = t; i < n; i += S) {
      double s = 0.0;
      for (int k = off[i]; k < off[i + 1]; ++k) s += val[k] * w.p[idx[k]];
      w.q[i] = -s;
      apq += w.p[i] * (-s);
    }
    const double pq = block_sum_all(apq, sh, &bc);
    if (!(pq > 0.0)) {
      status = -1;
      break;
    }
    const double alpha = rz / pq;
    arz = arr = 0.0;
    for (int i = t; i < n; i += S) {
      x[i] += alpha * w.p[i];
      const double ri = w.r[i] - alpha * w.q[i];
      w.r[i] = ri;
      const double zi = w.dinv[i] * ri;
      w.z[i] = zi;
      arz += ri * zi;
      arr += ri * ri;
    }
    const double rz1 = block_sum_all(arz, sh, &bc);
    const double rr = block_sum_all(arr, sh, &bc);
    ++it;
    if (rr <= tol2) {
      status = 1;
      break;
    }
    const double beta = rz1 / rz;
    rz = rz1;
    for (int i = t; i < n; i += S) w.p[i] = w.z[i] + beta * w.p[i];
    __syncthreads();
  }
  if (t == 0) {
    w.st->status = status;
    w.st->iters = it;
    w.st->done = 1;
  }
}

// Fixed-order grid reduction of two values (partials[2 * block + 0/1]);
// true in the last block to arrive, totals in t0, t1 (thread 0).
__device__ __forceinline__ bool grid_reduce2(double a, double b, double* __restrict__ partials, unsigned* counter, double* t0,
                                             double* t1, double* sh) {
  __shared__ int am_last;
  double sa = block_sum(a, sh);
  double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[2 * blockIdx.x], sa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&partials[2 * blockIdx.x + 1], sb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double x0 = 0.0, x1 = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
    x0 += ld_coherent(&partials[2 * i]);
    x1 += ld_coherent(&partials[2 * i + 1]);
  }
  x0 = block_sum(x0, sh);
  x1 = block_sum(x1, sh);
  if (threadIdx.x == 0) {
    *t0 = x0;
    *t1 = x1;
    *counter = 0u;
  }
  return true;
}

__global__ void k_csr_reset(CsrState* st, int maxit) {
  if (threadIdx.x == 0) {
    st->rz = st->rr = st->bb = st->pq = st->tol2 = 0.0;
    st->iters = 0;
    st->status = 0;
    st->done = 0;
    st->maxit = maxit;
    st->counter = 0u;
  }
}

// Per-row checks of the multi-block path; any bad row sets status -3.
__global__ void k_csr_check(int n, int nnz, const int* __restrict__ off, const int* __restrict__ idx, CsrState* st) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, step = gridDim.x * blockDim.x;
  if (bad_rows(n, nnz, off, idx, i0, step)) {
    __hip_atomic_store(&st->status, -3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// x = 0, r = -b, z = p = D^-1 r; r.z and r.r
__global__ void k_csr_init(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                           const double* __restrict__ b, double* __restrict__ x, CsrWork w, double rtol) {
  __shared__ double sh[16];
  if (w.st->done) return;
  double arz = 0.0, arr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double di = inv_diag(off, idx, val, i);
    const double ri = -b[i];
    const double zi = di * ri;
    w.dinv[i] = di;
    x[i] = 0.0;
    w.r[i] = ri;
    w.p[i] = zi;
    arz += ri * zi;
    arr += ri * ri;
  }
  double rz, rr;
  if (grid_reduce2(arz, arr, w.partials, &w.st->counter, &rz, &rr, sh) && threadIdx.x == 0) {
    w.st->rz = rz;
    w.st->rr = rr;
    w.st->bb = rr;
    w.st->tol2 = rtol * rtol * rr;
    if (rr == 0.0) {
      w.st->status = 1;
      w.st->done = 1;
    }
  }
}

// q = (-A) p, p.q
__global__ void k_csr_spmv(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                           CsrWork w) {
  __shared__ double sh[16];
  if (w.st->done) return;
  double apq = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k) s += val[k] * w.p[idx[k]];
    w.q[i] = -s;
    apq += w.p[i] * (-s);
  }
  double pq, unused;
  if (grid_reduce2(apq, 0.0, w.partials, &w.st->counter, &pq, &unused, sh) && threadIdx.x == 0) {
    w.st->pq = pq;
    if (!(pq > 0.0)) {
      w.st->status = -1;
      w.st->done = 1;
    }
  }
}

// x += a p, r -= a q, z = D^-1 r; r.z, r.r; convergence test
__global__ void k_csr_update(int n, double* __restrict__ x, CsrWork w) {
  __shared__ double sh[16];
  if (w.st->done) return;
  const double alpha = w.st->rz / w.st->pq;
  double arz = 0.0, arr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * w.p[i];
    const double ri = w.r[i] - alpha * w.q[i];
    w.r[i] = ri;
    const double zi = w.dinv[i] * ri;
    w.z[i] = zi;
    arz += ri * zi;
    arr += ri * ri;
  }
  double rz1, rr;
  if (grid_reduce2(arz, arr, w.partials, &w.st->counter, &rz1, &rr, sh) && threadIdx.x == 0) {
    CsrState* st = w.st;
    st->iters += 1;
    st->pq = rz1 / st->rz;  // beta for k_csr_pupdate
    st->rz = rz1;
    st->rr = rr;
    if (rr <= st->tol2) {
      st->status = 1;
      st->done = 1;
    } else if (st->iters >= st->maxit) {
      st->status = -2;
      st->done = 1;
    }
  }
}

// p = z + beta p
__global__ void k_csr_pupdate(int n, CsrWork w) {
  if (w.st->done) return;
  const double beta = w.st->pq;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w.p[i] = w.z[i] + beta * w.p[i];
}

// Gaussian elimination with partial pivoting of an n <= 3 CSR system, one
// thread (the device counterpart of small_solve below).
__global__ void k_csr_small(int n, int nnz, const int* __restrict__ off, const int* __restrict__ idx,
                            const double* __restrict__ val, const double* __restrict__ b, double* __restrict__ x,
                            CsrState* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double A[3][4] = {{0}};
  st->iters = 0;
  st->done = 1;
  if (off[0] != 0 || off[n] != nnz) {
    st->status = -3;
    return;
  }
  // compressed columns, UMFPACK's layout: column j holds entries off[j]..off[j+1]
  for (int j = 0; j < n; ++j) {
    if (off[j + 1] < off[j]) {
      st->status = -3;
      return;
    }
    for (int k = off[j]; k < off[j + 1]; ++k) {
      if (idx[k] < 0 || idx[k] >= n) {
        st->status = -3;
        return;
      }
      A[idx[k]][j] += val[k];
    }
    A[j][3] = b[j];
  }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int i = c + 1; i < n; ++i)
      if (fabs(A[i][c]) > fabs(A[piv][c])) piv = i;
    if (A[piv][c] == 0.0) {
      st->status = -4;
      return;
    }
    if (piv != c)
      for (int j = 0; j < 4; ++j) {
        double tmp = A[c][j];
        A[c][j] = A[piv][j];
        A[piv][j] = tmp;
      }
    for (int i = c + 1; i < n; ++i) {
      double m = A[i][c] / A[c][c];
      for (int j = c; j < 4; ++j) A[i][j] -= m * A[c][j];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = A[i][3];
    for (int j = i + 1; j < n; ++j) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
  st->status = 1;
}

// compressed columns (UMFPACK's layout: column j holds entries off[j]..off[j+1],
// row indices idx), which matters for the non-symmetric 3x3 corner systems
int small_solve(int n, const int* off, const int* idx, const double* val, const double* b, double* x) {
  double A[3][4] = {{0}};
  for (int j = 0; j < n; ++j) {
    for (int k = off[j]; k < off[j + 1]; ++k) A[idx[k]][j] += val[k];
    A[j][3] = b[j];
  }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(A[i][c]) > std::fabs(A[piv][c])) piv = i;
    if (A[piv][c] == 0.0) return MAC_E_NOCONV;
    if (piv != c)
      for (int j = 0; j < 4; ++j) std::swap(A[c][j], A[piv][j]);
    for (int i = c + 1; i < n; ++i) {
      double m = A[i][c] / A[c][c];
      for (int j = c; j < 4; ++j) A[i][j] -= m * A[c][j];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = A[i][3];
    for (int j = i + 1; j < n; ++j) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
  return 0;
}

int maxit_for(int n) {
  const long long m = 20ll * n + 1000;
  return m > 2000000000ll ? 2000000000 : (int)m;
}

#define CSR_HIPRET(e)                                                                  \
  do {                                                                                 \
    if ((e) != hipSuccess) return mac_internal_set_error(MAC_E_HIP, "HIP failure: " #e); \
  } while (0)

// Queue the solve of (-A) x = -b on stream st and wait for it.  All pointers
// are device pointers; w holds at least n rows.  hst: host copy of the final
// state (pinned or not).
int run_cg(hipStream_t st, int n, int nnz, const int* off, const int* idx, const double* val, const double* b, double* x,
           const CsrWork& w, double rtol, CsrState* hst, bool rows_checked) {
  const int maxit = maxit_for(n);
  k_csr_reset<<<1, 64, 0, st>>>(w.st, maxit);
  CSR_HIPRET(hipGetLastError());
  if (n <= CSR_SINGLE_MAX) {
    k_csr_cg_single<<<1, CSR_SINGLE_THREADS, 0, st>>>(n, nnz, off, idx, val, b, x, w, rtol, maxit);
    CSR_HIPRET(hipGetLastError());
    CSR_HIPRET(hipMemcpyAsync(hst, w.st, sizeof(CsrState), hipMemcpyDeviceToHost, st));
    CSR_HIPRET(hipStreamSynchronize(st));
    return 0;
  }
  int blocks = (n + CSR_THREADS - 1) / CSR_THREADS;
  if (blocks > CSR_GRID_MAX) blocks = CSR_GRID_MAX;
  if (!rows_checked) {
    k_csr_check<<<blocks, CSR_THREADS, 0, st>>>(n, nnz, off, idx, w.st);
    CSR_HIPRET(hipGetLastError());
  }
  k_csr_init<<<blocks, CSR_THREADS, 0, st>>>(n, off, idx, val, b, x, w, rtol);
  CSR_HIPRET(hipGetLastError());
  for (;;) {
    for (int k = 0; k < CSR_BATCH; ++k) {
      k_csr_spmv<<<blocks, CSR_THREADS, 0, st>>>(n, off, idx, val, w);
      k_csr_update<<<blocks, CSR_THREADS, 0, st>>>(n, x, w);
      k_csr_pupdate<<<blocks, CSR_THREADS, 0, st>>>(n, w);
    }
    CSR_HIPRET(hipGetLastError());
    CSR_HIPRET(hipMemcpyAsync(hst, w.st, sizeof(CsrState), hipMemcpyDeviceToHost, st));
    CSR_HIPRET(hipStreamSynchronize(st));
    if (hst->done) break;
  }
  return 0;
}

int status_rc(const CsrState& s) {
  switch (s.status) {
    case 1:
      return 0;
    case -1:
      return mac_internal_set_error(MAC_E_NOCONV, "CG breakdown: matrix not negative definite");
    case -2:
      return mac_internal_set_error(MAC_E_NOCONV, "CSR CG did not converge");
    case -3:
      return mac_internal_set_error(MAC_E_ARG, "malformed CSR: offsets or column indices out of range");
    case -4:
      return mac_internal_set_error(MAC_E_NOCONV, "singular small system");
    default:
      return mac_internal_set_error(MAC_E_HIP, "CSR solve ended in an unknown state");
  }
}

}  // namespace
