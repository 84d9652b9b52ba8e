// csr_kernels.h -- the device Jacobi-PCG behind slot A (the linear-solver
// policy, Policy_Solve_Linear_Umfpack.h:7-71): kernels, the device state of
// one solve and the host loop that queues them.  A is negative
// (semi)definite, so CG runs on (-A) x = -b; every CG scalar stays on the
// device.
//   n <= CSR_SINGLE_MAX  one workgroup runs the whole solve (setup, every
//                        iteration and the convergence test) in one launch;
//                        the vectors stay in that CU's caches, and no launch
//                        or host round trip separates the iterations.
//   larger n             two grid-stride launches per iteration (the new
//                        search direction is formed inside the product),
//                        scalars reduced on the device in fixed block order;
//                        a batch of CSR_BATCH iterations is queued before the
//                        host reads the convergence flag once (launches after
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
constexpr int CSR_BATCH = 16;       // iterations queued between host reads, at least
constexpr int CSR_BATCH_MAX = 512;  // and at most

// Device state of one solve.  status: 0 running, 1 converged, -1 breakdown
// (p.(-A)p <= 0: the matrix is not negative definite), -2 max_iter reached,
// -3 malformed CSR (device-side check), -4 singular small system.
struct CsrState {
  double rz, rr, bb, pq, tol2, beta;
  double gam[2], alp[2];  // gamma and alpha of iterations of even and odd parity
  int iters, status, done, maxit;
  unsigned counter;
  unsigned pad[3];
};
static_assert(sizeof(CsrState) <= 256, "state block");

// Device arrays of one solve.  The one-workgroup kernel uses r, z, p, q and
// dinv; the one-launch-per-iteration path uses r, s and w in two copies each
// (rb, sb, wb: read one, write the other), p and dinv.
struct CsrWork {
  CsrState* st;
  double *partials, *r, *z, *p, *q, *dinv, *r1, *s0, *s1, *w0, *w1;
};

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the work arrays for n rows (state, two sets of three partials per
// block, ten vectors).
constexpr size_t work_bytes(size_t n) {
  return 256 + align256(6 * CSR_GRID_MAX * sizeof(double)) + 10 * align256(n * sizeof(double));
}

CsrWork carve(void* base, size_t n) {
  char* b = (char*)base;
  CsrWork w;
  w.st = (CsrState*)b;
  b += 256;
  w.partials = (double*)b;
  b += align256(6 * CSR_GRID_MAX * sizeof(double));
  double** v[10] = {&w.r, &w.z, &w.p, &w.q, &w.dinv, &w.r1, &w.s0, &w.s1, &w.w0, &w.w1};
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

__global__ void k_csr_reset(CsrState* st, int maxit) {
  if (threadIdx.x == 0) {
    st->rz = st->rr = st->bb = st->pq = st->tol2 = st->beta = 0.0;
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

// The multi-workgroup path runs one launch per iteration: the
// single-reduction form of preconditioned CG (Chronopoulos and Gear, 1989),
// with u = D^-1 r and w = (-A) u carried as vectors, so that both inner
// products of an iteration, gamma = r.u and delta = w.u, come from one
// reduction.  Iteration k, with gamma_k, delta_k from the last launch:
//   beta = gamma_k / gamma_{k-1}, alpha = gamma_k / (delta_k - beta gamma_k / alpha_{k-1})
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = D^-1 r, w = (-A) u
// and the new r.u, w.u, r.r.  The product needs u at the neighbour rows; it
// is formed at every gathered column from the neighbour's old r, w and s
// (read from one copy while the new ones go to the other), so no launch
// separates the vector updates from the product.  No grid-wide atomics:
// each launch writes three partials per block with plain stores, and the
// next launch -- after the launch boundary has made them visible -- sums them
// in every block in the same fixed order, so every block takes the same
// decisions.  Two sets of partials alternate by iteration parity (a block
// that is done writes its set while others still read the other), and block
// 0 keeps gamma and alpha in the state in two parity slots likewise.

// Sums of a, b and c over the block, returned in every thread: the fixed
// order of block_sum, with two barriers for the three.  sh holds 3 doubles
// per wave.
__device__ __forceinline__ void block_sum3_all(double& a, double& b, double& c, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
    c += __shfl_xor(c, o);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();  // earlier readers of sh are done
  if (lane == 0) {
    sh[3 * wv] = a;
    sh[3 * wv + 1] = b;
    sh[3 * wv + 2] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
  for (int i = 0; i < nw; ++i) {
    a += sh[3 * i];
    b += sh[3 * i + 1];
    c += sh[3 * i + 2];
  }
}

// The three sums over the G blocks of a partial set (3 doubles per block),
// in every block.
__device__ __forceinline__ void sum_partials(const double* __restrict__ set, double& a, double& b, double& c, double* sh) {
  a = b = c = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
    a += set[3 * i];
    b += set[3 * i + 1];
    c += set[3 * i + 2];
  }
  block_sum3_all(a, b, c, sh);
}

// D^-1 of every row
__global__ void k_csr_diag(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                           CsrWork w) {
  if (w.st->done) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w.dinv[i] = inv_diag(off, idx, val, i);
}

// x = 0, r = -b, u = D^-1 r, w = (-A) u, p = s = 0; partials r.u, w.u, r.r
// into set 0.
__global__ __launch_bounds__(CSR_THREADS) void k_csr_cg_init(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                              const double* __restrict__ b, double* __restrict__ x, CsrWork w) {
  __shared__ double sh[3 * (CSR_THREADS / 64)];
  if (w.st->done) return;
  double g = 0.0, d = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k) {
      const int j = idx[k];
      acc += val[k] * (w.dinv[j] * -b[j]);
    }
    const double ri = -b[i], ui = w.dinv[i] * ri, wi = -acc;
    x[i] = 0.0;
    w.r[i] = ri;
    w.w0[i] = wi;
    w.s0[i] = 0.0;
    w.p[i] = 0.0;
    g += ri * ui;
    d += wi * ui;
    rr += ri * ri;
  }
  block_sum3_all(g, d, rr, sh);
  if (threadIdx.x == 0) {
    w.partials[3 * blockIdx.x] = g;
    w.partials[3 * blockIdx.x + 1] = d;
    w.partials[3 * blockIdx.x + 2] = rr;
  }
}

// Iteration k (partials of set k & 1 in, set (k + 1) & 1 out).  test_only:
// the convergence test alone (it writes nothing unless the solve has ended).
__global__ __launch_bounds__(CSR_THREADS) void k_csr_cg_iter(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                              double* __restrict__ x, CsrWork w, int k, double rtol, int test_only) {
  __shared__ double sh[3 * (CSR_THREADS / 64)];
  CsrState* st = w.st;
  if (st->done) return;
  const int G = gridDim.x;
  const double* pin = w.partials + (size_t)(k & 1) * 3 * G;
  double* pout = w.partials + (size_t)((k + 1) & 1) * 3 * G;
  double gam, del, rr;
  sum_partials(pin, gam, del, rr, sh);
  const double tol2 = k == 0 ? rtol * rtol * rr : st->tol2;
  int status = 0;
  double beta = 0.0, alpha = 0.0;
  if (rr <= tol2) {
    status = 1;
  } else if (k >= st->maxit) {
    status = -2;
  } else {
    beta = k == 0 ? 0.0 : gam / st->gam[(k - 1) & 1];
    const double den = k == 0 ? del : del - beta * gam / st->alp[(k - 1) & 1];
    if (!(gam > 0.0) || !(den > 0.0)) status = -1;
    else alpha = gam / den;
  }
  if (status != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = status;
      st->iters = k;
      st->rr = rr;
      st->done = 1;
    }
    return;
  }
  if (test_only) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rr;
      st->tol2 = tol2;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->gam[k & 1] = gam;
    st->alp[k & 1] = alpha;
    if (k == 0) {
      st->tol2 = tol2;
      st->bb = rr;
    }
  }
  const double* r_in = (k & 1) ? w.r1 : w.r;
  double* r_out = (k & 1) ? w.r : w.r1;
  const double* s_in = (k & 1) ? w.s1 : w.s0;
  double* s_out = (k & 1) ? w.s0 : w.s1;
  const double* w_in = (k & 1) ? w.w1 : w.w0;
  double* w_out = (k & 1) ? w.w0 : w.w1;
  double g = 0.0, d = 0.0, r2 = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double di = w.dinv[i];
    const double ui = di * r_in[i];
    const double pi = ui + beta * w.p[i];
    const double si = w_in[i] + beta * s_in[i];
    w.p[i] = pi;
    x[i] += alpha * pi;
    const double rn = r_in[i] - alpha * si;
    double acc = 0.0;
    for (int kk = off[i]; kk < off[i + 1]; ++kk) {
      const int j = idx[kk];
      const double sj = w_in[j] + beta * s_in[j];
      acc += val[kk] * (w.dinv[j] * (r_in[j] - alpha * sj));
    }
    const double un = di * rn, wn = -acc;
    s_out[i] = si;
    r_out[i] = rn;
    w_out[i] = wn;
    g += rn * un;
    d += wn * un;
    r2 += rn * rn;
  }
  block_sum3_all(g, d, r2, sh);
  if (threadIdx.x == 0) {
    pout[3 * blockIdx.x] = g;
    pout[3 * blockIdx.x + 1] = d;
    pout[3 * blockIdx.x + 2] = r2;
  }
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
  k_csr_diag<<<blocks, CSR_THREADS, 0, st>>>(n, off, idx, val, w);
  k_csr_cg_init<<<blocks, CSR_THREADS, 0, st>>>(n, off, idx, val, b, x, w);
  CSR_HIPRET(hipGetLastError());
  int it = 0, batch = CSR_BATCH, it_prev = 0;
  double rr_prev = 0.0;
  for (;;) {
    // a batch of iterations, then a launch that only tests the last one's
    // residual so that the host sees it
    for (int k = 0; k < batch; ++k, ++it) k_csr_cg_iter<<<blocks, CSR_THREADS, 0, st>>>(n, off, idx, val, x, w, it, rtol, 0);
    k_csr_cg_iter<<<blocks, CSR_THREADS, 0, st>>>(n, off, idx, val, x, w, it, rtol, 1);
    CSR_HIPRET(hipGetLastError());
    CSR_HIPRET(hipMemcpyAsync(hst, w.st, sizeof(CsrState), hipMemcpyDeviceToHost, st));
    CSR_HIPRET(hipStreamSynchronize(st));
    if (hst->done) break;
    // the next batch: about the iterations still needed at the mean decay
    // rate of the last batch (a host read costs about as much as a few
    // iterations; launches queued past convergence return at once)
    batch = CSR_BATCH;
    if (rr_prev > 0.0 && hst->rr > 0.0 && hst->rr < rr_prev && hst->tol2 > 0.0) {
      const double rate = std::log(hst->rr / rr_prev) / (it - it_prev);  // < 0
      const double need = std::log(hst->tol2 / hst->rr) / rate;
      batch = std::max(CSR_BATCH, std::min(CSR_BATCH_MAX, (int)(0.9 * need)));
    }
    rr_prev = hst->rr;
    it_prev = it;
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
