// csr_solve.hip -- slot A of the reference, the linear-solver policy
// (Policy_Solve_Linear_Umfpack.h:7-71), on the GPU: solve A x = b for the
// symmetric CSR (= CSC) matrix assembled by
// Algorithms_Solve_Pressure::Init_Iteration (:44-157).  UMFPACK factorises;
// this slot runs Jacobi-preconditioned CG in FP64 on the device to a tight
// relative residual, which at rtol <= 1e-12 agrees with the direct solve far
// inside the 1e-8 parity bound.  Systems of size <= 3 (the extrapolation
// boundary condition's corner solves, Algorithm_Extrapolate_Boundary_Condition.h:549-576)
// are solved on the host by Gaussian elimination with partial pivoting.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mac_amd.h"
#include "brick.h"

namespace {

struct CsrScal {
  double v[4];
  unsigned counter;
};

// y = -A x and, when want, partial x.y (the CG curvature with M = -A SPD)
__global__ void k_csr_negspmv_dot(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                                  const double* __restrict__ x, double* __restrict__ y, double* __restrict__ partials,
                                  CsrScal* s, int slot) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double t = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k) t += val[k] * x[idx[k]];
    y[i] = -t;
    acc += x[i] * (-t);
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) s->v[slot] = tot;
}

// z = r / diag(-A); partial r.z
__global__ void k_csr_jacobi_dot(int n, const int* __restrict__ off, const int* __restrict__ idx, const double* __restrict__ val,
                                 const double* __restrict__ r, double* __restrict__ z, double* __restrict__ partials, CsrScal* s,
                                 int slot) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double dg = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k)
      if (idx[k] == i) dg += -val[k];
    double zi = dg != 0.0 ? r[i] / dg : r[i];
    z[i] = zi;
    acc += r[i] * zi;
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) s->v[slot] = tot;
}

// x += a p; r -= a q; partial r.r
__global__ void k_csr_axpy(int n, double a, double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                           const double* __restrict__ q, double* __restrict__ partials, CsrScal* s, int slot) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += a * p[i];
    double rv = r[i] - a * q[i];
    r[i] = rv;
    acc += rv * rv;
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) s->v[slot] = tot;
}

__global__ void k_csr_pupdate(int n, double beta, const double* __restrict__ z, double* __restrict__ p) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

__global__ void k_neg_copy(int n, const double* __restrict__ b, double* __restrict__ r) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) r[i] = -b[i];
}

int small_solve(int n, const int* off, const int* idx, const double* val, const double* b, double* x) {
  double A[3][4] = {{0}};
  for (int i = 0; i < n; ++i) {
    for (int k = off[i]; k < off[i + 1]; ++k) {
      if (idx[k] < 0 || idx[k] >= n) return MAC_E_ARG;
      A[i][idx[k]] += val[k];
    }
    A[i][3] = b[i];
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

}  // namespace

extern "C" int mac_csr_set_error(const char* msg);

extern "C" int mac_solve_csr(mac_ctx* ctx, int n, const int* offset, const int* indice, const double* value, const double* b,
                             double* x, double rtol) {
  if (!ctx) return mac_csr_set_error("null context");
  if (n == 0) return 0;
  if (n < 0 || !offset || !indice || !value || !b || !x) return mac_csr_set_error("bad CSR arguments");
  int nnz = offset[n];
  if (offset[0] != 0 || nnz < 0) return mac_csr_set_error("bad CSR offsets");
  for (int i = 0; i < n; ++i)
    if (offset[i + 1] < offset[i]) return mac_csr_set_error("CSR offsets must be non-decreasing");
  for (int k = 0; k < nnz; ++k)
    if (indice[k] < 0 || indice[k] >= n) return mac_csr_set_error("CSR column index out of range");
  if (n <= 3) {
    int rc = small_solve(n, offset, indice, value, b, x);
    if (rc) mac_csr_set_error("singular small system");
    return rc;
  }
  int *doff = nullptr, *didx = nullptr;
  double *dval = nullptr, *db = nullptr, *dx = nullptr, *dr = nullptr, *dz = nullptr, *dp = nullptr, *dq = nullptr,
         *dpart = nullptr;
  CsrScal *ds = nullptr, hs;
  int rc = 0;
  const int threads = 256;
  int blocks = (n + threads - 1) / threads;
  if (blocks > 4096) blocks = 4096;
#define TRY(e)                                   \
  do {                                           \
    if ((e) != hipSuccess) {                     \
      rc = mac_csr_set_error("HIP failure: " #e); \
      goto done;                                 \
    }                                            \
  } while (0)
  TRY(hipMalloc(&doff, (n + 1) * sizeof(int)));
  TRY(hipMalloc(&didx, (nnz > 0 ? nnz : 1) * sizeof(int)));
  TRY(hipMalloc(&dval, (nnz > 0 ? nnz : 1) * sizeof(double)));
  TRY(hipMalloc(&db, n * sizeof(double)));
  TRY(hipMalloc(&dx, n * sizeof(double)));
  TRY(hipMalloc(&dr, n * sizeof(double)));
  TRY(hipMalloc(&dz, n * sizeof(double)));
  TRY(hipMalloc(&dp, n * sizeof(double)));
  TRY(hipMalloc(&dq, n * sizeof(double)));
  TRY(hipMalloc(&dpart, blocks * sizeof(double)));
  TRY(hipMalloc(&ds, sizeof(CsrScal)));
  TRY(hipMemcpy(doff, offset, (n + 1) * sizeof(int), hipMemcpyHostToDevice));
  if (nnz > 0) {
    TRY(hipMemcpy(didx, indice, nnz * sizeof(int), hipMemcpyHostToDevice));
    TRY(hipMemcpy(dval, value, nnz * sizeof(double), hipMemcpyHostToDevice));
  }
  TRY(hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice));
  TRY(hipMemset(dx, 0, n * sizeof(double)));
  TRY(hipMemset(dp, 0, n * sizeof(double)));
  TRY(hipMemset(ds, 0, sizeof(CsrScal)));
  {
    // (-A) x = -b, SPD; Jacobi-PCG
    k_neg_copy<<<blocks, threads>>>(n, db, dr);
    k_csr_jacobi_dot<<<blocks, threads>>>(n, doff, didx, dval, dr, dz, dpart, ds, 0);
    // ||b||^2 via r.r : reuse axpy with a = 0
    k_csr_axpy<<<blocks, threads>>>(n, 0.0, dx, dr, dp, dp, dpart, ds, 2);
    TRY(hipMemcpy(&hs, ds, sizeof(CsrScal), hipMemcpyDeviceToHost));
    double bb = hs.v[2];
    double rz = hs.v[0];
    if (bb == 0.0) goto copyback;
    double tol2 = rtol * rtol * bb;
    double beta = 0.0;
    int maxit = 20 * n + 1000;
    int it = 0;
    for (; it < maxit; ++it) {
      k_csr_pupdate<<<blocks, threads>>>(n, beta, dz, dp);
      k_csr_negspmv_dot<<<blocks, threads>>>(n, doff, didx, dval, dp, dq, dpart, ds, 1);
      TRY(hipMemcpy(&hs, ds, sizeof(CsrScal), hipMemcpyDeviceToHost));
      double pq = hs.v[1];
      if (!(pq > 0.0)) {
        rc = mac_csr_set_error("CG breakdown: matrix not negative definite");
        goto done;
      }
      double alpha = rz / pq;
      k_csr_axpy<<<blocks, threads>>>(n, alpha, dx, dr, dp, dq, dpart, ds, 2);
      k_csr_jacobi_dot<<<blocks, threads>>>(n, doff, didx, dval, dr, dz, dpart, ds, 0);
      TRY(hipMemcpy(&hs, ds, sizeof(CsrScal), hipMemcpyDeviceToHost));
      if (hs.v[2] <= tol2) break;
      beta = hs.v[0] / rz;
      rz = hs.v[0];
    }
    if (it >= maxit) {
      rc = mac_csr_set_error("CSR CG did not converge");
      rc = MAC_E_NOCONV;
    }
  }
copyback:
  TRY(hipMemcpy(x, dx, n * sizeof(double), hipMemcpyDeviceToHost));
done:
  hipFree(doff);
  hipFree(didx);
  hipFree(dval);
  hipFree(db);
  hipFree(dx);
  hipFree(dr);
  hipFree(dz);
  hipFree(dp);
  hipFree(dq);
  hipFree(dpart);
  hipFree(ds);
#undef TRY
  return rc;
}
