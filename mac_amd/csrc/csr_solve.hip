// csr_solve.hip -- slot A of the reference, the linear-solver policy
// (Policy_Solve_Linear_Umfpack.h:7-71), on the GPU: solve A x = b for a
// matrix in UMFPACK's compressed columns; the large systems are the symmetric
// pressure matrix assembled by Algorithms_Solve_Pressure::Init_Iteration
// (:44-157), where columns and rows coincide.  UMFPACK factorises;
// this slot runs Jacobi-preconditioned CG in FP64 on the device
// (csr_kernels.h) to a tight relative residual, which at rtol <= 1e-12
// agrees with the direct solve far inside the 1e-8 parity bound.
//
// The CSR copy and the CG vectors live in a device workspace owned by the
// context (grown, never shrunk, freed by mac_destroy), so only a call larger
// than every earlier one allocates; all work runs on the context's stream.
// Systems of size <= 3 (the extrapolation boundary condition's corner
// solves, Algorithm_Extrapolate_Boundary_Condition.h:549-576, which are not
// symmetric) are solved on the host by Gaussian elimination with partial
// pivoting, reading the arrays as compressed columns.
#include "csr_kernels.h"

extern "C" int mac_solve_csr(mac_ctx* ctx, int n, const int* offset, const int* indice, const double* value,
                             const double* b, double* x, double rtol) {
  if (!ctx) return mac_internal_set_error(MAC_E_ARG, "null context");
  if (n == 0) return 0;
  if (n < 0 || !offset || !indice || !value || !b || !x) return mac_internal_set_error(MAC_E_ARG, "bad CSR arguments");
  const int nnz = offset[n];
  if (offset[0] != 0 || nnz < 0) return mac_internal_set_error(MAC_E_ARG, "bad CSR offsets");
  for (int i = 0; i < n; ++i)
    if (offset[i + 1] < offset[i]) return mac_internal_set_error(MAC_E_ARG, "CSR offsets must be non-decreasing");
  for (int k = 0; k < nnz; ++k)
    if (indice[k] < 0 || indice[k] >= n) return mac_internal_set_error(MAC_E_ARG, "CSR column index out of range");
  if (n <= 3) {
    int rc = small_solve(n, offset, indice, value, b, x);
    if (rc) mac_internal_set_error(rc, "singular small system");
    return rc;
  }
  // workspace: the work arrays, then x, b, val, off, idx
  const size_t wb = work_bytes(n);
  const size_t bx = align256((size_t)n * sizeof(double));
  const size_t bval = align256((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  const size_t boff = align256((size_t)(n + 1) * sizeof(int));
  const size_t bidx = align256((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
  void* ws = nullptr;
  void* pin = nullptr;
  int rc = mac_internal_csr_workspace(ctx, wb + 2 * bx + bval + boff + bidx, &ws, &pin);
  if (rc) return rc;
  CsrState* hst = (CsrState*)pin;
  char* base = (char*)ws;
  CsrWork w = carve(base, n);
  double* dx = (double*)(base + wb);
  double* db = (double*)(base + wb + bx);
  double* dval = (double*)(base + wb + 2 * bx);
  int* doff = (int*)(base + wb + 2 * bx + bval);
  int* didx = (int*)(base + wb + 2 * bx + bval + boff);
  hipStream_t st = mac_internal_stream(ctx);
  CSR_HIPRET(hipMemcpyAsync(doff, offset, (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
  if (nnz > 0) {
    CSR_HIPRET(hipMemcpyAsync(didx, indice, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, st));
    CSR_HIPRET(hipMemcpyAsync(dval, value, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, st));
  }
  CSR_HIPRET(hipMemcpyAsync(db, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
  rc = run_cg(st, n, nnz, doff, didx, dval, db, dx, w, rtol, hst, true);
  if (rc) return rc;
  CSR_HIPRET(hipMemcpyAsync(x, dx, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  CSR_HIPRET(hipStreamSynchronize(st));
  return status_rc(*hst);
}
