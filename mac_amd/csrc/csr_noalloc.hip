// csr_noalloc.hip -- mac_solve_csr_noalloc: the slot A solve (the same
// Jacobi-PCG kernels as mac_solve_csr, csr_kernels.h) for callers that must
// not allocate: caller-owned device arrays in and out, and the solve's work
// arrays (state, reduction partials, ten vectors) in `scratch`, a fixed
// file-scope device array of work_bytes(MAC_CSR_NOALLOC_MAX_N) bytes (about
// 21.1 MB), held by every GPU the library runs on until the process ends.
//
// Host code never passes `scratch` itself (in host code the name is a
// host-side stand-in, not a device address): on the first call on a device,
// with that device current, hipGetSymbolAddress gives that device's copy,
// kept per device and used for every kernel argument.  Calls that overlap on
// one GPU share that array and race; k_csr_reset clears the solve state at
// the start of every call, and every vector is written before it is read.
#include <mutex>

#include "csr_kernels.h"

namespace {

constexpr size_t NOALLOC_BYTES = work_bytes(MAC_CSR_NOALLOC_MAX_N);
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEV = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEV] = {nullptr};

// This device's copy of scratch, looked up once per device.
int scratch_of(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= MAX_DEV) return mac_internal_set_error(MAC_E_ARG, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return mac_internal_set_error(MAC_E_HIP, "hipGetDevice failed");
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return mac_internal_set_error(MAC_E_HIP, "hipSetDevice failed");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) hipSetDevice(cur);
    if (e != hipSuccess || !p) return mac_internal_set_error(MAC_E_HIP, "hipGetSymbolAddress(scratch) failed");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int mac_solve_csr_noalloc(mac_ctx* ctx, int n, int nnz, const int* d_offset, const int* d_indice,
                                     const double* d_value, const double* d_b, double* d_x, double rtol, int* iters) {
  if (iters) *iters = 0;
  if (!ctx) return mac_internal_set_error(MAC_E_ARG, "null context");
  if (n == 0) return 0;
  if (n < 0 || nnz < 0 || !d_offset || (nnz > 0 && (!d_indice || !d_value)) || !d_b || !d_x)
    return mac_internal_set_error(MAC_E_ARG, "bad CSR arguments");
  if (n > MAC_CSR_NOALLOC_MAX_N)
    return mac_internal_set_error(MAC_E_ARG, "mac_solve_csr_noalloc: n exceeds MAC_CSR_NOALLOC_MAX_N (262144 rows)");
  void* base = nullptr;
  int rc = scratch_of(mac_internal_device(ctx), &base);
  if (rc) return rc;
  hipStream_t st = mac_internal_stream(ctx);
  CsrWork w = carve(base, n);
  CsrState hst;
  if (n <= 3) {
    k_csr_small<<<1, 64, 0, st>>>(n, nnz, d_offset, d_indice, d_value, d_b, d_x, w.st);
    CSR_HIPRET(hipGetLastError());
    CSR_HIPRET(hipMemcpyAsync(&hst, w.st, sizeof(CsrState), hipMemcpyDeviceToHost, st));
    CSR_HIPRET(hipStreamSynchronize(st));
  } else {
    rc = run_cg(st, n, nnz, d_offset, d_indice, d_value, d_b, d_x, w, rtol, &hst, false);
    if (rc) return rc;
  }
  if (iters) *iters = hst.iters;
  return status_rc(hst);
}
