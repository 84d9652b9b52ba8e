// brick.h -- device-side helpers shared by the MAC-grid kernels.
//
// A level of the grid is a list of 8x8x8 bricks.  Brick b holds cells
// b*512 + x + 8y + 64z (x fastest, as the reference's Offset, Offset.h:14-24).
// nbr27[b*27 + (ox+1) + 3(oy+1) + 9(oz+1)] is the brick at offset (ox,oy,oz)
// in brick units (the centre entry is b itself), or -1 when that brick does
// not exist; a missing brick's cells read as zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MAC_BS 8
#define MAC_BC 512

// Stencil code byte (one per cell per level).
//   bits 0..5: the face (-x,+x,-y,+y,-z,+z) enters the operator: it is not a
//              Neumann face, so it adds inv_h^2 to the diagonal and, when the
//              neighbour is in the domain, -inv_h^2 * value(neighbour) (the
//              neighbour value is 0 for every cell outside the domain, which is
//              how air and missing cells act as the p=0 Dirichlet ghost).
//   bit 6:     the cell is an unknown (in the domain).
//   bit 7:     (coarse levels) the cell's Galerkin coefficients are the
//              rediscretised ones that bits 0..5 and the spacings give.
#define CODE_DOM 0x40u
#define CODE_STD 0x80u

// Labels used to coarsen the cell types from one level to the next.
#define LAB_N 0  // Neumann-like: solid, missing
#define LAB_D 1  // Dirichlet-like: air, _0_Pres
#define LAB_F 2  // unknown

__device__ __forceinline__ int nbr_slot(int ox, int oy, int oz) { return (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1); }

// Brick and local cell of (x,y,z) given relative to brick b, each in [-8,15].
__device__ __forceinline__ int locate(const int* __restrict__ nbr27, int b, int x, int y, int z, int& cell) {
  int ox = ((x + 8) >> 3) - 1, oy = ((y + 8) >> 3) - 1, oz = ((z + 8) >> 3) - 1;
  cell = (x & 7) + 8 * (y & 7) + 64 * (z & 7);
  if ((ox | oy | oz) == 0) return b;
  return nbr27[b * 27 + nbr_slot(ox, oy, oz)];
}

template <typename T>
__device__ __forceinline__ T ld_rel(const T* __restrict__ v, const int* __restrict__ nbr27, int b, int x, int y, int z) {
  int cell;
  int nb = locate(nbr27, b, x, y, z, cell);
  return nb >= 0 ? v[(size_t)nb * MAC_BC + cell] : T(0);
}

// Sum of a double over a block (any multiple of 64 threads, <= 1024);
// fixed-order, so bitwise reproducible.  Result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s += sh[i];
  }
  return s;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s = fmax(s, sh[i]);
  }
  return s;
}

__device__ __forceinline__ double ld_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide reduction epilogue.  Thread 0 of every block stores its partial;
// the last block to arrive (counted with an atomic) reduces all partials in
// block order and returns true in every thread of that block, with the total
// in *total (thread 0).  Deterministic: the summation order never depends on
// arrival order.
__device__ __forceinline__ bool grid_reduce(double part, double* __restrict__ partials, unsigned* counter, bool is_max,
                                            double* total, double* sh) {
  __shared__ int am_last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[blockIdx.x], part, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double acc = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
    double p = ld_coherent(&partials[i]);
    acc = is_max ? fmax(acc, p) : acc + p;
  }
  double t = is_max ? block_max(acc, sh) : block_sum(acc, sh);
  if (threadIdx.x == 0) {
    *total = t;
    *counter = 0u;
  }
  return true;
}
