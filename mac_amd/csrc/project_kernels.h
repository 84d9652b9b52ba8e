// project_kernels.h -- HIP kernels of the free-surface pressure projection.
//
// Layout and launch shape.  One wavefront owns one brick at a time: lane
// = x + 8y holds the brick's column of 8 cells along z in registers, so the
// z-neighbours are register reads, the x/y neighbours inside the brick are
// cross-lane shuffles (lane +-1, +-8), and only the lanes on a brick face
// load halo values from the neighbour brick.  Every load of a brick's own
// cells is a fully coalesced 64-lane row.  A launch has one wave per active
// brick (WPB waves per 256-thread block), and the dispatcher balances the
// blocks over the 256 CUs; the loops keep a grid stride for launches capped
// at GRID_CAP blocks.
//
// All vectors are zero outside the domain, so the 7-point operator needs only
// the cell's own code byte (brick.h).  Reductions write one partial per block;
// a one-block finalize kernel (k_finalize) sums them in block order
// (bitwise reproducible) and applies the CG scalar logic on the device, so a
// CG iteration never waits on the host.
#pragma once
#include "brick.h"
#include <type_traits>

struct Scal {
  double ff;      // ||f||^2
  double tol2;    // rtol^2 * ||f||^2
  double rr;      // ||r||^2 (recursive residual)
  double rz;      // r . z
  double dq;      // d . q
  double alpha, beta;
  double maxdiv;  // max |div u| after the update
  double tmp;     // scratch total (true residual)
  int iter;       // CG iterations taken
  int done;       // 0 running, 1 converged, 2 max_iter, 3 breakdown
  int rz_valid;   // 0 before the first V-cycle of a (re)start
  int max_iter;
  unsigned counter;
  int pad;
};

#define KTHREADS 256

#define WPB 4          // waves per block
#ifndef GRID_CAP
#define GRID_CAP (1 << 30)  // no cap: one wave per brick, the dispatcher balances the CUs
#endif

enum FinOp { OP_FF = 0, OP_RR0 = 1, OP_TRUE = 2, OP_DQ = 3, OP_RR = 4, OP_RZ = 5, OP_MAX = 6, OP_FF_R0 = 7 };

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_in_block() { return threadIdx.x >> 6; }

// Block order by XCD.  The dispatcher hands workgroup b of a launch to XCD
// b % 8 (round robin over the eight XCDs, each with its own L2).  The kernels
// walk their lists in the order of xcd_block(), which renumbers the blocks so
// that XCD x runs one contiguous range of logical blocks: bricks that are
// close in the (z,y,x)-sorted lists, and so mostly neighbours in space, are
// then worked on by one XCD, and the face planes a wave reads from its
// neighbour bricks are mostly in that XCD's L2 already.  A bijection for any
// grid size; MAC_XCD_ORDER = 0 keeps the hardware order.
#ifndef MAC_XCD_ORDER
#define MAC_XCD_ORDER 1
#endif
#define NXCD 8
__device__ __forceinline__ int xcd_block() {
  if (!MAC_XCD_ORDER) return blockIdx.x;
  const int G = gridDim.x, b = blockIdx.x;
  const int x = b % NXCD, q = G / NXCD, rem = G % NXCD;
  return x * q + min(x, rem) + b / NXCD;
}
__device__ __forceinline__ int first_wave() { return xcd_block() * WPB + wave_in_block(); }

// Wave-uniform loop over the active bricks of a level.  meta holds, per
// active brick, MSTRIDE ints: the brick id and its 6 face neighbours
// (-x,+x,-y,+y,-z,+z; -1 = none), so a wave fetches everything it needs about
// its brick with one scalar load and no dependent table lookups.
#define MSTRIDE 8
#define FOR_EACH_BRICK(meta, n, M)                                                            \
  for (int i_ = first_wave(); i_ < (n); i_ += gridDim.x * WPB)                               \
    if (const int* M = (meta) + (size_t)MSTRIDE * __builtin_amdgcn_readfirstlane(i_); true)

// The CG convergence flag read in the same round trip as the wave's first
// brick id (both scalar loads issue before one wait), so the early exit after
// convergence adds no latency to a live launch.  The flag is uniform, so every
// wave of a block takes the same branch.  Brick ids are >= 0; the test on b
// only keeps its load ahead of the branch.
// Launches always have n >= 1.
__device__ __forceinline__ bool converged(const int* __restrict__ meta, int stride, int n, int i, const Scal* s) {
  const int j = __builtin_amdgcn_readfirstlane(max(0, min(i, n - 1)));
  const int b = meta[(size_t)stride * j];
  const int d = s->done;
  return (b < -1) | (d != 0);
}

// Walk the 8 cells of a lane's column with the 6 neighbour values of each.
// col holds the lane's own values; ld(brick, cell) fetches a value from a
// neighbour brick (already composed, e.g. z + beta d); f(z, xm, xp, ym, yp, zm, zp).
// The four side faces of the brick are fetched lane-parallel, one full
// 64-value face plane per load (x faces indexed (y, z), y faces (x, z)), and
// handed to the edge lanes by cross-lane shuffles; the z faces are ordinary
// rows.  The neighbour-brick tests are wave-uniform, so nothing diverges.
// The six face planes a column stencil needs from the neighbour bricks.
template <typename T>
struct Halo {
  T xm, xp, ym, yp, below, above;
};
template <typename T, typename Ld>
__device__ __forceinline__ Halo<T> load_halo(const int* M, int lane, Ld ld) {
  const int x = lane & 7, y = lane >> 3;
  const int bxm = M[1], bxp = M[2], bym = M[3], byp = M[4], bzm = M[5], bzp = M[6];
  Halo<T> h;
  h.xm = (bxm >= 0) ? ld(bxm, 7 + 8 * x + 64 * y) : T(0);  // lane = y' + 8 z'
  h.xp = (bxp >= 0) ? ld(bxp, 8 * x + 64 * y) : T(0);
  h.ym = (bym >= 0) ? ld(bym, x + 56 + 64 * y) : T(0);     // lane = x' + 8 z'
  h.yp = (byp >= 0) ? ld(byp, x + 64 * y) : T(0);
  h.below = (bzm >= 0) ? ld(bzm, lane + 448) : T(0);
  h.above = (bzp >= 0) ? ld(bzp, lane) : T(0);
  return h;
}
// col_stencil with the face planes already loaded.
template <typename T, typename F>
__device__ __forceinline__ void col_apply(int lane, const T (&col)[8], const Halo<T>& h, F f) {
  const int x = lane & 7, y = lane >> 3;
#pragma unroll
  for (int z = 0; z < 8; ++z) {
    T v = col[z];
    T xm = __shfl_up(v, 1), xp = __shfl_down(v, 1), ym = __shfl_up(v, 8), yp = __shfl_down(v, 8);
    const T gxm = __shfl(h.xm, y + 8 * z), gxp = __shfl(h.xp, y + 8 * z);
    const T gym = __shfl(h.ym, x + 8 * z), gyp = __shfl(h.yp, x + 8 * z);
    if (x == 0) xm = gxm;
    if (x == 7) xp = gxp;
    if (y == 0) ym = gym;
    if (y == 7) yp = gyp;
    T zm = (z > 0) ? col[z > 0 ? z - 1 : 0] : h.below;
    T zp = (z < 7) ? col[z < 7 ? z + 1 : 7] : h.above;
    f(z, xm, xp, ym, yp, zm, zp);
  }
}

template <typename T, typename Ld, typename F>
__device__ __forceinline__ void col_stencil(const int* M, int lane, const T (&col)[8], Ld ld, F f) {
  const int x = lane & 7, y = lane >> 3;
  const int bxm = M[1], bxp = M[2], bym = M[3], byp = M[4], bzm = M[5], bzp = M[6];
  const T hxm = (bxm >= 0) ? ld(bxm, 7 + 8 * x + 64 * y) : T(0);  // lane = y' + 8 z'
  const T hxp = (bxp >= 0) ? ld(bxp, 8 * x + 64 * y) : T(0);
  const T hym = (bym >= 0) ? ld(bym, x + 56 + 64 * y) : T(0);     // lane = x' + 8 z'
  const T hyp = (byp >= 0) ? ld(byp, x + 64 * y) : T(0);
  const T below = (bzm >= 0) ? ld(bzm, lane + 448) : T(0);
  const T above = (bzp >= 0) ? ld(bzp, lane) : T(0);
#pragma unroll
  for (int z = 0; z < 8; ++z) {
    T v = col[z];
    T xm = __shfl_up(v, 1), xp = __shfl_down(v, 1), ym = __shfl_up(v, 8), yp = __shfl_down(v, 8);
    const T gxm = __shfl(hxm, y + 8 * z), gxp = __shfl(hxp, y + 8 * z);
    const T gym = __shfl(hym, x + 8 * z), gyp = __shfl(hyp, x + 8 * z);
    if (x == 0) xm = gxm;
    if (x == 7) xp = gxp;
    if (y == 0) ym = gym;
    if (y == 7) yp = gyp;
    T zm = (z > 0) ? col[z > 0 ? z - 1 : 0] : below;
    T zp = (z < 7) ? col[z < 7 ? z + 1 : 7] : above;
    f(z, xm, xp, ym, yp, zm, zp);
  }
}

// L v at a cell from its code and neighbour values (L = -A, dg = diagonal).
template <typename T>
__device__ __forceinline__ T lap_cell(unsigned cd, T vc, T xm, T xp, T ym, T yp, T zm, T zp, T ihx2, T ihy2, T ihz2, T& dg) {
  T acc = T(0), d = T(0);
  if (cd & 1u) { d += ihx2; acc += ihx2 * xm; }
  if (cd & 2u) { d += ihx2; acc += ihx2 * xp; }
  if (cd & 4u) { d += ihy2; acc += ihy2 * ym; }
  if (cd & 8u) { d += ihy2; acc += ihy2 * yp; }
  if (cd & 16u) { d += ihz2; acc += ihz2 * zm; }
  if (cd & 32u) { d += ihz2; acc += ihz2 * zp; }
  dg = d;
  return d * vc - acc;
}

// a / d for a positive diagonal d: FP32 multiplies by the hardware reciprocal
// (within 1 ulp; the same per-cell factor in every pre- and post-sweep, so the
// smoother stays symmetric), FP64 divides.
template <typename T>
__device__ __forceinline__ T div_diag(T a, T d) {
  if constexpr (sizeof(T) == 4)
    return a * __builtin_amdgcn_rcpf(d);
  else
    return a / d;
}

template <typename T>
__device__ __forceinline__ T diag_of(unsigned cd, T ihx2, T ihy2, T ihz2) {
  T d = T(0);
  if (cd & 1u) d += ihx2;
  if (cd & 2u) d += ihx2;
  if (cd & 4u) d += ihy2;
  if (cd & 8u) d += ihy2;
  if (cd & 16u) d += ihz2;
  if (cd & 32u) d += ihz2;
  return d;
}

__device__ __forceinline__ void write_partial(double v, double* __restrict__ partials, double* sh, bool is_max) {
  double t = is_max ? block_max(v, sh) : block_sum(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// One block: total of the partials in block order, then the scalar step.
// mirror (pinned host memory, or null): receives a copy of the state, so a
// host that waits for this launch reads the state without a copy of its own.
// Thread 0 loads the whole state while every thread loads its partials, so
// the launch makes one round trip to memory (a converged solve's skipped
// steps sum the stale partials and drop the total).
__global__ void __launch_bounds__(1024) k_finalize(int n, const double* __restrict__ partials, Scal* s, int op, double rtol2,
                                                   Scal* mirror) {
  __shared__ double sh[16];
  const bool mx = (op == OP_MAX);
  Scal sv;
  if (threadIdx.x == 0) sv = *s;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc = mx ? fmax(acc, partials[i]) : acc + partials[i];
  double t = mx ? block_max(acc, sh) : block_sum(acc, sh);
  if (threadIdx.x != 0) return;
  if ((op == OP_DQ || op == OP_RR || op == OP_RZ) && sv.done) {
    if (mirror) *mirror = sv;
    return;
  }
  switch (op) {
    case OP_FF:
      sv.ff = t;
      sv.tol2 = rtol2 * t;
      sv.done = (t == 0.0) ? 1 : 0;
      break;
    case OP_FF_R0:
      // ||f||^2 of a divergence pass that also wrote r = f (a zero start):
      // the OP_FF and OP_RR0 steps with ||r||^2 = ||f||^2
      sv.ff = t;
      sv.tol2 = rtol2 * t;
      sv.rr = t;
      sv.rz_valid = 0;
      sv.done = (t == 0.0 || t <= sv.tol2) ? 1 : 0;
      break;
    case OP_RR0:
      sv.rr = t;
      sv.rz_valid = 0;
      sv.done = (t <= sv.tol2) ? 1 : 0;
      break;
    case OP_TRUE:
      sv.tmp = t;
      break;
    case OP_DQ:
      sv.dq = t;
      if (t > 0.0) {
        sv.alpha = sv.rz / t;
      } else {
        sv.alpha = 0.0;
        sv.done = 3;
      }
      break;
    case OP_RR:
      sv.rr = t;
      sv.iter += 1;
      if (t <= sv.tol2) sv.done = 1;
      else if (sv.iter >= sv.max_iter) sv.done = 2;
      break;
    case OP_RZ:
      sv.beta = sv.rz_valid ? t / sv.rz : 0.0;
      sv.rz = t;
      sv.rz_valid = 1;
      break;
    case OP_MAX:
      sv.maxdiv = t;
      break;
  }
  *s = sv;
  if (mirror) *mirror = sv;
}

// ---------------------------------------------------------------- setup ----

// Level-0 stencil code and coarsening label of every cell, from the flag
// byte.  Follows Algorithms_Solve_Pressure.h:103-133: for each face, a valid
// neighbour across a const face (Data_CellType_Interface_Constant.h:46-61:
// the neighbour's const bit for +1, the cell's own for -1) is a Neumann face
// and drops out; every other face adds inv_h^2 to the diagonal; an in-domain
// neighbour also couples.  In-domain = Fluid, Inflow or Interior
// (Data_CellType_Domain_Fluid_Without_0_Pres.h:59-62).
__global__ void k_code0(int nb, const int* __restrict__ nbr, const uint8_t* __restrict__ flags, uint8_t* __restrict__ code,
                        uint8_t* __restrict__ lab) {
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned fl = flags[c];
    bool ex = fl & 0x80u;
    unsigned mat = fl & 7u;
    bool dom = ex && (mat == 1u || mat == 3u || mat == 4u);
    unsigned cd = 0;
    if (dom) {
      cd = CODE_DOM;
      const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        unsigned nf = ld_rel(flags, nbr, b, x + dx[k], y + dy[k], z + dz[k]);
        bool nvalid = nf & 0x80u;
        unsigned cbit = 8u << (k >> 1);
        bool cst = (k & 1) ? (nf & cbit) : (fl & cbit);
        bool neumann = nvalid && cst;
        if (!neumann) cd |= (1u << k);
      }
    }
    code[c] = (uint8_t)cd;
    lab[c] = dom ? LAB_F : ((!ex || mat == 2u) ? LAB_N : LAB_D);
  }
}

// Coarse label: Dirichlet if any child is Dirichlet, else unknown if any
// child is an unknown, else Neumann.
__global__ void k_coarsen_lab(const int* __restrict__ child, const uint8_t* __restrict__ labf, uint8_t* __restrict__ labc) {
  int C = blockIdx.x;
  __shared__ int ch[8];
  if (threadIdx.x < 8) ch[threadIdx.x] = child[C * 8 + threadIdx.x];
  __syncthreads();
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int X = t & 7, Y = (t >> 3) & 7, Z = t >> 6;
    bool anyD = false, anyF = false;
    for (int k = 0; k < 8; ++k) {
      int fx = 2 * X + (k & 1), fy = 2 * Y + ((k >> 1) & 1), fz = 2 * Z + (k >> 2);
      int o = (fx >> 3) + 2 * (fy >> 3) + 4 * (fz >> 3);
      int fb = ch[o];
      if (fb < 0) continue;
      unsigned l = labf[(size_t)fb * MAC_BC + (fx & 7) + 8 * (fy & 7) + 64 * (fz & 7)];
      anyD |= (l == LAB_D);
      anyF |= (l == LAB_F);
    }
    labc[(size_t)C * MAC_BC + t] = anyD ? LAB_D : (anyF ? LAB_F : LAB_N);
  }
}

// Coarse code from labels: an unknown couples to D and F neighbours.
__global__ void k_code_c(const int* __restrict__ nbr, const uint8_t* __restrict__ lab, uint8_t* __restrict__ code) {
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned cd = 0;
    if (lab[c] == LAB_F) {
      cd = CODE_DOM;
      const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        unsigned nl = ld_rel(lab, nbr, b, x + dx[k], y + dy[k], z + dz[k]);  // missing -> LAB_N
        if (nl != LAB_N) cd |= (1u << k);
      }
    }
    code[c] = (uint8_t)cd;
  }
}

// Clears up to CAP double arrays in one launch; blockIdx.y picks the array.
struct ZeroList {
  static constexpr int CAP = 48;
  double* p[CAP];
  size_t n[CAP];
  int k;
};
__global__ void k_zero_list(ZeroList zl) {
  double* p = zl.p[blockIdx.y];
  const size_t n = zl.n[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Clears n 16-byte words: a plain fill kernel, about three times the rate of
// the runtime's fill on the pressure vector, and a kernel node when captured.
__global__ void __launch_bounds__(256) k_zero16(uint4* __restrict__ p, size_t n) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = z;
}

__global__ void k_brick_count(int nb, const uint8_t* __restrict__ code, int* __restrict__ cnt) {
  int b = blockIdx.x;
  __shared__ int s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  int my = 0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) my += (code[(size_t)b * MAC_BC + t] & CODE_DOM) ? 1 : 0;
  atomicAdd(&s, my);
  __syncthreads();
  if (threadIdx.x == 0) cnt[b] = s;
}

// ------------------------------------------------------------ projection ----

// Divergence of a column, operation order of Policy_Divergence.h:20-29:
// ret = 0; ret -= u(c)/h; ret += u(c+e_x)/h; ... (no contraction).
struct DivCol {
  double d[8];
};

__device__ __forceinline__ DivCol div_column(const int* M, int lane, const double* __restrict__ u, const double* __restrict__ v,
                                             const double* __restrict__ w, double ihx, double ihy, double ihz) {
#pragma clang fp contract(off)
  const int x = lane & 7, y = lane >> 3;
  const size_t base = (size_t)M[0] * MAC_BC + lane;
  const int bxp = M[2], byp = M[4], bzp = M[6];
  double wc[8];
#pragma unroll
  for (int z = 0; z < 8; ++z) wc[z] = w[base + 64 * z];
  const double wtop = (bzp >= 0) ? w[(size_t)bzp * MAC_BC + lane] : 0.0;
  const double hup = (bxp >= 0) ? u[(size_t)bxp * MAC_BC + 8 * x + 64 * y] : 0.0;  // +x face, lane = y' + 8 z'
  const double hvp = (byp >= 0) ? v[(size_t)byp * MAC_BC + x + 64 * y] : 0.0;      // +y face, lane = x' + 8 z'
  DivCol out;
#pragma unroll
  for (int z = 0; z < 8; ++z) {
    double uc = u[base + 64 * z], vc = v[base + 64 * z];
    double up = __shfl_down(uc, 1), vp = __shfl_down(vc, 8);
    const double gu = __shfl(hup, y + 8 * z), gv = __shfl(hvp, x + 8 * z);
    if (x == 7) up = gu;
    if (y == 7) vp = gv;
    double wp = (z < 7) ? wc[z < 7 ? z + 1 : 7] : wtop;
    double ret = 0.0;
    ret -= uc * ihx;
    ret += up * ihx;
    ret -= vc * ihy;
    ret += vp * ihy;
    ret -= wc[z] * ihz;
    ret += wp * ihz;
    out.d[z] = ret;
  }
  return out;
}

// f = -div(u) on in-domain cells (L p = f with L = -A), div optionally kept;
// partial ||f||^2.
__global__ void __launch_bounds__(KTHREADS) k_div(int nact, const int* __restrict__ meta,
                                                  const uint8_t* __restrict__ code, const double* __restrict__ u,
                                                  const double* __restrict__ v, const double* __restrict__ w, double ihx,
                                                  double ihy, double ihz, double* __restrict__ f, double* __restrict__ div_out,
                                                  double* __restrict__ r0, double* __restrict__ partials) {
  __shared__ double sh[16];
  const int lane = lane_id();
  double acc = 0.0;
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    DivCol dc = div_column(M, lane, u, v, w, ihx, ihy, ihz);
    const size_t base = (size_t)b * MAC_BC + lane;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      bool dom = code[base + 64 * z] & CODE_DOM;
      double dv = dom ? dc.d[z] : 0.0;
      if (div_out) div_out[base + 64 * z] = dv;
      f[base + 64 * z] = -dv;
      if (r0) r0[base + 64 * z] = -dv;  // the residual of a zero start
      acc += dv * dv;
    }
  }
  write_partial(acc, partials, sh, false);
}

// max |div u| over the domain (the reference's post-check, :209-219).
__global__ void __launch_bounds__(KTHREADS) k_maxdiv(int nact, const int* __restrict__ meta,
                                                     const uint8_t* __restrict__ code, const double* __restrict__ u,
                                                     const double* __restrict__ v, const double* __restrict__ w, double ihx,
                                                     double ihy, double ihz, double* __restrict__ partials) {
  __shared__ double sh[16];
  const int lane = lane_id();
  double m = 0.0;
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    DivCol dc = div_column(M, lane, u, v, w, ihx, ihy, ihz);
    const size_t base = (size_t)b * MAC_BC + lane;
#pragma unroll
    for (int z = 0; z < 8; ++z)
      if (code[base + 64 * z] & CODE_DOM) m = fmax(m, fabs(dc.d[z]));
  }
  write_partial(m, partials, sh, true);
}

// r = f - L x and partial ||r||^2.  write_r = 0: true-residual check only.
__global__ void __launch_bounds__(KTHREADS) k_resid0(int nact, const int* __restrict__ meta,
                                                     const uint8_t* __restrict__ code, const double* __restrict__ f,
                                                     const double* __restrict__ xv, double* __restrict__ r, double ihx2,
                                                     double ihy2, double ihz2, int write_r, double* __restrict__ partials) {
  __shared__ double sh[16];
  const int lane = lane_id();
  double acc = 0.0;
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    const size_t base = (size_t)b * MAC_BC + lane;
    double col[8], fv[8];
    unsigned cd[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      col[z] = xv[base + 64 * z];
      cd[z] = code[base + 64 * z];
      fv[z] = f[base + 64 * z];  // ahead of the per-cell branches
    }
    col_stencil<double>(M, lane, col, [&](int nb, int cell) { return xv[(size_t)nb * MAC_BC + cell]; },
                        [&](int z, double xm, double xp, double ym, double yp, double zm, double zp) {
                          if (cd[z] & CODE_DOM) {
                            double dg;
                            double lx = lap_cell<double>(cd[z], col[z], xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
                            double rv = fv[z] - lx;
                            if (write_r) r[base + 64 * z] = rv;
                            acc += rv * rv;
                          }
                        });
  }
  write_partial(acc, partials, sh, false);
}

// d_new = z + beta d_old (formed on the fly at the neighbours too), q = L d_new,
// partial d_new . q.  TD: the storage type of the search direction (FP32 in
// the mixed method): d_new is formed in FP64 and rounded to TD, and q, the
// dot product and the later x update all use the rounded value, so x and the
// recursive residual r stay consistent.
template <typename TZ, typename TD>
__global__ void __launch_bounds__(KTHREADS) k_pmatvec(int nact, const int* __restrict__ meta,
                                                      const uint8_t* __restrict__ code, const TZ* __restrict__ zv,
                                                      const TD* __restrict__ dold, TD* __restrict__ dnew,
                                                      double* __restrict__ q, double ihx2, double ihy2, double ihz2,
                                                      double* __restrict__ partials, const Scal* s) {
  if (converged(meta, MSTRIDE, nact, first_wave(), s)) return;
  __shared__ double sh[16];
  const double beta = s->beta;
  const int lane = lane_id();
  double acc = 0.0;
  // beta = 0 (the first iteration of a start): d = z, and d_old, whatever it
  // holds, is not read
  auto sweep = [&](auto first) {
    FOR_EACH_BRICK(meta, nact, M) {
      const int b = M[0];
      const size_t base = (size_t)b * MAC_BC + lane;
      auto dval = [&](size_t c) -> double {
        if constexpr (decltype(first)::value) return (double)(TD)zv[c];
        else return (double)(TD)((double)zv[c] + beta * (double)dold[c]);
      };
      double col[8];
      unsigned cd[8];
#pragma unroll
      for (int z = 0; z < 8; ++z) {
        col[z] = dval(base + 64 * z);
        cd[z] = code[base + 64 * z];
      }
      col_stencil<double>(
          M, lane, col, [&](int nb, int cell) { return dval((size_t)nb * MAC_BC + cell); },
          [&](int z, double xm, double xp, double ym, double yp, double zm, double zp) {
            if (cd[z] & CODE_DOM) {
              double dg;
              double qv = lap_cell<double>(cd[z], col[z], xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
              dnew[base + 64 * z] = (TD)col[z];
              q[base + 64 * z] = qv;
              acc += col[z] * qv;
            }
          });
    }
  };
  if (beta == 0.0) sweep(std::true_type{});
  else sweep(std::false_type{});
  write_partial(acc, partials, sh, false);
}

// x += alpha d, r -= alpha q, partial ||r||^2.
template <typename TD>
__global__ void __launch_bounds__(KTHREADS) k_update(int nact, const int* __restrict__ meta, const uint8_t* __restrict__ code,
                                                     double* __restrict__ x, double* __restrict__ r,
                                                     const TD* __restrict__ d, const double* __restrict__ q,
                                                     double* __restrict__ partials, const Scal* s) {
  if (converged(meta, MSTRIDE, nact, first_wave(), s)) return;
  __shared__ double sh[16];
  const double alpha = s->alpha;
  const int lane = lane_id();
  double acc = 0.0;
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    const size_t base = (size_t)b * MAC_BC + lane;
    // the column's loads issued together, ahead of the per-cell branches
    unsigned cd[8];
    double xv[8], rv0[8], qv[8];
    TD dv[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const size_t c = base + 64 * z;
      cd[z] = code[c];
      xv[z] = x[c];
      rv0[z] = r[c];
      qv[z] = q[c];
      dv[z] = d[c];
    }
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const size_t c = base + 64 * z;
      if (cd[z] & CODE_DOM) {
        x[c] = xv[z] + alpha * (double)dv[z];
        double rv = rv0[z] - alpha * qv[z];
        r[c] = rv;
        acc += rv * rv;
      }
    }
  }
  write_partial(acc, partials, sh, false);
}

// Damped Jacobi sweep: zout = zin + omega D^-1 (r - L zin); zin == nullptr
// means a zero start (zout = omega D^-1 r).  want_dot: partial r . zout.
template <typename TR, typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_jacobi(int nact, const int* __restrict__ meta,
                                                     const uint8_t* __restrict__ code, const TR* __restrict__ r,
                                                     const TZ* __restrict__ zin, TZ* __restrict__ zout, TZ ihx2, TZ ihy2,
                                                     TZ ihz2, TZ omega, int want_dot, double* __restrict__ partials,
                                                     const Scal* s) {
  if (converged(meta, MSTRIDE, nact, first_wave(), s)) return;
  __shared__ double sh[16];
  const int lane = lane_id();
  double acc = 0.0;
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    const size_t base = (size_t)b * MAC_BC + lane;
    unsigned cd[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) cd[z] = code[base + 64 * z];
    if (zin) {
      // r loaded with the codes and the column, ahead of the sweep (its
      // loads inside the per-cell branches would wait one by one)
      TR rr[8];
#pragma unroll
      for (int z = 0; z < 8; ++z) rr[z] = r[base + 64 * z];
      TZ col[8];
#pragma unroll
      for (int z = 0; z < 8; ++z) col[z] = zin[base + 64 * z];
      col_stencil<TZ>(M, lane, col, [&](int nb, int cell) { return zin[(size_t)nb * MAC_BC + cell]; },
                      [&](int z, TZ xm, TZ xp, TZ ym, TZ yp, TZ zm, TZ zp) {
                        TZ out = TZ(0);
                        if (cd[z] & CODE_DOM) {
                          TZ dg;
                          TZ lz = lap_cell<TZ>(cd[z], col[z], xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
                          TR rv = rr[z];
                          out = dg > TZ(0) ? col[z] + div_diag<TZ>(omega * ((TZ)rv - lz), dg) : TZ(0);
                          if (want_dot) acc += (double)rv * (double)out;
                        }
                        zout[base + 64 * z] = out;
                      });
    } else {
#pragma unroll
      for (int z = 0; z < 8; ++z) {
        TZ out = TZ(0);
        if (cd[z] & CODE_DOM) {
          TR rv = r[base + 64 * z];
          TZ dg = diag_of<TZ>(cd[z], ihx2, ihy2, ihz2);
          out = dg > TZ(0) ? div_diag<TZ>(omega * (TZ)rv, dg) : TZ(0);
          if (want_dot) acc += (double)rv * (double)out;
        }
        zout[base + 64 * z] = out;
      }
    }
  }
  if (want_dot) write_partial(acc, partials, sh, false);
}

// The coarse-level operator (levels >= 1).  Level l+1's operator is the
// Galerkin operator P0^T A_l P0 / 16 of the piecewise-constant aggregation P0
// (a coarse cell's value on its 8 children), which equals the rediscretised
// 7-point operator away from boundaries: the coupling to a face neighbour is
// the sum of the fine couplings across the coarse face / 16 (0 toward a cell
// that is not an unknown: its share stays in the diagonal), the diagonal
// (the children's diagonals - the couplings between children, both ways) /
// 16.  A Neumann (const) face or a Dirichlet cell inside a coarse cell so
// enters the coarse operator through the fine couplings it removes, and the
// coarse grids see the boundary where the fine grid has it; with labels alone
// a floor face inside the coarse cells made the multigrid take 24 and 34
// iterations on deep pools 256 and 512 cells wide.  Stored per level as NCF
// arrays of one value per cell (couplings -x,+x,-y,+y,-z,+z, then the
// diagonal), in FP64 and FP32.
#define NCF 7
template <typename T>
struct Coef {
  T c[6];
  T dg;
};
template <typename T>
__device__ __forceinline__ Coef<T> load_coef(const T* __restrict__ cf, size_t ncell, size_t i) {
  Coef<T> k;
#pragma unroll
  for (int q = 0; q < 6; ++q) k.c[q] = cf[(size_t)q * ncell + i];
  k.dg = cf[(size_t)6 * ncell + i];
  return k;
}
// L v at a coarse cell: dg v - sum of couplings x neighbours
template <typename T>
__device__ __forceinline__ T lap_coef(const Coef<T>& k, T vc, T xm, T xp, T ym, T yp, T zm, T zp) {
  T acc = k.c[0] * xm;
  acc += k.c[1] * xp;
  acc += k.c[2] * ym;
  acc += k.c[3] * yp;
  acc += k.c[4] * zm;
  acc += k.c[5] * zp;
  return k.dg * vc - acc;
}

// A coarse cell whose coefficients are the rediscretised 7-point ones (the
// code byte's CODE_STD bit: couplings inv_h^2 on the faces of the code, the
// code's diagonal) takes them from the code and the level's squared inverse
// spacings instead of reading the NCF arrays; most coarse cells are such
// cells, so the arrays are read near boundaries only.
template <typename T>
struct Ih {
  T x, y, z;
};
template <typename T>
__device__ __forceinline__ Coef<T> coef_at(unsigned cd, const T* __restrict__ cf, size_t ncell, size_t i, Ih<T> ih) {
  if (cd & CODE_STD) {
    Coef<T> k;
    k.c[0] = (cd & 1u) ? ih.x : T(0);
    k.c[1] = (cd & 2u) ? ih.x : T(0);
    k.c[2] = (cd & 4u) ? ih.y : T(0);
    k.c[3] = (cd & 8u) ? ih.y : T(0);
    k.c[4] = (cd & 16u) ? ih.z : T(0);
    k.c[5] = (cd & 32u) ? ih.z : T(0);
    k.dg = diag_of<T>(cd, ih.x, ih.y, ih.z);
    return k;
  }
  return load_coef<T>(cf, ncell, i);
}

// One coarse brick per block, one coarse cell per thread: level l+1's
// coefficients from level l's (cff, or for l = 0 the stencil codes and the
// squared inverse spacings).  childc: the 8 level-l bricks of each level-(l+1)
// brick (octant order, -1 = none); nbrf / nbrc: the 27-neighbour tables.
__global__ void k_coarse_coef(const uint8_t* __restrict__ codef, const int* __restrict__ nbrf, const double* __restrict__ cff,
                              size_t ncf, double ihx2, double ihy2, double ihz2, const int* __restrict__ childc,
                              uint8_t* __restrict__ codec, const int* __restrict__ nbrc, double* __restrict__ cfc,
                              float* __restrict__ cfc32, size_t ncc) {
  const int C = blockIdx.x;
  const double ih2[3] = {ihx2, ihy2, ihz2};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t ci = (size_t)C * MAC_BC + t;
    double cs[6] = {0, 0, 0, 0, 0, 0}, d = 0.0;
    const int X = t & 7, Y = (t >> 3) & 7, Z = t >> 6;
    if (codec[ci] & CODE_DOM) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, e = q >> 2;
        const int fx = 2 * X + a, fy = 2 * Y + b, fz = 2 * Z + e;  // 0..15 over the 8 children
        const int fb = childc[(size_t)C * 8 + (fx >> 3) + 2 * (fy >> 3) + 4 * (fz >> 3)];
        if (fb < 0) continue;
        const int lx = fx & 7, ly = fy & 7, lz = fz & 7;
        const size_t fi = (size_t)fb * MAC_BC + lx + 8 * ly + 64 * lz;
        const unsigned cdf = codef[fi];
        if (!(cdf & CODE_DOM)) continue;
        d += cff ? cff[(size_t)6 * ncf + fi] : diag_of<double>(cdf, ihx2, ihy2, ihz2);
        const bool leaves[6] = {a == 0, a == 1, b == 0, b == 1, e == 0, e == 1};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          double w;
          if (cff) {
            w = cff[(size_t)k * ncf + fi];
          } else {
            w = 0.0;
            if (cdf & (1u << k)) {
              const int sx = (k == 0) ? -1 : (k == 1) ? 1 : 0, sy = (k == 2) ? -1 : (k == 3) ? 1 : 0,
                        sz = (k == 4) ? -1 : (k == 5) ? 1 : 0;
              const unsigned cg = ld_rel(codef, nbrf, fb, lx + sx, ly + sy, lz + sz);
              if (cg & CODE_DOM) w = ih2[k >> 1];
            }
          }
          if (leaves[k]) cs[k] += w;
          else d -= w;
        }
      }
      // the coarse level's own spacings; a cell within 1e-12 of the
      // rediscretised operator (on every face toward an unknown) takes it
      // exactly and is marked CODE_STD
      const unsigned cc = codec[ci];
      const double ic[3] = {0.25 * ihx2, 0.25 * ihy2, 0.25 * ihz2};
      bool stdc = true;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int sx = (k == 0) ? -1 : (k == 1) ? 1 : 0, sy = (k == 2) ? -1 : (k == 3) ? 1 : 0,
                  sz = (k == 4) ? -1 : (k == 5) ? 1 : 0;
        const unsigned cn = ld_rel(codec, nbrc, C, X + sx, Y + sy, Z + sz);
        cs[k] = (cn & CODE_DOM) ? cs[k] / 16.0 : 0.0;
        const double want = ((cn & CODE_DOM) && (cc & (1u << k))) ? ic[k >> 1] : 0.0;
        stdc = stdc && fabs(cs[k] - want) <= 1e-12 * ic[k >> 1];
      }
      d /= 16.0;
      const double dstd = diag_of<double>(cc, ic[0], ic[1], ic[2]);
      stdc = stdc && fabs(d - dstd) <= 1e-12 * dstd;
      if (stdc) {
#pragma unroll
        for (int k = 0; k < 6; ++k) cs[k] = cs[k] != 0.0 ? ic[k >> 1] : 0.0;
        d = dstd;
      }
      // codec is read by the neighbouring threads' ld_rel above only for
      // CODE_DOM, which the mark leaves alone
      if (stdc) codec[ci] = (uint8_t)(cc | CODE_STD);
    } else {
      codec[ci] = (uint8_t)(codec[ci] | CODE_STD);  // all-zero coefficients, as the code gives
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      cfc[(size_t)k * ncc + ci] = cs[k];
      cfc32[(size_t)k * ncc + ci] = (float)cs[k];
    }
    cfc[(size_t)6 * ncc + ci] = d;
    cfc32[(size_t)6 * ncc + ci] = (float)d;
  }
}

// Block-Jacobi smoothing step: the brick's unknowns take m damped-Jacobi
// sweeps (weights om[0..m-1]) among themselves with the neighbour bricks'
// face planes frozen at zin, i.e. zout = zin + q(D^-1 A_BB) D^-1 (r - A zin)
// on each brick (q the sweeps' polynomial) scaled by w.damp, a symmetric
// block smoother.  zin == nullptr: from zero.
struct BlockW {
  int m;
  float om[8];
  double om64[8];
  double damp;  // zout = zin + damp (block correction)
};
// The block step runs with one brick per 256-thread block (the slab form):
// wave w holds z-planes 2w and 2w+1 of the brick in the lane = x + 8y
// columns, so a sweep's serial chain is two cells, and the coarse levels,
// whose few bricks would leave most of the chip idle with a wave per brick,
// get four waves per brick.  Each sweep publishes the wave's two planes in
// LDS (double buffered, one barrier per sweep) and takes the planes 2w-1 and
// 2w+2 of the neighbouring waves from there; in-brick x/y neighbours are
// cross-lane shuffles, and the frozen face values of the neighbour bricks
// come in as gx / gy (x / y edge lanes), below (wave 0) and above (wave 3).
// idg: the inverse diagonal per cell (FP32; FP64 divides), 0 off the domain.
template <typename T>
struct SlabLDS {
  T pl[2][WPB][2][64];  // [buffer][wave][low/high plane][lane]
};
template <typename TZ>
__device__ __forceinline__ void slab_sweeps(int wv, int lane, const unsigned (&cd)[2], const Coef<TZ> (&K)[2],
                                            const TZ (&rv)[2], const TZ (&idg)[2], TZ (&col)[2], const TZ (&gx)[2],
                                            const TZ (&gy)[2], TZ below, TZ above, const BlockW& w, SlabLDS<TZ>& L) {
  const int x = lane & 7, y = lane >> 3;
  for (int k = 0; k < w.m; ++k) {
    const TZ om = sizeof(TZ) == 4 ? (TZ)w.om[k] : (TZ)w.om64[k];
    auto& P = L.pl[k & 1];
    P[wv][0][lane] = col[0];
    P[wv][1][lane] = col[1];
    __syncthreads();
    const TZ lo = wv > 0 ? P[wv - 1][1][lane] : below;
    const TZ hi = wv < WPB - 1 ? P[wv + 1][0][lane] : above;
    TZ nw[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const TZ v = col[j];
      TZ xm = __shfl_up(v, 1), xp = __shfl_down(v, 1), ym = __shfl_up(v, 8), yp = __shfl_down(v, 8);
      if (x == 0) xm = gx[j];
      if (x == 7) xp = gx[j];
      if (y == 0) ym = gy[j];
      if (y == 7) yp = gy[j];
      const TZ zm = j ? col[0] : lo;
      const TZ zp = j ? hi : col[1];
      TZ out = TZ(0);
      if (cd[j] & CODE_DOM) {
        const TZ lz = lap_coef<TZ>(K[j], v, xm, xp, ym, yp, zm, zp);
        if constexpr (sizeof(TZ) == 4)
          out = idg[j] != TZ(0) ? v + om * (rv[j] - lz) * idg[j] : TZ(0);
        else
          out = K[j].dg > TZ(0) ? v + om * (rv[j] - lz) / K[j].dg : TZ(0);
      }
      nw[j] = out;
    }
    col[0] = nw[0];
    col[1] = nw[1];
  }
}

template <typename TZ>
__device__ __forceinline__ void inv_diag_slab(const unsigned (&cd)[2], const Coef<TZ> (&K)[2], TZ (&idg)[2]) {
#pragma unroll
  for (int j = 0; j < 2; ++j) idg[j] = ((cd[j] & CODE_DOM) && K[j].dg > TZ(0)) ? TZ(1) / K[j].dg : TZ(0);
}

// The block step, one block per active brick.
template <typename TR, typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_jacobi_block(int nact, const int* __restrict__ meta,
                                                                const uint8_t* __restrict__ code, const TR* __restrict__ r,
                                                                const TZ* __restrict__ zin, TZ* __restrict__ zout,
                                                                const TZ* __restrict__ cf, size_t ncell, Ih<TZ> ih,
                                                                BlockW w, const Scal* s) {
  const int i = xcd_block();
  if (converged(meta, MSTRIDE, nact, i, s)) return;
  __shared__ SlabLDS<TZ> L;
  const int lane = lane_id(), wv = wave_in_block();
  const int x = lane & 7, y = lane >> 3;
  const int* M = meta + (size_t)MSTRIDE * i;
  const size_t base = (size_t)M[0] * MAC_BC + lane + 128 * wv;
  unsigned cd[2];
  TZ col[2], rv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    cd[j] = code[base + 64 * j];
    rv[j] = (TZ)r[base + 64 * j];
    col[j] = zin ? zin[base + 64 * j] : TZ(0);
  }
  TZ gx[2] = {TZ(0), TZ(0)}, gy[2] = {TZ(0), TZ(0)}, below = TZ(0), above = TZ(0);
  if (zin) {
    auto ld = [&](int nb, int cell) { return zin[(size_t)nb * MAC_BC + cell]; };
    const TZ hxm = M[1] >= 0 ? ld(M[1], 7 + 8 * x + 64 * y) : TZ(0);  // lane = y' + 8 z'
    const TZ hxp = M[2] >= 0 ? ld(M[2], 8 * x + 64 * y) : TZ(0);
    const TZ hym = M[3] >= 0 ? ld(M[3], x + 56 + 64 * y) : TZ(0);  // lane = x' + 8 z'
    const TZ hyp = M[4] >= 0 ? ld(M[4], x + 64 * y) : TZ(0);
    if (wv == 0 && M[5] >= 0) below = ld(M[5], lane + 448);
    if (wv == WPB - 1 && M[6] >= 0) above = ld(M[6], lane);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int z = 2 * wv + j;
      const TZ gxm = __shfl(hxm, y + 8 * z), gxp = __shfl(hxp, y + 8 * z);
      const TZ gym = __shfl(hym, x + 8 * z), gyp = __shfl(hyp, x + 8 * z);
      gx[j] = x == 0 ? gxm : gxp;
      gy[j] = y == 0 ? gym : gyp;
    }
  }
  const TZ col0[2] = {col[0], col[1]};
  const Coef<TZ> K[2] = {coef_at<TZ>(cd[0], cf, ncell, base, ih), coef_at<TZ>(cd[1], cf, ncell, base + 64, ih)};
  TZ idg[2];
  inv_diag_slab<TZ>(cd, K, idg);
  slab_sweeps<TZ>(wv, lane, cd, K, rv, idg, col, gx, gy, below, above, w, L);
  const TZ dmp = (TZ)w.damp;
#pragma unroll
  for (int j = 0; j < 2; ++j) zout[base + 64 * j] = col0[j] + dmp * (col[j] - col0[j]);
}

// A coarse brick in the slab form without sweeps: wave w's planes 2w, 2w+1
// of zv with their six neighbours (the planes 2w-1 and 2w+2 read from
// memory, the x / y face planes of the neighbour bricks lane-parallel).
template <typename TZ>
struct SlabCells {
  unsigned cd[2];
  TZ v[2], xm[2], xp[2], ym[2], yp[2], zm[2], zp[2];
  __device__ __forceinline__ void load(const int* M, int lane, int wv, const uint8_t* __restrict__ code,
                                       const TZ* __restrict__ zv) {
    const int x = lane & 7, y = lane >> 3;
    const int b = M[0];
    const size_t base = (size_t)b * MAC_BC + lane + 128 * wv;
    auto ld = [&](int nb, int cell) { return nb >= 0 ? zv[(size_t)nb * MAC_BC + cell] : TZ(0); };
    const TZ hxm = ld(M[1], 7 + 8 * x + 64 * y), hxp = ld(M[2], 8 * x + 64 * y);  // lane = y' + 8 z'
    const TZ hym = ld(M[3], x + 56 + 64 * y), hyp = ld(M[4], x + 64 * y);         // lane = x' + 8 z'
    const TZ below = wv > 0 ? zv[base - 64] : ld(M[5], lane + 448);
    const TZ above = wv < WPB - 1 ? zv[base + 128] : ld(M[6], lane);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cd[j] = code[base + 64 * j];
      v[j] = zv[base + 64 * j];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int z = 2 * wv + j;
      TZ a = __shfl_up(v[j], 1), bq = __shfl_down(v[j], 1), c = __shfl_up(v[j], 8), d = __shfl_down(v[j], 8);
      const TZ gxm = __shfl(hxm, y + 8 * z), gxp = __shfl(hxp, y + 8 * z);
      const TZ gym = __shfl(hym, x + 8 * z), gyp = __shfl(hyp, x + 8 * z);
      xm[j] = x == 0 ? gxm : a;
      xp[j] = x == 7 ? gxp : bq;
      ym[j] = y == 0 ? gym : c;
      yp[j] = y == 7 ? gyp : d;
    }
    zm[0] = below;
    zp[0] = v[1];
    zm[1] = v[0];
    zp[1] = above;
  }
};

// res = r - L z on a coarse level, one block per active brick.
template <typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_residual_coef(int nact, const int* __restrict__ meta,
                                                            const uint8_t* __restrict__ code, const TZ* __restrict__ r,
                                                            const TZ* __restrict__ zv, TZ* __restrict__ res,
                                                            const TZ* __restrict__ cf, size_t ncell, Ih<TZ> ih,
                                                            const Scal* s) {
  const int i = xcd_block();
  if (converged(meta, MSTRIDE, nact, i, s)) return;
  const int lane = lane_id(), wv = wave_in_block();
  const int* M = meta + (size_t)MSTRIDE * i;
  const size_t base = (size_t)M[0] * MAC_BC + lane + 128 * wv;
  const TZ rv[2] = {r[base], r[base + 64]};
  SlabCells<TZ> sc;
  sc.load(M, lane, wv, code, zv);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    TZ out = TZ(0);
    if (sc.cd[j] & CODE_DOM) {
      const Coef<TZ> K = coef_at<TZ>(sc.cd[j], cf, ncell, base + 64 * j, ih);
      out = rv[j] - lap_coef<TZ>(K, sc.v[j], sc.xm[j], sc.xp[j], sc.ym[j], sc.yp[j], sc.zm[j], sc.zp[j]);
    }
    res[base + 64 * j] = out;
  }
}

// One damped-Jacobi sweep on a coarse level: zout = zin + omega D^-1 (r - L zin);
// zin == nullptr: from zero (zout = omega D^-1 r).  One block per active brick.
template <typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_jacobi_coef(int nact, const int* __restrict__ meta,
                                                          const uint8_t* __restrict__ code, const TZ* __restrict__ r,
                                                          const TZ* __restrict__ zin, TZ* __restrict__ zout,
                                                          const TZ* __restrict__ cf, size_t ncell, Ih<TZ> ih, TZ omega,
                                                          const Scal* s) {
  const int i = xcd_block();
  if (converged(meta, MSTRIDE, nact, i, s)) return;
  const int lane = lane_id(), wv = wave_in_block();
  const int* M = meta + (size_t)MSTRIDE * i;
  const size_t base = (size_t)M[0] * MAC_BC + lane + 128 * wv;
  SlabCells<TZ> sc;
  if (zin) {
    sc.load(M, lane, wv, code, zin);
  } else {
#pragma unroll
    for (int j = 0; j < 2; ++j) sc.cd[j] = code[base + 64 * j];
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    TZ out = TZ(0);
    if (sc.cd[j] & CODE_DOM) {
      const Coef<TZ> K = coef_at<TZ>(sc.cd[j], cf, ncell, base + 64 * j, ih);
      const TZ rv = r[base + 64 * j];
      if (K.dg > TZ(0)) {
        if (zin)
          out = sc.v[j] + div_diag<TZ>(omega * (rv - lap_coef<TZ>(K, sc.v[j], sc.xm[j], sc.xp[j], sc.ym[j], sc.yp[j], sc.zm[j],
                                                                      sc.zp[j])),
                                       K.dg);
        else
          out = div_diag<TZ>(omega * rv, K.dg);
      }
    }
    zout[base + 64 * j] = out;
  }
}

// Two damped-Jacobi sweeps from a zero start in one pass:
//   z1 = omega D^-1 r (formed on the fly, also at the halo), zout = z1 + omega2 D^-1 (r - L z1).
// Bitwise equal to two k_jacobi launches (same expressions, same order).
template <typename TR, typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_jacobi2(int nact, const int* __restrict__ meta,
                                                      const uint8_t* __restrict__ code, const TR* __restrict__ r,
                                                      TZ* __restrict__ zout, TZ ihx2, TZ ihy2, TZ ihz2, TZ omega,
                                                      TZ omega2, const Scal* s) {
  if (converged(meta, MSTRIDE, nact, first_wave(), s)) return;
  const int lane = lane_id();
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    const size_t base = (size_t)b * MAC_BC + lane;
    unsigned cd[8];
    TZ z1[8], rv[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      cd[z] = code[base + 64 * z];
      rv[z] = (TZ)r[base + 64 * z];
      TZ dg = diag_of<TZ>(cd[z], ihx2, ihy2, ihz2);
      z1[z] = ((cd[z] & CODE_DOM) && dg > TZ(0)) ? div_diag<TZ>(omega * rv[z], dg) : TZ(0);
    }
    col_stencil<TZ>(
        M, lane, z1,
        [&](int nb, int cell) {
          const size_t j = (size_t)nb * MAC_BC + cell;
          const unsigned c2 = code[j];
          const TZ rj = (TZ)r[j];  // loaded with the code, not behind its test
          TZ d2 = diag_of<TZ>(c2, ihx2, ihy2, ihz2);
          return ((c2 & CODE_DOM) && d2 > TZ(0)) ? div_diag<TZ>(omega * rj, d2) : TZ(0);
        },
        [&](int z, TZ xm, TZ xp, TZ ym, TZ yp, TZ zm, TZ zp) {
          TZ out = TZ(0);
          if (cd[z] & CODE_DOM) {
            TZ dg;
            TZ lz = lap_cell<TZ>(cd[z], z1[z], xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
            out = dg > TZ(0) ? z1[z] + div_diag<TZ>(omega2 * (rv[z] - lz), dg) : TZ(0);
          }
          zout[base + 64 * z] = out;
        });
  }
}

// Restriction, separable: r_c = (1/8) W^T res over the coarse domain with the
// 1-D weights (1/4, 3/4, 3/4, 1/4) on fine cells 2X-1 .. 2X+2.  One block per
// active coarse brick, in two halves of 4 coarse z-planes: the 18x18x10 fine
// residual block a half needs is staged in LDS with all loads in flight at
// once, then reduced along x, y and z in LDS.
// rmeta per active coarse brick (RSTRIDE ints): brick id, then at +4 the
// 64 fine bricks covering fine cells -1..16 relative to the coarse brick's
// first child, fb[bx + 4 by + 16 bz] with b* = 0..3 for fine brick offsets
// -1..2 (-1 = none).
// zsm != nullptr fuses the coarse level's first block-smoothing step from
// zero (k_jacobi_block with zin == nullptr: the neighbour bricks are frozen
// at 0, so the step needs nothing outside the brick): wave 0 takes the
// restricted brick from LDS and writes zsm.
#define RSTRIDE 68
// The coarse level's first block step from zero on the restricted brick in
// LDS (rcl), in the slab form, run by all four waves of the block.
template <typename T>
__device__ __forceinline__ void smooth_from_zero(int C, const uint8_t* __restrict__ codec, const T* rcl,
                                                      T* __restrict__ zsm, const T* __restrict__ cf, size_t ncell,
                                                      Ih<T> ih, const BlockW& w, SlabLDS<T>& L) {
  const int lane = lane_id(), wv = wave_in_block();
  const size_t base = (size_t)C * MAC_BC + lane + 128 * wv;
  unsigned cd[2];
  T rv[2], col[2], gz[2], idg[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    cd[j] = codec[base + 64 * j];
    rv[j] = rcl[lane + 128 * wv + 64 * j];
    col[j] = T(0);
    gz[j] = T(0);
  }
  const Coef<T> K[2] = {coef_at<T>(cd[0], cf, ncell, base, ih), coef_at<T>(cd[1], cf, ncell, base + 64, ih)};
  inv_diag_slab<T>(cd, K, idg);
  slab_sweeps<T>(wv, lane, cd, K, rv, idg, col, gz, gz, T(0), T(0), w, L);
  const T dmp = (T)w.damp;
#pragma unroll
  for (int j = 0; j < 2; ++j) zsm[base + 64 * j] = T(0) + dmp * (col[j] - T(0));
}

template <typename T>
__global__ void __launch_bounds__(KTHREADS) k_restrict_sep(const int* __restrict__ rmeta, const uint8_t* __restrict__ codec,
                                                           const T* __restrict__ resf, T* __restrict__ rc, const Scal* s,
                                                           T* __restrict__ zsm, const T* __restrict__ cfc, size_t ncc, Ih<T> ih,
                                                           BlockW w) {
  if (converged(rmeta, RSTRIDE, (int)gridDim.x, xcd_block(), s)) return;
  __shared__ T rcl[MAC_BC];
  if constexpr (sizeof(T) == 4) {
    // FP32: both halves at once (the 18^3 fine region fits in LDS), so the
    // global loads issue together and the three passes need three barriers
    constexpr int NA = 18 * 18 * 18;  // [fz][fy][fx], fine x, y, z in -1..16
    constexpr int NA1 = 18 * 18 * 8;  // [fz][fy][X]
    __shared__ int fba[64];
    __shared__ T FA[NA];
    __shared__ T ta1[NA1];
    T* ta2 = FA;  // [fz][Y][X], 18*64, after FA is consumed
    const int* M = rmeta + (size_t)RSTRIDE * xcd_block();
    const int C = M[0];
    if (threadIdx.x < 64) fba[threadIdx.x] = M[4 + threadIdx.x];
    __syncthreads();
    // Each thread loads whole fine z-columns (fz = -1..16) of the region: the
    // 16x16 interior columns one per thread, the 68 columns of the ring
    // around them one each for threads 0..67; along a column only the brick
    // slot (at fz = 0 and 16) and the cell offset (+64) change.
    const int t = threadIdx.x;
    const int ix = t & 15, iy = t >> 4;
    int rx, ry;  // the ring column of threads 0..67
    if (t < 18) {
      rx = t - 1;
      ry = -1;
    } else if (t < 36) {
      rx = t - 19;
      ry = 16;
    } else if (t < 52) {
      rx = -1;
      ry = t - 36;
    } else {
      rx = 16;
      ry = t - 52;
    }
    const bool ring = t < 68;
    auto col_base = [&](int gx, int gy, int& slot, int& cell) {
      slot = ((gx + 8) >> 3) + 4 * ((gy + 8) >> 3);
      cell = (gx & 7) + 8 * (gy & 7);
    };
    int si, ci, sr, cr;
    col_base(ix, iy, si, ci);
    col_base(rx, ry, sr, cr);
    T vi[18], vr[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      const int gz = k - 1, zs = 16 * ((gz + 8) >> 3), zc = 64 * (gz & 7);
      const int fi = fba[si + zs];
      vi[k] = fi >= 0 ? resf[(size_t)fi * MAC_BC + ci + zc] : T(0);
      vr[k] = T(0);
      if (ring) {
        const int fr = fba[sr + zs];
        if (fr >= 0) vr[k] = resf[(size_t)fr * MAC_BC + cr + zc];
      }
    }
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      FA[k * 324 + (iy + 1) * 18 + ix + 1] = vi[k];
      if (ring) FA[k * 324 + (ry + 1) * 18 + rx + 1] = vr[k];
    }
    __syncthreads();
    // x pass: thread t takes X = t & 7 on the rows (fy, fz) = r, r + 32, ...
    {
      const int X = t & 7;
      int fy = t >> 3, fz = 0;  // r = t >> 3 < 32: fy = r % 18, fz = r / 18
      if (fy >= 18) {
        fy -= 18;
        fz = 1;
      }
      for (int r = t >> 3; r < 324; r += 32) {
        const T* row = &FA[(fz * 18 + fy) * 18 + 2 * X];
        ta1[r * 8 + X] = T(0.25) * row[0] + T(0.75) * row[1] + T(0.75) * row[2] + T(0.25) * row[3];
        fy += 14;  // r += 32 = one fz and 14 fy
        fz += 1;
        if (fy >= 18) {
          fy -= 18;
          fz += 1;
        }
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 18 * 64; i += blockDim.x) {
      const int X = i & 7, Y = (i >> 3) & 7, fz = i >> 6;
      const T* col = &ta1[(fz * 18 + 2 * Y) * 8 + X];
      ta2[i] = T(0.25) * col[0] + T(0.75) * col[8] + T(0.75) * col[16] + T(0.25) * col[24];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
      const int X = t & 7, Y = (t >> 3) & 7, Z = t >> 6;
      const size_t c = (size_t)C * MAC_BC + t;
      T acc = T(0);
      if (codec[c] & CODE_DOM) {
        const T* col = &ta2[(2 * Z) * 64 + Y * 8 + X];
        acc = (T(0.25) * col[0] + T(0.75) * col[64] + T(0.75) * col[128] + T(0.25) * col[192]) * T(0.125);
      }
      rc[c] = acc;
      rcl[t] = acc;
    }
    if (zsm) {
      __shared__ SlabLDS<T> L;
      __syncthreads();
      smooth_from_zero<T>(C, codec, rcl, zsm, cfc, ncc, ih, w, L);
    }
    return;
  }
  constexpr int NF = 10 * 18 * 18;  // [fz][fy][fx], fine x,y in -1..16, 10 fine z per half
  constexpr int NT1 = 10 * 18 * 8;  // [fz][fy][X]
  __shared__ int fb[64];
  __shared__ T F[NF];
  __shared__ T t1[NT1];
  T* t2 = F;  // [fz][Y][X], 10*64, after F is consumed
  const int* M = rmeta + (size_t)RSTRIDE * xcd_block();
  const int C = M[0];
  if (threadIdx.x < 64) fb[threadIdx.x] = M[4 + threadIdx.x];
  for (int half = 0; half < 2; ++half) {
    __syncthreads();
    T v[(NF + KTHREADS - 1) / KTHREADS];
#pragma unroll
    for (int k = 0; k < (NF + KTHREADS - 1) / KTHREADS; ++k) {
      const int i = threadIdx.x + KTHREADS * k;
      v[k] = T(0);
      if (i < NF) {
        const int gx = i % 18 - 1, gy = (i / 18) % 18 - 1, gz = 8 * half + i / 324 - 1;
        const int f = fb[((gx + 8) >> 3) + 4 * ((gy + 8) >> 3) + 16 * ((gz + 8) >> 3)];
        if (f >= 0) v[k] = resf[(size_t)f * MAC_BC + (gx & 7) + 8 * (gy & 7) + 64 * (gz & 7)];
      }
    }
#pragma unroll
    for (int k = 0; k < (NF + KTHREADS - 1) / KTHREADS; ++k) {
      const int i = threadIdx.x + KTHREADS * k;
      if (i < NF) F[i] = v[k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NT1; i += blockDim.x) {
      const int X = i & 7, fy = (i >> 3) % 18, fz = (i >> 3) / 18;
      const T* row = &F[(fz * 18 + fy) * 18 + 2 * X];
      t1[i] = T(0.25) * row[0] + T(0.75) * row[1] + T(0.75) * row[2] + T(0.25) * row[3];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 10 * 64; i += blockDim.x) {
      const int X = i & 7, Y = (i >> 3) & 7, fz = i >> 6;
      const T* col = &t1[(fz * 18 + 2 * Y) * 8 + X];
      t2[i] = T(0.25) * col[0] + T(0.75) * col[8] + T(0.75) * col[16] + T(0.25) * col[24];
    }
    __syncthreads();
    {
      const int t = threadIdx.x;  // 256 = 4 coarse z-planes of 64
      const int X = t & 7, Y = (t >> 3) & 7, zz = t >> 6;
      const size_t c = (size_t)C * MAC_BC + X + 8 * Y + 64 * (4 * half + zz);
      T acc = T(0);
      if (codec[c] & CODE_DOM) {
        const T* col = &t2[(2 * zz) * 64 + Y * 8 + X];
        acc = (T(0.25) * col[0] + T(0.75) * col[64] + T(0.75) * col[128] + T(0.25) * col[192]) * T(0.125);
      }
      rc[c] = acc;
      rcl[X + 8 * Y + 64 * (4 * half + zz)] = acc;
    }
  }
  if (zsm) {
    __shared__ SlabLDS<T> L;
    __syncthreads();
    smooth_from_zero<T>(C, codec, rcl, zsm, cfc, ncc, ih, w, L);
  }
}

// Prolongation, column form: z_f += W z_c on the fine domain.  Each lane
// gathers the 4 coarse columns (X0|X1, Y0|Y1) over the 6 coarse z-levels its
// 8 fine cells need, then forms the trilinear weights 3/4 (parent side) and
// 1/4 per direction.  The coarse cells a fine brick touches lie in the parent
// brick and its neighbours toward the brick's octant (2x2x2 bricks), whose
// ids are wave-uniform, so every gather is a plain indexed load.
// pmeta per active fine brick (PSTRIDE ints): brick id, octant in its
// parent, then the 8 coarse bricks bb[sx + 2 sy + 4 sz] (sx = 1: one step
// from the parent toward the octant's side in x, etc.; -1 = none).
#define PSTRIDE 16
template <typename T>
__global__ void __launch_bounds__(KTHREADS) k_prolong_col(int nact, const int* __restrict__ pmeta,
                                                          const uint8_t* __restrict__ codef, const T* __restrict__ zc,
                                                          T* __restrict__ zf, const Scal* s) {
  if (converged(pmeta, PSTRIDE, nact, first_wave(), s)) return;
  const int lane = lane_id();
  const int x = lane & 7, y = lane >> 3;
  for (int i_ = first_wave(); i_ < nact; i_ += gridDim.x * WPB) {
    const int* M = pmeta + (size_t)PSTRIDE * __builtin_amdgcn_readfirstlane(i_);
    const int B = M[0], o = M[1];
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    const int bb0 = M[2], bb1 = M[3], bb2 = M[4], bb3 = M[5], bb4 = M[6], bb5 = M[7], bb6 = M[8], bb7 = M[9];
    const int bx = 4 * ox, by = 4 * oy, bz = 4 * oz;
    const int X0 = bx + (x >> 1), Y0 = by + (y >> 1);
    const int X1 = X0 + ((x & 1) ? 1 : -1), Y1 = Y0 + ((y & 1) ? 1 : -1);
    const bool sx = (X1 < 0 || X1 > 7), sy = (Y1 < 0 || Y1 > 7);
    // brick of each of the 4 coarse columns, for z inside / outside the parent
    const int k00 = bb0, k10 = sx ? bb1 : bb0, k01 = sy ? bb2 : bb0, k11 = sx ? (sy ? bb3 : bb1) : (sy ? bb2 : bb0);
    const int z00 = bb4, z10 = sx ? bb5 : bb4, z01 = sy ? bb6 : bb4, z11 = sx ? (sy ? bb7 : bb5) : (sy ? bb6 : bb4);
    auto gather = [&](int brk, int X, int Y, int Z) -> T {
      return brk >= 0 ? zc[(size_t)brk * MAC_BC + (X & 7) + 8 * (Y & 7) + 64 * (Z & 7)] : T(0);
    };
    T c00[6], c10[6], c01[6], c11[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int Z = bz - 1 + k;
      const bool out = (Z < 0 || Z > 7);
      c00[k] = gather(out ? z00 : k00, X0, Y0, Z);
      c10[k] = gather(out ? z10 : k10, X1, Y0, Z);
      c01[k] = gather(out ? z01 : k01, X0, Y1, Z);
      c11[k] = gather(out ? z11 : k11, X1, Y1, Z);
    }
    const size_t base = (size_t)B * MAC_BC + lane;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const size_t c = base + 64 * z;
      if (!(codef[c] & CODE_DOM)) continue;
      const int k0 = (z >> 1) + 1, k1 = (z & 1) ? k0 + 1 : k0 - 1;
      const T a00 = T(0.75) * c00[k0] + T(0.25) * c00[k1];
      const T a10 = T(0.75) * c10[k0] + T(0.25) * c10[k1];
      const T a01 = T(0.75) * c01[k0] + T(0.25) * c01[k1];
      const T a11 = T(0.75) * c11[k0] + T(0.25) * c11[k1];
      zf[c] += T(0.75) * (T(0.75) * a00 + T(0.25) * a01) + T(0.25) * (T(0.75) * a10 + T(0.25) * a11);
    }
  }
}

// The prolongation of one fine brick, held by a wave: the four coarse
// columns a lane's cells read (as in k_prolong_col) and the trilinear value
// at the lane's own cells and at the cells just across the brick's faces.
// A face cell of a neighbour fine brick lies between the same coarse cells
// as the edge lane's own cell, with the two weights of the crossed direction
// swapped, so the edge lanes form the prolongated halo themselves, with the
// same operations in the same order as the neighbour forms its own cell.
// pmeta rows (PSTRIDE ints): brick id, octant, the 8 coarse bricks of
// k_prolong_col, then the 6 face-neighbour fine bricks (-x,+x,-y,+y,-z,+z).
template <typename T>
struct ProlongCols {
  T c00[6], c10[6], c01[6], c11[6];
  __device__ __forceinline__ void gather(const int* M, int lane, const T* __restrict__ zc) {
    const int x = lane & 7, y = lane >> 3;
    const int o = M[1];
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    const int bb0 = M[2], bb1 = M[3], bb2 = M[4], bb3 = M[5], bb4 = M[6], bb5 = M[7], bb6 = M[8], bb7 = M[9];
    const int bx = 4 * ox, by = 4 * oy, bz = 4 * oz;
    const int X0 = bx + (x >> 1), Y0 = by + (y >> 1);
    const int X1 = X0 + ((x & 1) ? 1 : -1), Y1 = Y0 + ((y & 1) ? 1 : -1);
    const bool sx = (X1 < 0 || X1 > 7), sy = (Y1 < 0 || Y1 > 7);
    const int k00 = bb0, k10 = sx ? bb1 : bb0, k01 = sy ? bb2 : bb0, k11 = sx ? (sy ? bb3 : bb1) : (sy ? bb2 : bb0);
    const int z00 = bb4, z10 = sx ? bb5 : bb4, z01 = sy ? bb6 : bb4, z11 = sx ? (sy ? bb7 : bb5) : (sy ? bb6 : bb4);
    auto g = [&](int brk, int X, int Y, int Z) -> T {
      return brk >= 0 ? zc[(size_t)brk * MAC_BC + (X & 7) + 8 * (Y & 7) + 64 * (Z & 7)] : T(0);
    };
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int Z = bz - 1 + k;
      const bool out = (Z < 0 || Z > 7);
      c00[k] = g(out ? z00 : k00, X0, Y0, Z);
      c10[k] = g(out ? z10 : k10, X1, Y0, Z);
      c01[k] = g(out ? z01 : k01, X0, Y1, Z);
      c11[k] = g(out ? z11 : k11, X1, Y1, Z);
    }
  }
  // z = -1 .. 8: coarse levels k0 (weight 3/4) and k1 (1/4)
  __device__ __forceinline__ static void zlev(int z, int& k0, int& k1) {
    if (z < 0) {
      k0 = 0;
      k1 = 1;
    } else if (z > 7) {
      k0 = 5;
      k1 = 4;
    } else {
      k0 = (z >> 1) + 1;
      k1 = (z & 1) ? k0 + 1 : k0 - 1;
    }
  }
  // face: 0 own cell, 1 across an x face, 2 across a y face
  __device__ __forceinline__ T at(int z, int face) const {
    int k0, k1;
    zlev(z, k0, k1);
    const T a00 = T(0.75) * c00[k0] + T(0.25) * c00[k1];
    const T a10 = T(0.75) * c10[k0] + T(0.25) * c10[k1];
    const T a01 = T(0.75) * c01[k0] + T(0.25) * c01[k1];
    const T a11 = T(0.75) * c11[k0] + T(0.25) * c11[k1];
    if (face == 1) return T(0.75) * (T(0.75) * a10 + T(0.25) * a11) + T(0.25) * (T(0.75) * a00 + T(0.25) * a01);
    if (face == 2) return T(0.75) * (T(0.75) * a01 + T(0.25) * a00) + T(0.25) * (T(0.75) * a11 + T(0.25) * a10);
    return T(0.75) * (T(0.75) * a00 + T(0.25) * a01) + T(0.25) * (T(0.75) * a10 + T(0.25) * a11);
  }
};

// The prolongated iterate zin = zfin + P zc of a fine brick: the wave's own
// column and the frozen face values (edge lanes), with the face planes of
// zfin and of the codes loaded lane-parallel from the neighbour bricks.
template <typename T>
struct ProlongedBrick {
  unsigned cd[8];
  T col[8], gx[8], gy[8], below, above;
  __device__ __forceinline__ void load(const int* M, int lane, const uint8_t* __restrict__ code, const T* __restrict__ zfin,
                                       const T* __restrict__ zc) {
    const int x = lane & 7, y = lane >> 3;
    const int B = M[0];
    const int* F = M + 10;  // face neighbours, the order of meta
    const size_t base = (size_t)B * MAC_BC + lane;
    ProlongCols<T> pc;
    pc.gather(M, lane, zc);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      cd[z] = code[base + 64 * z];
      const T v = zfin[base + 64 * z];
      col[z] = (cd[z] & CODE_DOM) ? v + pc.at(z, 0) : v;
    }
    const Halo<T> hz = load_halo<T>(F - 1, lane, [&](int nb, int cell) { return zfin[(size_t)nb * MAC_BC + cell]; });
    const Halo<unsigned> hc =
        load_halo<unsigned>(F - 1, lane, [&](int nb, int cell) { return (unsigned)code[(size_t)nb * MAC_BC + cell]; });
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const T vxm = __shfl(hz.xm, y + 8 * z), vxp = __shfl(hz.xp, y + 8 * z);
      const T vym = __shfl(hz.ym, x + 8 * z), vyp = __shfl(hz.yp, x + 8 * z);
      const unsigned cxm = __shfl(hc.xm, y + 8 * z), cxp = __shfl(hc.xp, y + 8 * z);
      const unsigned cym = __shfl(hc.ym, x + 8 * z), cyp = __shfl(hc.yp, x + 8 * z);
      const T vx = x == 0 ? vxm : vxp;
      const unsigned cx = x == 0 ? cxm : cxp;
      const T vy = y == 0 ? vym : vyp;
      const unsigned cy = y == 0 ? cym : cyp;
      gx[z] = (cx & CODE_DOM) ? vx + pc.at(z, 1) : vx;
      gy[z] = (cy & CODE_DOM) ? vy + pc.at(z, 2) : vy;
    }
    below = (hc.below & CODE_DOM) ? hz.below + pc.at(-1, 0) : hz.below;
    above = (hc.above & CODE_DOM) ? hz.above + pc.at(8, 0) : hz.above;
  }
};

// The coarse columns of ProlongCols that wave w of a k_prolong_block block reads:
// the 3 coarse z-levels bz-1+w .. bz+1+w (local levels 0..2) cover its planes
// 2w, 2w+1 and the halo plane below (w = 0) or above (w = 3).  at() takes
// its two levels as constants, so the columns stay in registers, and forms
// the value with the expressions of ProlongCols::at (identical bits).
template <typename T>
struct SlabCols {
  T c00[3], c10[3], c01[3], c11[3];
  __device__ __forceinline__ void gather(const int* M, int lane, int wv, const T* __restrict__ zc) {
    const int x = lane & 7, y = lane >> 3;
    const int o = M[1];
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    const int bb0 = M[2], bb1 = M[3], bb2 = M[4], bb3 = M[5], bb4 = M[6], bb5 = M[7], bb6 = M[8], bb7 = M[9];
    const int bx = 4 * ox, by = 4 * oy, bz = 4 * oz;
    const int X0 = bx + (x >> 1), Y0 = by + (y >> 1);
    const int X1 = X0 + ((x & 1) ? 1 : -1), Y1 = Y0 + ((y & 1) ? 1 : -1);
    const bool sx = (X1 < 0 || X1 > 7), sy = (Y1 < 0 || Y1 > 7);
    const int k00 = bb0, k10 = sx ? bb1 : bb0, k01 = sy ? bb2 : bb0, k11 = sx ? (sy ? bb3 : bb1) : (sy ? bb2 : bb0);
    const int z00 = bb4, z10 = sx ? bb5 : bb4, z01 = sy ? bb6 : bb4, z11 = sx ? (sy ? bb7 : bb5) : (sy ? bb6 : bb4);
    auto g = [&](int brk, int X, int Y, int Z) -> T {
      return brk >= 0 ? zc[(size_t)brk * MAC_BC + (X & 7) + 8 * (Y & 7) + 64 * (Z & 7)] : T(0);
    };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int Z = bz - 1 + wv + k;
      const bool out = (Z < 0 || Z > 7);
      c00[k] = g(out ? z00 : k00, X0, Y0, Z);
      c10[k] = g(out ? z10 : k10, X1, Y0, Z);
      c01[k] = g(out ? z01 : k01, X0, Y1, Z);
      c11[k] = g(out ? z11 : k11, X1, Y1, Z);
    }
  }
  template <int K0, int K1>
  __device__ __forceinline__ T at(int face) const {
    const T a00 = T(0.75) * c00[K0] + T(0.25) * c00[K1];
    const T a10 = T(0.75) * c10[K0] + T(0.25) * c10[K1];
    const T a01 = T(0.75) * c01[K0] + T(0.25) * c01[K1];
    const T a11 = T(0.75) * c11[K0] + T(0.25) * c11[K1];
    if (face == 1) return T(0.75) * (T(0.75) * a10 + T(0.25) * a11) + T(0.25) * (T(0.75) * a00 + T(0.25) * a01);
    if (face == 2) return T(0.75) * (T(0.75) * a01 + T(0.25) * a00) + T(0.25) * (T(0.75) * a11 + T(0.25) * a10);
    return T(0.75) * (T(0.75) * a00 + T(0.25) * a01) + T(0.25) * (T(0.75) * a10 + T(0.25) * a11);
  }
  // plane 2w + j (j = 0, 1)
  __device__ __forceinline__ T at_plane(int j, int face) const { return j ? at<1, 2>(face) : at<1, 0>(face); }
};

// Prolongation fused with the first post-smoothing block step on a coarse
// level: zout = zin + damp (block correction of zin), zin = zfin + P zc,
// equal to k_prolong_col followed by k_jacobi_block(zin) without writing and
// re-reading zin.  One block per active fine brick, wave w on z-planes 2w,
// 2w+1; the prolongated values (own cells and frozen halo) are formed as in
// ProlongedBrick.
template <typename T>
__global__ void __launch_bounds__(KTHREADS) k_prolong_block(int nact, const int* __restrict__ pmeta,
                                                                 const uint8_t* __restrict__ code, const T* __restrict__ r,
                                                                 const T* __restrict__ zc, const T* __restrict__ zfin,
                                                                 T* __restrict__ zout, const T* __restrict__ cf, size_t ncell,
                                                                 Ih<T> ih, BlockW w, const Scal* s) {
  const int i = xcd_block();
  if (converged(pmeta, PSTRIDE, nact, i, s)) return;
  __shared__ SlabLDS<T> L;
  const int lane = lane_id(), wv = wave_in_block();
  const int x = lane & 7, y = lane >> 3;
  const int* M = pmeta + (size_t)PSTRIDE * i;
  const int* F = M + 10;  // face neighbours, the order of meta
  const size_t base = (size_t)M[0] * MAC_BC + lane + 128 * wv;
  SlabCols<T> pc;
  pc.gather(M, lane, wv, zc);
  unsigned cd[2];
  T col[2], rv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    cd[j] = code[base + 64 * j];
    rv[j] = r[base + 64 * j];
    const T v = zfin[base + 64 * j];
    col[j] = (cd[j] & CODE_DOM) ? v + pc.at_plane(j, 0) : v;
  }
  auto ldz = [&](int nb, int cell) { return nb >= 0 ? zfin[(size_t)nb * MAC_BC + cell] : T(0); };
  auto ldc = [&](int nb, int cell) { return nb >= 0 ? (unsigned)code[(size_t)nb * MAC_BC + cell] : 0u; };
  const T zxm = ldz(F[0], 7 + 8 * x + 64 * y), zxp = ldz(F[1], 8 * x + 64 * y);
  const T zym = ldz(F[2], x + 56 + 64 * y), zyp = ldz(F[3], x + 64 * y);
  const unsigned cxm = ldc(F[0], 7 + 8 * x + 64 * y), cxp = ldc(F[1], 8 * x + 64 * y);
  const unsigned cym = ldc(F[2], x + 56 + 64 * y), cyp = ldc(F[3], x + 64 * y);
  T gx[2], gy[2], below = T(0), above = T(0);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int z = 2 * wv + j;
    const T vxm = __shfl(zxm, y + 8 * z), vxp = __shfl(zxp, y + 8 * z);
    const T vym = __shfl(zym, x + 8 * z), vyp = __shfl(zyp, x + 8 * z);
    const unsigned gcxm = __shfl(cxm, y + 8 * z), gcxp = __shfl(cxp, y + 8 * z);
    const unsigned gcym = __shfl(cym, x + 8 * z), gcyp = __shfl(cyp, x + 8 * z);
    const T vx = x == 0 ? vxm : vxp;
    const unsigned cx = x == 0 ? gcxm : gcxp;
    const T vy = y == 0 ? vym : vyp;
    const unsigned cy = y == 0 ? gcym : gcyp;
    gx[j] = (cx & CODE_DOM) ? vx + pc.at_plane(j, 1) : vx;
    gy[j] = (cy & CODE_DOM) ? vy + pc.at_plane(j, 2) : vy;
  }
  if (wv == 0) {
    const T zb = ldz(F[4], lane + 448);
    below = (ldc(F[4], lane + 448) & CODE_DOM) ? zb + pc.template at<0, 1>(0) : zb;
  }
  if (wv == WPB - 1) {
    const T za = ldz(F[5], lane);
    above = (ldc(F[5], lane) & CODE_DOM) ? za + pc.template at<2, 1>(0) : za;
  }
  const T col0[2] = {col[0], col[1]};
  const Coef<T> K[2] = {coef_at<T>(cd[0], cf, ncell, base, ih), coef_at<T>(cd[1], cf, ncell, base + 64, ih)};
  T idg[2];
  inv_diag_slab<T>(cd, K, idg);
  slab_sweeps<T>(wv, lane, cd, K, rv, idg, col, gx, gy, below, above, w, L);
  const T dmp = (T)w.damp;
#pragma unroll
  for (int j = 0; j < 2; ++j) zout[base + 64 * j] = col0[j] + dmp * (col[j] - col0[j]);
}

// Prolongation into level 0 fused with the first post-smoothing Jacobi sweep:
// zout = zin + omega D^-1 (r - L zin), zin = zfin + P zc (k_prolong_col then
// k_jacobi, without writing and re-reading zin).  want_dot: partial r . zout.
template <typename TZ>
#ifndef MAC_PJ_WAVES
#define MAC_PJ_WAVES 4
#endif
__global__ void __launch_bounds__(KTHREADS) __attribute__((amdgpu_waves_per_eu(MAC_PJ_WAVES ? MAC_PJ_WAVES : 1)))
k_prolong_jacobi(int nact, const int* __restrict__ pmeta,
                                                             const uint8_t* __restrict__ code, const double* __restrict__ r,
                                                             const TZ* __restrict__ zc, const TZ* __restrict__ zfin,
                                                             TZ* __restrict__ zout, TZ ihx2, TZ ihy2, TZ ihz2, TZ omega,
                                                             int want_dot, double* __restrict__ partials, const Scal* s) {
  if (converged(pmeta, PSTRIDE, nact, first_wave(), s)) return;
  __shared__ double sh[16];
  const int lane = lane_id();
  const int x = lane & 7, y = lane >> 3;
  double acc = 0.0;
  for (int i_ = first_wave(); i_ < nact; i_ += gridDim.x * WPB) {
    const int* M = pmeta + (size_t)PSTRIDE * __builtin_amdgcn_readfirstlane(i_);
    const size_t base = (size_t)M[0] * MAC_BC + lane;
    ProlongedBrick<TZ> pb;
    pb.load(M, lane, code, zfin, zc);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      const TZ v = pb.col[z];
      TZ xm = __shfl_up(v, 1), xp = __shfl_down(v, 1), ym = __shfl_up(v, 8), yp = __shfl_down(v, 8);
      if (x == 0) xm = pb.gx[z];
      if (x == 7) xp = pb.gx[z];
      if (y == 0) ym = pb.gy[z];
      if (y == 7) yp = pb.gy[z];
      const TZ zm = (z > 0) ? pb.col[z > 0 ? z - 1 : 0] : pb.below;
      const TZ zp = (z < 7) ? pb.col[z < 7 ? z + 1 : 7] : pb.above;
      TZ out = TZ(0);
      if (pb.cd[z] & CODE_DOM) {
        TZ dg;
        const TZ lz = lap_cell<TZ>(pb.cd[z], v, xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
        const double rv = r[base + 64 * z];
        out = dg > TZ(0) ? v + div_diag<TZ>(omega * ((TZ)rv - lz), dg) : TZ(0);
        if (want_dot) acc += rv * (double)out;
      }
      zout[base + 64 * z] = out;
    }
  }
  if (want_dot) write_partial(acc, partials, sh, false);
}

// k_prolong_jacobi in the slab form: one 256-thread block per active fine
// brick, wave w on z-planes 2w, 2w+1 (a quarter of the column form's
// registers, four times the waves).  The prolongated values of the wave's
// planes and of the cells across the x / y faces are formed as in
// k_prolong_block (the face planes of zfin and of the codes loaded only on
// the 16 lanes that hold the wave's planes); the planes 2w-1 and 2w+2 come
// from the neighbouring waves through LDS (one barrier).  Same expressions
// as k_prolong_jacobi, so zout is bitwise the same; want_dot: one partial
// r . zout per block.
template <typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_prolong_jacobi_slab(int nact, const int* __restrict__ pmeta,
                                                                  const uint8_t* __restrict__ code, const double* __restrict__ r,
                                                                  const TZ* __restrict__ zc, const TZ* __restrict__ zfin,
                                                                  TZ* __restrict__ zout, TZ ihx2, TZ ihy2, TZ ihz2, TZ omega,
                                                                  int want_dot, double* __restrict__ partials, const Scal* s) {
  const int i = xcd_block();
  if (converged(pmeta, PSTRIDE, nact, i, s)) return;
  __shared__ TZ pl[WPB][2][64];
  __shared__ double sh[16];
  const int lane = lane_id(), wv = wave_in_block();
  const int x = lane & 7, y = lane >> 3;
  const int* M = pmeta + (size_t)PSTRIDE * i;
  const int* F = M + 10;  // face neighbours, the order of meta
  const size_t base = (size_t)M[0] * MAC_BC + lane + 128 * wv;
  SlabCols<TZ> pc;
  pc.gather(M, lane, wv, zc);
  unsigned cd[2];
  TZ col[2];
  double rv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    cd[j] = code[base + 64 * j];
    rv[j] = r[base + 64 * j];
    const TZ v = zfin[base + 64 * j];
    col[j] = (cd[j] & CODE_DOM) ? v + pc.at_plane(j, 0) : v;
  }
  // face planes: lane = y' + 8 z' (x faces) or x' + 8 z' (y faces); the
  // wave's planes z' = 2w, 2w+1 are lanes 16w .. 16w+15
  const bool mine = (lane >> 4) == wv;
  auto ldz = [&](int nb, int cell) { return (mine && nb >= 0) ? zfin[(size_t)nb * MAC_BC + cell] : TZ(0); };
  auto ldc = [&](int nb, int cell) { return (mine && nb >= 0) ? (unsigned)code[(size_t)nb * MAC_BC + cell] : 0u; };
  const TZ zxm = ldz(F[0], 7 + 8 * x + 64 * y), zxp = ldz(F[1], 8 * x + 64 * y);
  const TZ zym = ldz(F[2], x + 56 + 64 * y), zyp = ldz(F[3], x + 64 * y);
  const unsigned cxm = ldc(F[0], 7 + 8 * x + 64 * y), cxp = ldc(F[1], 8 * x + 64 * y);
  const unsigned cym = ldc(F[2], x + 56 + 64 * y), cyp = ldc(F[3], x + 64 * y);
  TZ gx[2], gy[2], below = TZ(0), above = TZ(0);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int z = 2 * wv + j;
    const TZ vxm = __shfl(zxm, y + 8 * z), vxp = __shfl(zxp, y + 8 * z);
    const TZ vym = __shfl(zym, x + 8 * z), vyp = __shfl(zyp, x + 8 * z);
    const unsigned gcxm = __shfl(cxm, y + 8 * z), gcxp = __shfl(cxp, y + 8 * z);
    const unsigned gcym = __shfl(cym, x + 8 * z), gcyp = __shfl(cyp, x + 8 * z);
    const TZ vx = x == 0 ? vxm : vxp;
    const unsigned cx = x == 0 ? gcxm : gcxp;
    const TZ vy = y == 0 ? vym : vyp;
    const unsigned cy = y == 0 ? gcym : gcyp;
    gx[j] = (cx & CODE_DOM) ? vx + pc.at_plane(j, 1) : vx;
    gy[j] = (cy & CODE_DOM) ? vy + pc.at_plane(j, 2) : vy;
  }
  if (wv == 0 && F[4] >= 0) {
    const TZ zb = zfin[(size_t)F[4] * MAC_BC + lane + 448];
    below = (code[(size_t)F[4] * MAC_BC + lane + 448] & CODE_DOM) ? zb + pc.template at<0, 1>(0) : zb;
  }
  if (wv == WPB - 1 && F[5] >= 0) {
    const TZ za = zfin[(size_t)F[5] * MAC_BC + lane];
    above = (code[(size_t)F[5] * MAC_BC + lane] & CODE_DOM) ? za + pc.template at<2, 1>(0) : za;
  }
  pl[wv][0][lane] = col[0];
  pl[wv][1][lane] = col[1];
  __syncthreads();
  const TZ zlo = wv > 0 ? pl[wv - 1][1][lane] : below;
  const TZ zhi = wv < WPB - 1 ? pl[wv + 1][0][lane] : above;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const TZ v = col[j];
    TZ xm = __shfl_up(v, 1), xp = __shfl_down(v, 1), ym = __shfl_up(v, 8), yp = __shfl_down(v, 8);
    if (x == 0) xm = gx[j];
    if (x == 7) xp = gx[j];
    if (y == 0) ym = gy[j];
    if (y == 7) yp = gy[j];
    const TZ zm = j ? col[0] : zlo;
    const TZ zp = j ? zhi : col[1];
    TZ out = TZ(0);
    if (cd[j] & CODE_DOM) {
      TZ dg;
      const TZ lz = lap_cell<TZ>(cd[j], v, xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
      out = dg > TZ(0) ? v + div_diag<TZ>(omega * ((TZ)rv[j] - lz), dg) : TZ(0);
      if (want_dot) acc += rv[j] * (double)out;
    }
    zout[base + 64 * j] = out;
  }
  if (want_dot) write_partial(acc, partials, sh, false);
}

// res = r - L z.
template <typename TR, typename TZ>
__global__ void __launch_bounds__(KTHREADS) k_residual(int nact, const int* __restrict__ meta,
                                                       const uint8_t* __restrict__ code, const TR* __restrict__ r,
                                                       const TZ* __restrict__ zv, TZ* __restrict__ res, TZ ihx2, TZ ihy2,
                                                       TZ ihz2, const Scal* s) {
  if (converged(meta, MSTRIDE, nact, first_wave(), s)) return;
  const int lane = lane_id();
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    const size_t base = (size_t)b * MAC_BC + lane;
    unsigned cd[8];
    TZ col[8];
    TR rr[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      cd[z] = code[base + 64 * z];
      col[z] = zv[base + 64 * z];
      rr[z] = r[base + 64 * z];
    }
    col_stencil<TZ>(M, lane, col, [&](int nb, int cell) { return zv[(size_t)nb * MAC_BC + cell]; },
                    [&](int z, TZ xm, TZ xp, TZ ym, TZ yp, TZ zm, TZ zp) {
                      TZ out = TZ(0);
                      if (cd[z] & CODE_DOM) {
                        TZ dg;
                        out = (TZ)rr[z] - lap_cell<TZ>(cd[z], col[z], xm, xp, ym, yp, zm, zp, ihx2, ihy2, ihz2, dg);
                      }
                      res[base + 64 * z] = out;
                    });
  }
}

// Exact solve on the coarsest level (a single brick): z = Linv r with the
// dense inverse precomputed on the host (symmetric, so column j of Linv is
// read coalesced as row j).  cellof[i] = cell of unknown i, rowof[cell] = i
// or -1.  want_dot: partial r . z.
template <typename TR, typename TZ>
__global__ void k_coarse_solve(int nc, const int* __restrict__ rowof, const int* __restrict__ cellof,
                               const TZ* __restrict__ cinv, const TR* __restrict__ rc, TZ* __restrict__ zc, int want_dot,
                               double* __restrict__ partials, const Scal* s) {
  if (s->done) return;
  __shared__ TZ rl[MAC_BC];
  __shared__ double sh[16];
  for (int i = threadIdx.x; i < nc; i += blockDim.x) rl[i] = (TZ)rc[cellof[i]];
  __syncthreads();
  int t = blockIdx.x * blockDim.x + threadIdx.x;  // cell of the coarsest brick
  double acc = 0.0;
  if (t < MAC_BC) {
    int row = rowof[t];
    TZ out = TZ(0);
    if (row >= 0) {
      for (int j = 0; j < nc; ++j) out += cinv[(size_t)j * nc + row] * rl[j];
      if (want_dot) acc = (double)rc[t] * (double)out;
    }
    zc[t] = out;
  }
  if (want_dot) write_partial(acc, partials, sh, false);
}

// The last two levels of the V-cycle in one launch of one 1024-thread block.
// The level above the top brick holds at most its 8 children, a 16^3 box of
// cells; the box lives in LDS (z, its sweep partner, the restriction rows),
// r, the codes and the operator's coefficients in registers (4 cells per
// thread), so every sweep is one
// pass over LDS between two barriers instead of a launch with its chain of
// dependent global loads.  The steps and their expressions are those of
// k_jacobi2 / k_jacobi (pre-sweeps), k_residual, k_restrict_sep,
// k_coarse_solve, k_prolong_col and k_jacobi (post-sweeps); cells outside the
// box and in missing bricks read as 0, as missing bricks do there.
struct TailArgs {
  int child[8];           // level-t bricks of the box, octant order (-1 = none)
  int nu;                 // sweeps before and after the correction
  int nc;                 // unknowns of the top brick
  double ih2[3];          // level t's squared inverse spacings (CODE_STD cells)
  double om_pre[8], om_post[8];
};
#define TAIL_THREADS 1024
template <typename TZ>
__global__ void __launch_bounds__(TAIL_THREADS) k_vcycle_tail(TailArgs a, const uint8_t* __restrict__ code_t,
                                                              const TZ* __restrict__ cf_t, size_t nct,
                                                              const TZ* __restrict__ r_t, TZ* __restrict__ z_t,
                                                              const uint8_t* __restrict__ code_c, const int* __restrict__ rowof,
                                                              const int* __restrict__ cellof, const TZ* __restrict__ cinv,
                                                              const Scal* s) {
  if (s->done) return;
  __shared__ TZ buf[2][4096];
  __shared__ TZ t1[2048];
  __shared__ TZ rc[MAC_BC], zc[MAC_BC];
  const int tid = threadIdx.x;
  long long g[4];
  unsigned cd[4];
  TZ rv[4];
  Coef<TZ> K[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = tid + TAIL_THREADS * k;
    const int X = i & 15, Y = (i >> 4) & 15, Z = i >> 8;
    const int b = a.child[(X >> 3) + 2 * (Y >> 3) + 4 * (Z >> 3)];
    g[k] = b >= 0 ? (long long)b * MAC_BC + (X & 7) + 8 * (Y & 7) + 64 * (Z & 7) : -1;
    cd[k] = g[k] >= 0 ? code_t[g[k]] : 0u;
    rv[k] = g[k] >= 0 ? r_t[g[k]] : TZ(0);
    if (cd[k] & CODE_DOM) K[k] = coef_at<TZ>(cd[k], cf_t, nct, (size_t)g[k], Ih<TZ>{(TZ)a.ih2[0], (TZ)a.ih2[1], (TZ)a.ih2[2]});
  }
  auto at = [](const TZ* v, int X, int Y, int Z) -> TZ {
    return ((unsigned)X < 16u && (unsigned)Y < 16u && (unsigned)Z < 16u) ? v[X + 16 * Y + 256 * Z] : TZ(0);
  };
  // one damped-Jacobi sweep zin -> zout (zin == nullptr: from zero)
  auto sweep = [&](const TZ* zin, TZ* zout, TZ om) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = tid + TAIL_THREADS * k;
      const int X = i & 15, Y = (i >> 4) & 15, Z = i >> 8;
      TZ out = TZ(0);
      if (cd[k] & CODE_DOM) {
        const TZ dg = K[k].dg;
        if (zin) {
          const TZ lz = lap_coef<TZ>(K[k], zin[i], at(zin, X - 1, Y, Z), at(zin, X + 1, Y, Z), at(zin, X, Y - 1, Z),
                                     at(zin, X, Y + 1, Z), at(zin, X, Y, Z - 1), at(zin, X, Y, Z + 1));
          out = dg > TZ(0) ? zin[i] + om * (rv[k] - lz) / dg : TZ(0);
        } else {
          out = dg > TZ(0) ? om * rv[k] / dg : TZ(0);
        }
      }
      zout[i] = out;
    }
    __syncthreads();
  };
  int cur = 0;
  sweep(nullptr, buf[0], (TZ)a.om_pre[0]);
  for (int k = 1; k < a.nu; ++k) {
    sweep(buf[cur], buf[cur ^ 1], (TZ)a.om_pre[k]);
    cur ^= 1;
  }
  TZ* z = buf[cur];
  TZ* res = buf[cur ^ 1];
  // res = r - L z
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = tid + TAIL_THREADS * k;
    const int X = i & 15, Y = (i >> 4) & 15, Z = i >> 8;
    TZ out = TZ(0);
    if (cd[k] & CODE_DOM)
      out = rv[k] - lap_coef<TZ>(K[k], z[i], at(z, X - 1, Y, Z), at(z, X + 1, Y, Z), at(z, X, Y - 1, Z), at(z, X, Y + 1, Z),
                                 at(z, X, Y, Z - 1), at(z, X, Y, Z + 1));
    res[i] = out;
  }
  __syncthreads();
  // restriction with the 1-D weights (1/4, 3/4, 3/4, 1/4): x, then y, then z
  for (int i = tid; i < 2048; i += TAIL_THREADS) {  // t1[fz][fy][X]
    const int X = i & 7, fy = (i >> 3) & 15, fz = i >> 7;
    t1[i] = TZ(0.25) * at(res, 2 * X - 1, fy, fz) + TZ(0.75) * at(res, 2 * X, fy, fz) + TZ(0.75) * at(res, 2 * X + 1, fy, fz) +
            TZ(0.25) * at(res, 2 * X + 2, fy, fz);
  }
  __syncthreads();
  TZ* t2 = res;  // [fz][Y][X]; res is consumed
  if (tid < 1024) {
    const int X = tid & 7, Y = (tid >> 3) & 7, fz = tid >> 6;
    auto r1 = [&](int fy) -> TZ { return (unsigned)fy < 16u ? t1[(fz * 16 + fy) * 8 + X] : TZ(0); };
    t2[tid] = TZ(0.25) * r1(2 * Y - 1) + TZ(0.75) * r1(2 * Y) + TZ(0.75) * r1(2 * Y + 1) + TZ(0.25) * r1(2 * Y + 2);
  }
  __syncthreads();
  if (tid < MAC_BC) {
    const int X = tid & 7, Y = (tid >> 3) & 7, Zc = tid >> 6;
    auto r2 = [&](int fz) -> TZ { return (unsigned)fz < 16u ? t2[fz * 64 + Y * 8 + X] : TZ(0); };
    TZ acc = TZ(0);
    if (code_c[tid] & CODE_DOM)
      acc = (TZ(0.25) * r2(2 * Zc - 1) + TZ(0.75) * r2(2 * Zc) + TZ(0.75) * r2(2 * Zc + 1) + TZ(0.25) * r2(2 * Zc + 2)) * TZ(0.125);
    rc[tid] = acc;
  }
  __syncthreads();
  // exact solve on the top brick
  if (tid < a.nc) t1[tid] = rc[cellof[tid]];
  __syncthreads();
  if (tid < MAC_BC) {
    const int row = rowof[tid];
    TZ out = TZ(0);
    if (row >= 0)
      for (int j = 0; j < a.nc; ++j) out += cinv[(size_t)j * a.nc + row] * t1[j];
    zc[tid] = out;
  }
  __syncthreads();
  // z += P zc (trilinear, 3/4 on the parent side)
  auto czv = [&](int X, int Y, int Z) -> TZ {
    return ((unsigned)X < 8u && (unsigned)Y < 8u && (unsigned)Z < 8u) ? zc[X + 8 * Y + 64 * Z] : TZ(0);
  };
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = tid + TAIL_THREADS * k;
    if (!(cd[k] & CODE_DOM)) continue;
    const int X = i & 15, Y = (i >> 4) & 15, Z = i >> 8;
    const int X0 = X >> 1, Y0 = Y >> 1, Z0 = Z >> 1;
    const int X1 = X0 + ((X & 1) ? 1 : -1), Y1 = Y0 + ((Y & 1) ? 1 : -1), Z1 = Z0 + ((Z & 1) ? 1 : -1);
    const TZ a00 = TZ(0.75) * czv(X0, Y0, Z0) + TZ(0.25) * czv(X0, Y0, Z1);
    const TZ a10 = TZ(0.75) * czv(X1, Y0, Z0) + TZ(0.25) * czv(X1, Y0, Z1);
    const TZ a01 = TZ(0.75) * czv(X0, Y1, Z0) + TZ(0.25) * czv(X0, Y1, Z1);
    const TZ a11 = TZ(0.75) * czv(X1, Y1, Z0) + TZ(0.25) * czv(X1, Y1, Z1);
    z[i] += TZ(0.75) * (TZ(0.75) * a00 + TZ(0.25) * a01) + TZ(0.25) * (TZ(0.75) * a10 + TZ(0.25) * a11);
  }
  __syncthreads();
  for (int k = 0; k < a.nu; ++k) {
    sweep(buf[cur], buf[cur ^ 1], (TZ)a.om_post[k]);
    cur ^= 1;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (g[k] >= 0) z_t[g[k]] = buf[cur][tid + TAIL_THREADS * k];
}

// u_i -= p(c)/h_i - p(c-e_i)/h_i on every face of an existing cell whose
// cell or whose -e_i neighbour is in the domain, unless the face is const
// (Algorithms_Solve_Pressure.h:198-207; Policy_Pressure_If_Correction.h:33-49;
// Policy_Gradiant.h:32-37; Data_Grid_Speed.tpp:20-26).  Same operation order.
// One wave per listed brick in the column layout: the -x and -y neighbours
// of a cell are cross-lane shuffles (the brick's -x / -y face planes of p and
// of the codes loaded lane-parallel from the neighbour bricks), the -z
// neighbour the column's previous cell (or the -z brick's top plane).
__global__ void __launch_bounds__(KTHREADS) k_gradient(int n, const int* __restrict__ list, const int* __restrict__ nbr,
                                                       const uint8_t* __restrict__ flags, const uint8_t* __restrict__ code,
                                                       const double* __restrict__ p, double* __restrict__ u,
                                                       double* __restrict__ v, double* __restrict__ w, double ihx, double ihy,
                                                       double ihz) {
#pragma clang fp contract(off)
  const int lane = lane_id();
  const int x = lane & 7, y = lane >> 3;
  for (int i_ = first_wave(); i_ < n; i_ += gridDim.x * WPB) {
    const int b = list[__builtin_amdgcn_readfirstlane(i_)];
    const int* N = nbr + (size_t)b * 27;
    const int bxm = N[12], bym = N[10], bzm = N[4];
    const size_t base = (size_t)b * MAC_BC + lane;
    unsigned fl[8], cd[8];
    double pc[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
      fl[z] = flags[base + 64 * z];
      cd[z] = code[base + 64 * z];
      pc[z] = p[base + 64 * z];
    }
    // face planes of the low-side neighbour bricks (x faces indexed (y, z), y faces (x, z))
    const double hpx = bxm >= 0 ? p[(size_t)bxm * MAC_BC + 7 + 8 * x + 64 * y] : 0.0;
    const unsigned hcx = bxm >= 0 ? code[(size_t)bxm * MAC_BC + 7 + 8 * x + 64 * y] : 0u;
    const double hpy = bym >= 0 ? p[(size_t)bym * MAC_BC + x + 56 + 64 * y] : 0.0;
    const unsigned hcy = bym >= 0 ? code[(size_t)bym * MAC_BC + x + 56 + 64 * y] : 0u;
    const double hpz = bzm >= 0 ? p[(size_t)bzm * MAC_BC + lane + 448] : 0.0;
    const unsigned hcz = bzm >= 0 ? code[(size_t)bzm * MAC_BC + lane + 448] : 0u;
    // per half column: the updates each cell takes and their gradients,
    // then the half's loads of u, v, w in flight at once, then the stores
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double gx[4], gy[4], gz[4];
      unsigned todo[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int z = 4 * h + k;
        double pxm = __shfl_up(pc[z], 1), pym = __shfl_up(pc[z], 8);
        unsigned cxm = __shfl_up(cd[z], 1), cym = __shfl_up(cd[z], 8);
        const double gpx = __shfl(hpx, y + 8 * z), gpy = __shfl(hpy, x + 8 * z);
        const unsigned gcx = __shfl(hcx, y + 8 * z), gcy = __shfl(hcy, x + 8 * z);
        if (x == 0) {
          pxm = gpx;
          cxm = gcx;
        }
        if (y == 0) {
          pym = gpy;
          cym = gcy;
        }
        const double pzm = z > 0 ? pc[z > 0 ? z - 1 : 0] : hpz;
        const unsigned czm = z > 0 ? cd[z > 0 ? z - 1 : 0] : hcz;
        const unsigned f = fl[z];
        const bool dom = cd[z] & CODE_DOM;
        unsigned t = 0;
        if (f & 0x80u) {
          if ((dom || (cxm & CODE_DOM)) && !(f & 0x08u)) t |= 1u;
          if ((dom || (cym & CODE_DOM)) && !(f & 0x10u)) t |= 2u;
          if ((dom || (czm & CODE_DOM)) && !(f & 0x20u)) t |= 4u;
        }
        todo[k] = t;
        double g = pc[z] * ihx;
        g -= pxm * ihx;
        gx[k] = g;
        g = pc[z] * ihy;
        g -= pym * ihy;
        gy[k] = g;
        g = pc[z] * ihz;
        g -= pzm * ihz;
        gz[k] = g;
      }
      double uu[4], vv[4], ww[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t c = base + 64 * (4 * h + k);
        uu[k] = (todo[k] & 1u) ? u[c] : 0.0;
        vv[k] = (todo[k] & 2u) ? v[c] : 0.0;
        ww[k] = (todo[k] & 4u) ? w[c] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const size_t c = base + 64 * (4 * h + k);
        if (todo[k] & 1u) u[c] = uu[k] - gx[k];
        if (todo[k] & 2u) v[c] = vv[k] - gy[k];
        if (todo[k] & 4u) w[c] = ww[k] - gz[k];
      }
    }
  }
}

// y = A x (reference sign) for the parity entry point.
__global__ void k_apply_A(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                          const double* __restrict__ xv, double* __restrict__ y, double ihx2, double ihy2, double ihz2) {
  int b = act[blockIdx.x];
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, yy = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned cd = code[c];
    if (cd & CODE_DOM) {
      double dg;
      double lx = lap_cell<double>(cd, xv[c], ld_rel(xv, nbr, b, x - 1, yy, z), ld_rel(xv, nbr, b, x + 1, yy, z),
                                   ld_rel(xv, nbr, b, x, yy - 1, z), ld_rel(xv, nbr, b, x, yy + 1, z),
                                   ld_rel(xv, nbr, b, x, yy, z - 1), ld_rel(xv, nbr, b, x, yy, z + 1), ihx2, ihy2, ihz2, dg);
      y[c] = -lx;
    }
  }
}

// Zero the out-of-domain entries of a host-supplied vector (so that the
// operator sees the zero-outside-domain convention).
// ------------------------------------------------ viscous air pressure ----
// Policy_Dirichlet_Boundary.h with Algorithms_Solve_Pressure_Dirichlet.h.
// "Air" is Air or Solid (GetIsAir); a Dirichlet boundary cell is an existing
// air cell with an existing non-air face neighbour.  Its pressure
//   p_air = - sum over those neighbours (i, s) of 2 nu (u_i(c+e_i) - u_i(c)) / h_i * s
// (same operation order; the velocity is the real one, and a face of a cell
// that does not exist reads what its brick holds, 0 for a missing brick).
// pD gets p_air on Dirichlet boundary cells, 0 elsewhere.
__device__ __forceinline__ bool cell_valid(const uint8_t* flags, const int* nbr27, int b, int x, int y, int z, size_t& idx) {
  int cell;
  const int nb = locate(nbr27, b, x, y, z, cell);
  if (nb < 0) return false;
  idx = (size_t)nb * MAC_BC + cell;
  return (flags[idx] & 0x80u) != 0;
}
__device__ __forceinline__ bool mat_air(unsigned f) { return (f & 7u) == 0u || (f & 7u) == 2u; }

__global__ void k_air_pressure(const int* __restrict__ nbr27, const uint8_t* __restrict__ flags, const double* __restrict__ u,
                               const double* __restrict__ v, const double* __restrict__ w, double ihx, double ihy, double ihz,
                               double nu, double* __restrict__ pD) {
#pragma clang fp contract(off)
  const int b = blockIdx.x;
  const double* U[3] = {u, v, w};
  const double ih[3] = {ihx, ihy, ihz};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t c = (size_t)b * MAC_BC + t;
    const unsigned f = flags[c];
    double ret = 0.0;
    if ((f & 0x80u) && mat_air(f)) {
      const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
      for (int i = 0; i < 3; ++i)
        for (int s = -1; s <= 1; s += 2) {
          size_t n;
          if (!cell_valid(flags, nbr27, b, x + (i == 0) * s, y + (i == 1) * s, z + (i == 2) * s, n) || mat_air(flags[n]))
            continue;
          const double v1 = U[i][c];
          const double v2 = ld_rel(U[i], nbr27, b, x + (i == 0), y + (i == 1), z + (i == 2));
          ret -= 2.0 * nu * (v2 - v1) * ih[i] * (double)s;
        }
    }
    pD[c] = ret;
  }
}

// f -= the Dirichlet term of b (b -= sum of p_air h_i^-2 over the Dirichlet
// boundary neighbours, f = -b), partial ||f||^2.
__global__ void __launch_bounds__(KTHREADS) k_dirichlet_rhs(int nact, const int* __restrict__ meta, const int* __restrict__ nbr27,
                                                            const uint8_t* __restrict__ flags, const uint8_t* __restrict__ code,
                                                            const double* __restrict__ pD, double ihx, double ihy, double ihz,
                                                            double* __restrict__ f, double* __restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double sh[16];
  const int lane = lane_id();
  const double ih[3] = {ihx, ihy, ihz};
  double acc = 0.0;
  FOR_EACH_BRICK(meta, nact, M) {
    const int b = M[0];
    for (int z = 0; z < 8; ++z) {
      const int t = lane + 64 * z;
      const size_t c = (size_t)b * MAC_BC + t;
      if (!(code[c] & CODE_DOM)) continue;
      const int x = t & 7, y = (t >> 3) & 7;
      double ret = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int s = -1; s <= 1; s += 2) {
          size_t n;
          if (!cell_valid(flags, nbr27, b, x + (i == 0) * s, y + (i == 1) * s, z + (i == 2) * s, n) || !mat_air(flags[n]))
            continue;
          ret += pD[n] * ih[i] * ih[i];
        }
      const double fv = f[c] + ret;
      f[c] = fv;
      acc += fv * fv;
    }
  }
  write_partial(acc, partials, sh, false);
}

// Pressure write-back of Divergence_Projection: p_air on the Dirichlet
// boundary cells (0 on the other cells outside the domain).
__global__ void k_dirichlet_store(int nb, const uint8_t* __restrict__ code, const double* __restrict__ pD, double* __restrict__ x) {
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t c = (size_t)blockIdx.x * MAC_BC + t;
    if (!(code[c] & CODE_DOM)) x[c] = pD[c];
  }
}

__global__ void k_mask(int nb, const uint8_t* __restrict__ code, double* __restrict__ v) {
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)blockIdx.x * MAC_BC + t;
    if (!(code[c] & CODE_DOM)) v[c] = 0.0;
  }
}
