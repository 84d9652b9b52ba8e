// project_kernels.h -- HIP kernels of the free-surface pressure projection.
//
// Every kernel runs one workgroup per active brick (a brick holding at least
// one unknown) of its level, 256 threads, two cells per thread.  All vectors
// are zero outside the domain, so the 7-point operator needs only the cell's
// own code byte (see brick.h).  Reductions are grid-wide and deterministic
// (grid_reduce), and their scalar consequences (alpha, beta, convergence) are
// computed on the device by the last block, so a CG iteration never waits on
// the host.
#pragma once
#include "brick.h"

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

// Operator row of a cell: returns L v at the cell (L = -A, A the reference's
// assembled matrix), dg = diagonal of L.  ih2 = inv_h^2 per direction.
template <typename T, typename TV>
__device__ __forceinline__ T lap_row(const TV* __restrict__ v, const int* __restrict__ nbr, int b, int x, int y, int z,
                                     unsigned cd, T vc, T ihx2, T ihy2, T ihz2, T& dg) {
  T acc = T(0), d = T(0);
  if (cd & 1u) { d += ihx2; acc += ihx2 * (T)ld_rel(v, nbr, b, x - 1, y, z); }
  if (cd & 2u) { d += ihx2; acc += ihx2 * (T)ld_rel(v, nbr, b, x + 1, y, z); }
  if (cd & 4u) { d += ihy2; acc += ihy2 * (T)ld_rel(v, nbr, b, x, y - 1, z); }
  if (cd & 8u) { d += ihy2; acc += ihy2 * (T)ld_rel(v, nbr, b, x, y + 1, z); }
  if (cd & 16u) { d += ihz2; acc += ihz2 * (T)ld_rel(v, nbr, b, x, y, z - 1); }
  if (cd & 32u) { d += ihz2; acc += ihz2 * (T)ld_rel(v, nbr, b, x, y, z + 1); }
  dg = d;
  return d * vc - acc;
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

// f = -div(u) on in-domain cells (L p = f with L = -A), div optionally kept.
// Divergence in the exact operation order of Policy_Divergence.h:20-29.
__global__ void k_div(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                      const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w, double ihx,
                      double ihy, double ihz, double* __restrict__ f, double* __restrict__ div_out,
                      double* __restrict__ partials, Scal* s) {
#pragma clang fp contract(off)
  __shared__ double sh[16];
  int b = act[blockIdx.x];
  double acc = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    double dv = 0.0;
    if (code[c] & CODE_DOM) {
      double ret = 0.0;
      ret -= u[c] * ihx;
      ret += ld_rel(u, nbr, b, x + 1, y, z) * ihx;
      ret -= v[c] * ihy;
      ret += ld_rel(v, nbr, b, x, y + 1, z) * ihy;
      ret -= w[c] * ihz;
      ret += ld_rel(w, nbr, b, x, y, z + 1) * ihz;
      dv = ret;
    }
    if (div_out) div_out[c] = dv;
    f[c] = -dv;
    acc += dv * dv;
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) {
    s->ff = tot;
    s->done = (tot == 0.0) ? 1 : 0;
  }
}

// max |div u| over the domain (the reference's post-check, :209-219).
__global__ void k_maxdiv(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                         const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                         double ihx, double ihy, double ihz, double* __restrict__ partials, Scal* s) {
#pragma clang fp contract(off)
  __shared__ double sh[16];
  int b = act[blockIdx.x];
  double m = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    if (code[c] & CODE_DOM) {
      double ret = 0.0;
      ret -= u[c] * ihx;
      ret += ld_rel(u, nbr, b, x + 1, y, z) * ihx;
      ret -= v[c] * ihy;
      ret += ld_rel(v, nbr, b, x, y + 1, z) * ihy;
      ret -= w[c] * ihz;
      ret += ld_rel(w, nbr, b, x, y, z + 1) * ihz;
      m = fmax(m, fabs(ret));
    }
  }
  double part = block_max(m, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, true, &tot, sh) && threadIdx.x == 0) s->maxdiv = tot;
}

// r = f - L x and ||r||^2 (x = starting guess; zero outside the domain).
// mode 0: start of a solve (sets rr, convergence); mode 1: true-residual
// check only (total to s->tmp, r untouched).
__global__ void k_resid0(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                         const double* __restrict__ f, const double* __restrict__ x, double* __restrict__ r, double ihx2,
                         double ihy2, double ihz2, double* __restrict__ partials, Scal* s, int mode) {
  __shared__ double sh[16];
  int b = act[blockIdx.x];
  double acc = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int xx = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned cd = code[c];
    if (cd & CODE_DOM) {
      double dg;
      double lx = lap_row<double, double>(x, nbr, b, xx, y, z, cd, x[c], ihx2, ihy2, ihz2, dg);
      double rv = f[c] - lx;
      if (mode == 0) r[c] = rv;
      acc += rv * rv;
    }
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) {
    if (mode == 0) {
      s->rr = tot;
      s->rz_valid = 0;
      s->done = (tot <= s->tol2) ? 1 : 0;
    } else {
      s->tmp = tot;
    }
  }
}

// d_new = z + beta d_old (formed on the fly at the neighbours too), q = L d_new,
// and alpha = (r.z) / (d_new . q).
template <typename TZ>
__global__ void k_pmatvec(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                          const TZ* __restrict__ zv, const double* __restrict__ dold, double* __restrict__ dnew,
                          double* __restrict__ q, double ihx2, double ihy2, double ihz2, double* __restrict__ partials,
                          Scal* s) {
  if (s->done) return;
  __shared__ double sh[16];
  const double beta = s->beta;
  int b = act[blockIdx.x];
  double acc = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned cd = code[c];
    if (cd & CODE_DOM) {
      double dc = (double)zv[c] + beta * dold[c];
      double a = 0.0, dg = 0.0;
#define NBD(X, Y, Z) ((double)ld_rel(zv, nbr, b, X, Y, Z) + beta * ld_rel(dold, nbr, b, X, Y, Z))
      if (cd & 1u) { dg += ihx2; a += ihx2 * NBD(x - 1, y, z); }
      if (cd & 2u) { dg += ihx2; a += ihx2 * NBD(x + 1, y, z); }
      if (cd & 4u) { dg += ihy2; a += ihy2 * NBD(x, y - 1, z); }
      if (cd & 8u) { dg += ihy2; a += ihy2 * NBD(x, y + 1, z); }
      if (cd & 16u) { dg += ihz2; a += ihz2 * NBD(x, y, z - 1); }
      if (cd & 32u) { dg += ihz2; a += ihz2 * NBD(x, y, z + 1); }
#undef NBD
      double qv = dg * dc - a;
      dnew[c] = dc;
      q[c] = qv;
      acc += dc * qv;
    }
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) {
    s->dq = tot;
    if (tot > 0.0) {
      s->alpha = s->rz / tot;
    } else {
      s->alpha = 0.0;
      s->done = 3;
    }
  }
}

// x += alpha d, r -= alpha q, ||r||^2, convergence test.
__global__ void k_update(const int* __restrict__ act, const uint8_t* __restrict__ code, double* __restrict__ x,
                         double* __restrict__ r, const double* __restrict__ d, const double* __restrict__ q,
                         double* __restrict__ partials, Scal* s) {
  if (s->done) return;
  __shared__ double sh[16];
  const double alpha = s->alpha;
  int b = act[blockIdx.x];
  double acc = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    if (code[c] & CODE_DOM) {
      x[c] += alpha * d[c];
      double rv = r[c] - alpha * q[c];
      r[c] = rv;
      acc += rv * rv;
    }
  }
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) {
    s->rr = tot;
    s->iter += 1;
    if (tot <= s->tol2) s->done = 1;
    else if (s->iter >= s->max_iter) s->done = 2;
  }
}

// Damped Jacobi sweep: zout = zin + omega D^-1 (r - L zin); zin == nullptr
// means a zero start (zout = omega D^-1 r).  With want_dot the sweep also
// forms r.z and the CG beta.
template <typename TR, typename TZ>
__global__ void k_jacobi(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                         const TR* __restrict__ r, const TZ* __restrict__ zin, TZ* __restrict__ zout, TZ ihx2, TZ ihy2,
                         TZ ihz2, TZ omega, int want_dot, double* __restrict__ partials, Scal* s) {
  if (s->done) return;
  __shared__ double sh[16];
  int b = act[blockIdx.x];
  double acc = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned cd = code[c];
    TZ out = TZ(0);
    if (cd & CODE_DOM) {
      TZ rc = (TZ)r[c];
      if (zin) {
        TZ dg;
        TZ zc = zin[c];
        TZ lz = lap_row<TZ, TZ>(zin, nbr, b, x, y, z, cd, zc, ihx2, ihy2, ihz2, dg);
        out = dg > TZ(0) ? zc + omega * (rc - lz) / dg : TZ(0);
      } else {
        TZ dg = diag_of<TZ>(cd, ihx2, ihy2, ihz2);
        out = dg > TZ(0) ? omega * rc / dg : TZ(0);
      }
      if (want_dot) acc += (double)r[c] * (double)out;
    }
    zout[c] = out;
  }
  if (!want_dot) return;
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) {
    s->beta = s->rz_valid ? tot / s->rz : 0.0;
    s->rz = tot;
    s->rz_valid = 1;
  }
}

// res = r - L z.
template <typename TR, typename TZ>
__global__ void k_residual(const int* __restrict__ act, const int* __restrict__ nbr, const uint8_t* __restrict__ code,
                           const TR* __restrict__ r, const TZ* __restrict__ zv, TZ* __restrict__ res, TZ ihx2, TZ ihy2,
                           TZ ihz2, Scal* s) {
  if (s->done) return;
  int b = act[blockIdx.x];
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned cd = code[c];
    TZ out = TZ(0);
    if (cd & CODE_DOM) {
      TZ dg;
      TZ lz = lap_row<TZ, TZ>(zv, nbr, b, x, y, z, cd, zv[c], ihx2, ihy2, ihz2, dg);
      out = (TZ)r[c] - lz;
    }
    res[c] = out;
  }
}

// r_c = (1/8) W^T res_f over the coarse domain: cell-centred full weighting,
// weights (1/4, 3/4, 3/4, 1/4) per direction over fine cells 2X-1 .. 2X+2.
template <typename T>
__global__ void k_restrict(const int* __restrict__ actc, const int* __restrict__ nbrc, const int* __restrict__ childc,
                           const uint8_t* __restrict__ codec, const T* __restrict__ resf, T* __restrict__ rc, Scal* s) {
  if (s->done) return;
  __shared__ int fb[64];
  int C = actc[blockIdx.x];
  if (threadIdx.x < 64) {
    int rx = (threadIdx.x & 3) - 1, ry = ((threadIdx.x >> 2) & 3) - 1, rz = (threadIdx.x >> 4) - 1;
    int cn = nbrc[C * 27 + nbr_slot(rx >> 1, ry >> 1, rz >> 1)];
    fb[threadIdx.x] = cn >= 0 ? childc[cn * 8 + (rx & 1) + 2 * (ry & 1) + 4 * (rz & 1)] : -1;
  }
  __syncthreads();
  const T wt[4] = {T(0.25), T(0.75), T(0.75), T(0.25)};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int X = t & 7, Y = (t >> 3) & 7, Z = t >> 6;
    size_t c = (size_t)C * MAC_BC + t;
    T acc = T(0);
    if (codec[c] & CODE_DOM) {
      for (int dz = 0; dz < 4; ++dz) {
        int gz = 2 * Z + dz - 1;
        int bz = ((gz + 8) >> 3);
        for (int dy = 0; dy < 4; ++dy) {
          int gy = 2 * Y + dy - 1;
          int by = ((gy + 8) >> 3);
          T wyz = wt[dy] * wt[dz];
          for (int dx = 0; dx < 4; ++dx) {
            int gx = 2 * X + dx - 1;
            int bx = ((gx + 8) >> 3);
            int f = fb[bx + 4 * by + 16 * bz];
            if (f >= 0) acc += wt[dx] * wyz * resf[(size_t)f * MAC_BC + (gx & 7) + 8 * (gy & 7) + 64 * (gz & 7)];
          }
        }
      }
      acc *= T(0.125);
    }
    rc[c] = acc;
  }
}

// z_f += W z_c on the fine domain: cell-centred trilinear interpolation,
// weight 3/4 from the parent cell and 1/4 from its neighbour toward the
// fine cell, per direction.
template <typename T>
__global__ void k_prolong(const int* __restrict__ actf, const int* __restrict__ parent, const int* __restrict__ oct,
                          const int* __restrict__ nbrc, const uint8_t* __restrict__ codef, const T* __restrict__ zc,
                          T* __restrict__ zf, Scal* s) {
  if (s->done) return;
  int B = actf[blockIdx.x];
  int P = parent[B], o = oct[B];
  int bx = 4 * (o & 1), by = 4 * ((o >> 1) & 1), bz = 4 * (o >> 2);
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)B * MAC_BC + t;
    if (!(codef[c] & CODE_DOM)) continue;
    int X0 = bx + (x >> 1), Y0 = by + (y >> 1), Z0 = bz + (z >> 1);
    int X1 = X0 + ((x & 1) ? 1 : -1), Y1 = Y0 + ((y & 1) ? 1 : -1), Z1 = Z0 + ((z & 1) ? 1 : -1);
    T acc = T(0);
    acc += T(27.0 / 64.0) * ld_rel(zc, nbrc, P, X0, Y0, Z0);
    acc += T(9.0 / 64.0) * (ld_rel(zc, nbrc, P, X1, Y0, Z0) + ld_rel(zc, nbrc, P, X0, Y1, Z0) + ld_rel(zc, nbrc, P, X0, Y0, Z1));
    acc += T(3.0 / 64.0) * (ld_rel(zc, nbrc, P, X1, Y1, Z0) + ld_rel(zc, nbrc, P, X1, Y0, Z1) + ld_rel(zc, nbrc, P, X0, Y1, Z1));
    acc += T(1.0 / 64.0) * ld_rel(zc, nbrc, P, X1, Y1, Z1);
    zf[c] += acc;
  }
}

// Exact solve on the coarsest level (a single brick): z = Linv r with the
// dense inverse precomputed on the host (symmetric, so column j of Linv is
// read coalesced as row j).  cellof[i] = cell of unknown i, rowof[cell] = i
// or -1.
template <typename TR, typename TZ>
__global__ void k_coarse_solve(int nc, const int* __restrict__ rowof, const int* __restrict__ cellof,
                               const TZ* __restrict__ cinv, const TR* __restrict__ rc, TZ* __restrict__ zc,
                               int want_dot, double* __restrict__ partials, Scal* s) {
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
  if (!want_dot) return;
  double part = block_sum(acc, sh);
  double tot;
  if (grid_reduce(part, partials, &s->counter, false, &tot, sh) && threadIdx.x == 0) {
    s->beta = s->rz_valid ? tot / s->rz : 0.0;
    s->rz = tot;
    s->rz_valid = 1;
  }
}

// u_i -= p(c)/h_i - p(c-e_i)/h_i on every face of an existing cell whose
// cell or whose -e_i neighbour is in the domain, unless the face is const
// (Algorithms_Solve_Pressure.h:198-207; Policy_Pressure_If_Correction.h:33-49;
// Policy_Gradiant.h:32-37; Data_Grid_Speed.tpp:20-26).  Same operation order.
__global__ void k_gradient(const int* __restrict__ list, const int* __restrict__ nbr, const uint8_t* __restrict__ flags,
                           const uint8_t* __restrict__ code, const double* __restrict__ p, double* __restrict__ u,
                           double* __restrict__ v, double* __restrict__ w, double ihx, double ihy, double ihz) {
#pragma clang fp contract(off)
  int b = list[blockIdx.x];
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    size_t c = (size_t)b * MAC_BC + t;
    unsigned fl = flags[c];
    if (!(fl & 0x80u)) continue;
    bool dom = code[c] & CODE_DOM;
    double pc = p[c];
    {
      unsigned nd = ld_rel(code, nbr, b, x - 1, y, z);
      if ((dom || (nd & CODE_DOM)) && !(fl & 0x08u)) {
        double g = pc * ihx;
        g -= ld_rel(p, nbr, b, x - 1, y, z) * ihx;
        u[c] = u[c] - g;
      }
    }
    {
      unsigned nd = ld_rel(code, nbr, b, x, y - 1, z);
      if ((dom || (nd & CODE_DOM)) && !(fl & 0x10u)) {
        double g = pc * ihy;
        g -= ld_rel(p, nbr, b, x, y - 1, z) * ihy;
        v[c] = v[c] - g;
      }
    }
    {
      unsigned nd = ld_rel(code, nbr, b, x, y, z - 1);
      if ((dom || (nd & CODE_DOM)) && !(fl & 0x20u)) {
        double g = pc * ihz;
        g -= ld_rel(p, nbr, b, x, y, z - 1) * ihz;
        w[c] = w[c] - g;
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
      y[c] = -lap_row<double, double>(xv, nbr, b, x, yy, z, cd, xv[c], ihx2, ihy2, ihz2, dg);
    }
  }
}

// Zero the out-of-domain entries of a host-supplied vector (so that the
// operator sees the zero-outside-domain convention).
__global__ void k_mask(int nb, const uint8_t* __restrict__ code, double* __restrict__ v) {
  size_t c = (size_t)blockIdx.x * MAC_BC + threadIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    c = (size_t)blockIdx.x * MAC_BC + t;
    if (!(code[c] & CODE_DOM)) v[c] = 0.0;
  }
}
