// surface_bc.h -- the free-surface velocity boundary condition
// (Algorithm_Extrapolate_Boundary_Condition.h:44-986) on the brick grid.
//
// The reference visits the cells in hash order and its rules write faces that
// other cells' rules read, so its result depends on that order.  Here the
// rules run as seven phases, one or two launches each, every phase reading
// what the phases before it wrote:
//   P1 apply_set_nan (:145-175): NaN on every face between two air cells and
//      on the low face of an air cell whose low neighbour does not exist;
//   P2 the zero-divergence face of every cell with one air neighbour
//      (Apply_Div_3d :920-938 with the sum over both tangential directions;
//      Apply_Div_2d :402-420 in 2-D, on cells with a plane partner);
//   P3 the tangential-stress faces of the face-cell pairs (Apply_3d_Plane
//      :905-918 / Apply_2d_Plane :428-448), all read from the state P2
//      left; a face two pairs write takes the larger (normal axis, side);
//   P4 the 45-degree corners (Apply_3d_Plane_45 :940-957);
//   P5 the three-axis corners (SolveMatrix :578-778 on CreateMatrix's 3x3
//      system :548-576, compressed columns, Gaussian elimination with
//      partial pivoting), all read from the state P4 left;
//   P6 NaN on the air faces outside the domain (Get_Is_Speed_In_Domain,
//      Policy_Is_In_Domain_Speed.h:15-49) of the cells with four or more air
//      neighbours (three or more in 2-D, and one-air cells without a plane
//      partner);
//   P7 apply_nan_extrap (:182-245) on every boundary cell, its means read
//      from the state P6 left (GetNeighbour(d, 0) is the low neighbour).
// P3, P5 and P7 compute into a staging array (one value and one pending byte
// per face) that a second launch writes back; the others write in place,
// which is race-free because no face they write is read or written by
// another cell of the same phase.  Every write keeps const faces
// (non-forced Set) and skips cells that do not exist; a missing brick reads
// as zero, a missing cell of an existing brick reads its memory.
//
// dim = 3, or 2 for the x-y plane (Algorithm_Extrapolate_Boundary_Condition_Dim<2>;
// z faces untouched).
#pragma once
#include <math.h>
#include <stdint.h>

#include "brick.h"

struct BcArgs {
  int nb;
  const int* nbr;
  const uint8_t* flags;
  double* U[3];       // the face field (a speed slot or the acceleration)
  double* tmp[3];     // staged values
  uint8_t* pend[3];   // staged-value flags (zero between launches)
  const uint8_t* active;  // per brick: the brick or a neighbour holds both air and non-air cells
  uint8_t* cls;           // per cell: bits 0-5 its air neighbours (-x,+x,-y,+y,-z,+z) if it is an
                          // existing non-air cell, bit 6 existing non-air, bit 7 air (k_bc_classify)
  double h[3], ih[3];
  int dim;
  // per-axis members chosen by a select, not an index: an index into a
  // kernel argument's array would copy the argument to scratch memory
  __device__ __forceinline__ double* Uc(int c) const { return c == 0 ? U[0] : (c == 1 ? U[1] : U[2]); }
  __device__ __forceinline__ double* tmpc(int c) const { return c == 0 ? tmp[0] : (c == 1 ? tmp[1] : tmp[2]); }
  __device__ __forceinline__ uint8_t* pendc(int c) const { return c == 0 ? pend[0] : (c == 1 ? pend[1] : pend[2]); }
  __device__ __forceinline__ double hc(int c) const { return c == 0 ? h[0] : (c == 1 ? h[1] : h[2]); }
  __device__ __forceinline__ double ihc(int c) const { return c == 0 ? ih[0] : (c == 1 ? ih[1] : ih[2]); }
};

namespace bc {

__device__ __forceinline__ bool f_valid(unsigned f) { return (f & 0x80u) != 0u; }
__device__ __forceinline__ bool f_air(unsigned f) { return (f & 0x80u) && ((f & 7u) == 0u || (f & 7u) == 2u); }
__device__ __forceinline__ bool f_fluid(unsigned f) {
  const unsigned m = f & 7u;
  return (f & 0x80u) && (m == 1u || m == 3u || m == 4u || m == 5u);
}

// The air neighbours of a cell as a bit set (bit 2 axis + (side > 0)); the
// k-th of them in axis-major order found by clearing the lower set bits, so
// no array is indexed by a run-time value (that would live in scratch memory).
struct Bnd {
  unsigned m;
  int n;
  __device__ __forceinline__ int bit(int k) const {
    unsigned v = m;
    for (int i = 0; i < k; ++i) v &= v - 1u;
    return __builtin_ctz(v);
  }
  __device__ __forceinline__ int dir(int k) const { return bit(k) >> 1; }
  __device__ __forceinline__ int sign(int k) const { return (bit(k) & 1) ? 1 : -1; }
};

// A cell of brick b at local (x, y, z) and lookups at offsets from it.
struct Cell {
  const BcArgs& a;
  int b, x, y, z;
  __device__ Cell(const BcArgs& a_, int b_, int t) : a(a_), b(b_), x(t & 7), y((t >> 3) & 7), z(t >> 6) {}
  __device__ long long slot(int dx, int dy, int dz) const {
    int cell;
    const int nb = locate(a.nbr, b, x + dx, y + dy, z + dz, cell);
    return nb < 0 ? -1 : (long long)nb * MAC_BC + cell;
  }
  __device__ unsigned flag(int dx, int dy, int dz) const {
    const long long s = slot(dx, dy, dz);
    return s < 0 ? 0u : a.flags[s];
  }
  __device__ double F(int c, int dx, int dy, int dz) const {
    const long long s = slot(dx, dy, dz);
    return s < 0 ? 0.0 : a.Uc(c)[s];
  }
  // the face may be written: the cell exists and the face is not const
  __device__ long long target(int c, int dx, int dy, int dz) const {
    const long long s = slot(dx, dy, dz);
    if (s < 0) return -1;
    const unsigned f = a.flags[s];
    return ((f & 0x80u) && !(f & (0x08u << c))) ? s : -1;
  }
  __device__ void set(int c, int dx, int dy, int dz, double v) const {
    const long long s = target(c, dx, dy, dz);
    if (s >= 0) a.Uc(c)[s] = v;
  }
  __device__ void stage(int c, int dx, int dy, int dz, double v) const {
    const long long s = slot(dx, dy, dz);
    if (s < 0) return;
    a.tmpc(c)[s] = v;
    a.pendc(c)[s] = 1;
  }
  // IsBoundary (:88-119): the air neighbours of a valid non-air cell,
  // axis-major, from the cell's class byte
  __device__ Bnd bnd(int dx, int dy, int dz) const {
    Bnd r;
    const long long s = slot(dx, dy, dz);
    r.m = s < 0 ? 0u : (a.cls[s] & 0x3fu);
    r.n = __builtin_popcount(r.m);
    return r;
  }
  // the same from the cell's own flags and its neighbours' (k_bc_classify)
  __device__ unsigned classify() const {
    const unsigned f = flag(0, 0, 0);
    if (!f_valid(f)) return 0u;
    if (f_air(f)) return 0x80u;
    unsigned m = 0x40u;
    for (int i = 0; i < a.dim; ++i)
      for (int s = -1; s <= 1; s += 2)
        if (f_air(flag(s * (i == 0), s * (i == 1), s * (i == 2)))) m |= 1u << (2 * i + (s > 0));
    return m;
  }
  // a one-air cell with air axis d at (dx, dy, dz) + i e_t
  __device__ bool partner(int dx, int dy, int dz, int d, int t, int i) const {
    const Bnd p = bnd(dx + i * (t == 0), dy + i * (t == 1), dz + i * (t == 2));
    return p.n == 1 && p.dir(0) == d;
  }
  __device__ bool has_partner2d(int dx, int dy, int dz, const Bnd& r) const {
    const int t = r.dir(0) == 0 ? 1 : 0;
    return partner(dx, dy, dz, r.dir(0), t, -1) || partner(dx, dy, dz, r.dir(0), t, 1);
  }
  // Get_Is_Speed_In_Domain(N, d) for N at (dx, dy, dz): the face between N and N - e_d
  __device__ bool in_domain(int dx, int dy, int dz, int d) const {
    const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
    const unsigned f2 = flag(dx - ex, dy - ey, dz - ez);
    const bool f1 = f_fluid(flag(dx, dy, dz));
    if (!f_valid(f2)) return f1;
    const bool fl2 = f_fluid(f2);
    if (f1 && fl2) return true;
    if (!f1 && !fl2) return false;
    const unsigned f3 = f1 ? flag(dx + ex, dy + ey, dz + ez) : flag(dx - 2 * ex, dy - 2 * ey, dz - 2 * ez);
    if (!f_valid(f3)) return true;
    return !f_fluid(f3);
  }
};

}  // namespace bc

// One brick per block; each thread takes cells t, t + blockDim.x, ...
#define BC_FOR_CELLS        \
  const int b = blockIdx.x; \
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x)

// Per brick: bit 0 an existing air cell, bit 1 an existing non-air cell.
__global__ void k_bc_summary(int nb, const uint8_t* __restrict__ flags, uint8_t* __restrict__ summ) {
  const int b = blockIdx.x;
  bool air = false, other = false;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const unsigned f = flags[(size_t)b * MAC_BC + t];
    if (f & 0x80u) {
      if (bc::f_air(f))
        air = true;
      else
        other = true;
    }
  }
  const unsigned m = (__ballot(air) ? 1u : 0u) | (__ballot(other) ? 2u : 0u);
  __shared__ unsigned sm;
  if (threadIdx.x == 0) sm = 0;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && m) atomicOr(&sm, m);
  __syncthreads();
  if (threadIdx.x == 0) summ[b] = (uint8_t)sm;
}

// A brick is active when it and its 26 neighbours hold both kinds: every
// rule after P1 reads and writes within two cells of a cell with an air
// neighbour, which never leaves the 3x3x3 block of bricks.
__global__ void k_bc_active(int nb, const int* __restrict__ nbr, const uint8_t* __restrict__ summ, uint8_t* __restrict__ active) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  unsigned m = 0;
  for (int k = 0; k < 27; ++k) {
    const int o = nbr[(size_t)b * 27 + k];
    if (o >= 0) m |= summ[o];
  }
  active[b] = (m == 3u) ? 1 : 0;
}

// The class byte of every cell (BcArgs::cls); in an inactive brick no cell
// has a neighbour of the other kind, so only its own kind is recorded.
__global__ void k_bc_classify(BcArgs a) {
  const bool act = a.active[blockIdx.x] != 0;
  BC_FOR_CELLS {
    const size_t e = (size_t)b * MAC_BC + t;
    if (act) {
      bc::Cell c(a, b, t);
      a.cls[e] = (uint8_t)c.classify();
    } else {
      const unsigned f = a.flags[e];
      a.cls[e] = (uint8_t)(!bc::f_valid(f) ? 0u : bc::f_air(f) ? 0x80u : 0x40u);
    }
  }
}

// P1
__global__ void k_bc_set_nan(BcArgs a) {
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    if (!bc::f_air(a.flags[(size_t)b * MAC_BC + t])) continue;
    for (int i = 0; i < a.dim; ++i) {
      const unsigned lo = c.flag(-(i == 0), -(i == 1), -(i == 2));
      if (!bc::f_valid(lo) || bc::f_air(lo)) c.set(i, 0, 0, 0, NAN);
    }
  }
}

// P2
__global__ void k_bc_div(BcArgs a) {
#pragma clang fp contract(off)
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    const bc::Bnd r = c.bnd(0, 0, 0);
    if (r.n != 1) continue;
    if (a.dim == 2 && !c.has_partner2d(0, 0, 0, r)) continue;
    const int d = r.dir(0), sg = r.sign(0);
    const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
    double v;
    int tx, ty, tz;
    if (sg == -1) {
      v = c.F(d, ex, ey, ez);
      tx = ty = tz = 0;
    } else {
      v = c.F(d, 0, 0, 0);
      tx = ex;
      ty = ey;
      tz = ez;
    }
    double res;
    if (a.dim == 2) {
      const int tt = d == 0 ? 1 : 0;
      const double v2 = c.F(tt, 0, 0, 0), v3 = c.F(tt, tt == 0, tt == 1, 0);
      res = v - sg * a.ihc(tt) * a.hc(d) * (v3 - v2);
    } else {
      double sum = 0.0;
      for (int tt = 0; tt < 3; ++tt) {
        if (tt == d) continue;
        const double v2 = c.F(tt, 0, 0, 0), v3 = c.F(tt, tt == 0, tt == 1, tt == 2);
        sum += a.ihc(tt) * (v3 - v2);
      }
      res = v - sg * sum * a.hc(d);
    }
    c.set(d, tx, ty, tz, res);
  }
}

// P3, gathered per target face: component t on the low t side of cell X,
// from the pair (A, B) = (X - s e_d - e_t, X - s e_d) of one-air cells with
// air axis d whose writing cell has side s.
__global__ void k_bc_plane(BcArgs a) {
#pragma clang fp contract(off)
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    // a pair writes X's face when its writing cell's air neighbour is X (X
    // air) or X - e_t (then X's low t neighbour is air)
    const unsigned cx = a.cls[(size_t)b * MAC_BC + t];
    for (int tt = 0; tt < a.dim; ++tt) {
      if (!(cx & 0x80u) && !(cx & (1u << (2 * tt)))) continue;
      bool have = false;
      double val = 0.0;
      for (int d = 0; d < a.dim; ++d) {
        if (d == tt) continue;
        for (int s = -1; s <= 1; s += 2) {
          const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
          const int bx = -s * ex, by = -s * ey, bz = -s * ez;
          const int ax = bx - (tt == 0), ay = by - (tt == 1), az = bz - (tt == 2);
          const bc::Bnd rB = c.bnd(bx, by, bz);
          if (rB.n != 1 || rB.dir(0) != d) continue;
          const bc::Bnd rA = c.bnd(ax, ay, az);
          if (rA.n != 1 || rA.dir(0) != d) continue;
          if (rB.sign(0) != s && rA.sign(0) != s) continue;
          if (a.dim == 2 && !(c.has_partner2d(bx, by, bz, rB) || c.has_partner2d(ax, ay, az, rA))) continue;
          const int px = (s == 1) ? ex : 0, py = (s == 1) ? ey : 0, pz = (s == 1) ? ez : 0;
          const double v1 = c.F(d, ax + px, ay + py, az + pz), v2 = c.F(d, bx + px, by + py, bz + pz);
          const double v3 = c.F(tt, bx, by, bz);
          val = -s * a.ihc(tt) * a.hc(d) * (v2 - v1) + v3;
          have = true;
        }
      }
      if (have) c.stage(tt, 0, 0, 0, val);
    }
  }
}

// Writes the staged faces back (non-forced, existing cells) and clears the flags.
__global__ void k_bc_apply(BcArgs a) {
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    const size_t s = (size_t)b * MAC_BC + t;
    const unsigned f = a.flags[s];
    for (int i = 0; i < 3; ++i) {
      if (!a.pendc(i)[s]) continue;
      a.pendc(i)[s] = 0;
      if ((f & 0x80u) && !(f & (0x08u << i))) a.Uc(i)[s] = a.tmpc(i)[s];
    }
  }
}

// P4
__global__ void k_bc_corner45(BcArgs a) {
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    const bc::Bnd r = c.bnd(0, 0, 0);
    if (r.n != 2 || r.dir(0) == r.dir(1)) continue;
    for (int k = 0; k < 2; ++k) {
      const int d = r.dir(k);
      const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
      if (r.sign(k) == -1)
        c.set(d, 0, 0, 0, c.F(d, ex, ey, ez));
      else
        c.set(d, ex, ey, ez, c.F(d, 0, 0, 0));
    }
  }
}

// 3x3 Gaussian elimination with partial pivoting on [A | b]; NaN on a zero pivot.
__device__ inline void bc_gauss3(double A[3][4], double x[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    int piv = col;
    for (int i = col + 1; i < 3; ++i)
      if (fabs(A[i][col]) > fabs(A[piv][col])) piv = i;
    if (A[piv][col] == 0.0) {
      x[0] = x[1] = x[2] = NAN;
      return;
    }
    // the swap with the pivot row, row by row (no run-time row index)
#pragma unroll
    for (int i = col + 1; i < 3; ++i)
      if (i == piv)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double tmp = A[col][j];
          A[col][j] = A[i][j];
          A[i][j] = tmp;
        }
    for (int i = col + 1; i < 3; ++i) {
      const double m = A[i][col] / A[col][col];
      for (int j = col; j < 4; ++j) A[i][j] -= m * A[col][j];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double s = A[i][3];
    for (int j = i + 1; j < 3; ++j) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
}

// P5: the corner system of SolveMatrix / CreateMatrix for air on side s_i of
// every axis i.  Terms in the reference's order (b[0] terms 1-14, b[1] terms
// 1-14, b[2]); the matrix m_value[k] is column k / 3, row k % 3.
__global__ void k_bc_corner3(BcArgs a) {
#pragma clang fp contract(off)
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    const bc::Bnd r = c.bnd(0, 0, 0);
    if (r.n != 3) continue;
    // the side of each axis's air neighbour (the low one first, as in the
    // axis-major list; 0 without one)
    auto side = [&](int i) { return (r.m & (1u << (2 * i))) ? -1 : ((r.m & (2u << (2 * i))) ? 1 : 0); };
    const int s1 = side(0), s2 = side(1), s3 = side(2);
    if (!s1 || !s2 || !s3) continue;
    const double i1 = a.ihc(0), i2 = a.ihc(1), i3 = a.ihc(2);
    auto U = [&](int comp, int dx, int dy, int dz) { return c.F(comp, dx, dy, dz); };
    double b0 = 0;
    b0 += (s1 == 1) ? -2 * i1 * U(0, 0, 0, 0) : 2 * i1 * U(0, 1, 0, 0);
    b0 += (s2 == 1) ? 2 * i2 * U(1, 0, 0, 0) : -2 * i2 * U(1, 0, 1, 0);
    b0 += (s3 == 1) ? 0.5 * i1 * U(2, 0, 0, 0) : -0.5 * i1 * U(2, 0, 0, 1);
    b0 += s3 * 0.5 * i1 * (-U(2, -s1, 0, 0) - U(2, -s1, 0, 1));
    b0 += (s1 == 1) ? 0.5 * i3 * U(0, 0, 0, 0) : -0.5 * i3 * U(0, 1, 0, 0);
    b0 += s1 * 0.5 * i3 * (-U(0, 0, 0, -s3) - U(0, 1, 0, -s3));
    b0 += (s3 == 1) ? -0.5 * i2 * U(2, 0, 0, 0) : 0.5 * i2 * U(2, 0, 0, 1);
    b0 += s3 * 0.5 * i2 * (U(2, 0, -s2, 0) + U(2, 0, -s2, 1));
    b0 += (s2 == 1) ? -0.5 * i3 * U(1, 0, 0, 0) : 0.5 * i3 * U(1, 0, 1, 0);
    b0 += s2 * 0.5 * i3 * (U(1, 0, 0, -s3) + U(1, 0, 1, -s3));
    double b1 = 0;
    b1 += (s3 == 1) ? -2 * i3 * U(2, 0, 0, 0) : 2 * i3 * U(2, 0, 0, -s3);
    b1 += (s2 == 1) ? 2 * i2 * U(1, 0, 0, 0) : -2 * i2 * U(1, 0, 1, 0);
    b1 += (s3 == 1) ? 0.5 * i1 * U(2, 0, 0, 0) : -0.5 * i1 * U(2, 0, 0, 1);
    b1 += s3 * 0.5 * i1 * (-U(2, -s1, 0, 0) - U(2, -s1, 0, 1));
    b1 += (s1 == 1) ? 0.5 * i3 * U(0, 0, 0, 0) : -0.5 * i3 * U(0, 1, 0, 0);
    b1 += s1 * 0.5 * i3 * (-U(0, 0, 0, -s3) - U(0, 1, 0, -s3));
    b1 += (s1 == 1) ? -0.5 * i2 * U(0, 0, 0, 0) : -0.5 * i2 * U(0, 1, 0, 0);
    b1 += s1 * 0.5 * i2 * (U(0, 0, -s2, 0) + U(0, 1, -s2, 0));
    b1 += (s2 == 1) ? -0.5 * i1 * U(1, 0, 0, 0) : 0.5 * i1 * U(1, 0, 1, 0);
    b1 += s2 * 0.5 * i1 * (U(1, -s1, 0, 0) + U(1, -s1, 1, 0));
    double b2 = 0;
    b2 += (s1 == 1) ? i1 * U(0, 0, 0, 0) : -i1 * U(0, 1, 0, 0);
    b2 += (s2 == 1) ? i2 * U(1, 0, 0, 0) : -i2 * U(1, 0, 1, 0);
    b2 += (s3 == 1) ? i3 * U(2, 0, 0, 0) : -i3 * U(2, 0, 0, 1);
    double res[3];
    if (isnan(b0) || isnan(b1) || isnan(b2)) {
      res[0] = res[1] = res[2] = NAN;
    } else {
      double A[3][4];
      A[0][0] = -s1 * 2 * i1 - s1 * 0.5 * i3;
      A[1][0] = -s1 * 0.5 * i3 + s1 * 0.5 * i2;
      A[2][0] = s1 * i1;
      A[0][1] = s2 * 2 * i2 + s2 * 0.5 * i3;
      A[1][1] = s2 * 2 * i2 + s2 * 0.5 * i2;
      A[2][1] = s2 * i2;
      A[0][2] = -s3 * 0.5 * i1 + s2 * s3 * 0.5 * i2;
      A[1][2] = -s3 * 2 * i3 - s3 * 0.5 * i1;
      A[2][2] = s3 * i3;
      A[0][3] = b0;
      A[1][3] = b1;
      A[2][3] = b2;
      bc_gauss3(A, res);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (side(i) == 1)
        c.stage(i, i == 0, i == 1, i == 2, res[i]);
      else
        c.stage(i, 0, 0, 0, res[i]);
    }
  }
}

// P6
__global__ void k_bc_mark(BcArgs a) {
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    const bc::Bnd r = c.bnd(0, 0, 0);
    bool mark = a.dim == 3 ? r.n >= 4 : r.n >= 3;
    if (a.dim == 2 && r.n == 1 && !c.has_partner2d(0, 0, 0, r)) mark = true;
    if (!mark) continue;
    for (int k = 0; k < r.n; ++k) {
      const int d = r.dir(k);
      const int ox = r.sign(k) == 1 ? (d == 0) : 0, oy = r.sign(k) == 1 ? (d == 1) : 0, oz = r.sign(k) == 1 ? (d == 2) : 0;
      if (!c.in_domain(ox, oy, oz, d)) c.set(d, ox, oy, oz, NAN);
    }
  }
}

// P7: a boundary cell's NaN faces take the mean of the finite same-component
// faces of the face's neighbours (the low neighbour counted for offsets -1
// and 0 of every axis, the high one for +1), then share the cell's
// divergence equally.
__global__ void k_bc_nan_extrap(BcArgs a) {
#pragma clang fp contract(off)
  if (!a.active[blockIdx.x]) return;
  BC_FOR_CELLS {
    bc::Cell c(a, b, t);
    if (c.bnd(0, 0, 0).n == 0) continue;
    int nn = 0, tdir[6], tsgn[6];
    double cur[3][2];  // the cell's faces: [axis][low, high]
    for (int i = 0; i < a.dim; ++i) {
      const int ex = (i == 0), ey = (i == 1), ez = (i == 2);
      cur[i][0] = c.F(i, 0, 0, 0);
      cur[i][1] = c.F(i, ex, ey, ez);
      if (!bc::f_valid(c.flag(ex, ey, ez))) continue;
      if (isnan(cur[i][0])) {
        tdir[nn] = i;
        tsgn[nn] = 0;
        ++nn;
      }
      if (isnan(cur[i][1])) {
        tdir[nn] = i;
        tsgn[nn] = 1;
        ++nn;
      }
    }
    if (nn == 0) continue;
    bool wr[6];
    for (int k = 0; k < nn; ++k) {
      const int d = tdir[k];
      const int ox = tsgn[k] ? (d == 0) : 0, oy = tsgn[k] ? (d == 1) : 0, oz = tsgn[k] ? (d == 2) : 0;
      double vt = 0.0;
      int cnt = 0;
      for (int d2 = 0; d2 < a.dim; ++d2)
        for (int s = -1; s <= 1; ++s) {
          const int ss = s == 0 ? -1 : s;
          const int nx = ox + ss * (d2 == 0), ny = oy + ss * (d2 == 1), nz = oz + ss * (d2 == 2);
          if (!bc::f_valid(c.flag(nx, ny, nz))) continue;
          const double v = c.F(d, nx, ny, nz);
          if (!isnan(v)) {
            vt += v;
            ++cnt;
          }
        }
      if (cnt != 0) vt /= cnt;
      wr[k] = c.target(d, ox, oy, oz) >= 0;
      if (wr[k]) cur[d][tsgn[k]] = vt;
    }
    double div = 0.0;
    for (int i = 0; i < a.dim; ++i) div += a.ihc(i) * (cur[i][1] - cur[i][0]);
    const double cor = div / nn;
    for (int k = 0; k < nn; ++k) {
      const int d = tdir[k];
      const int ox = tsgn[k] ? (d == 0) : 0, oy = tsgn[k] ? (d == 1) : 0, oz = tsgn[k] ? (d == 2) : 0;
      const double nv = cur[d][tsgn[k]] - (2 * tsgn[k] - 1) * a.hc(d) * cor;
      if (wr[k]) {
        cur[d][tsgn[k]] = nv;
        c.stage(d, ox, oy, oz, nv);
      }
    }
  }
}
