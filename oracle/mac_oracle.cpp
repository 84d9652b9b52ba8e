// mac_oracle.cpp -- CPU oracle for the MAC-grid hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker
// (or the timed CPU baseline); the product (mac_amd/libmac_amd.so) never
// calls into it.
//
// A from-scratch FP64 restatement of the reference's projection path in plain
// C++ loops, each function citing the reference file:line it follows.  The
// grid is the brick format of include/mac_amd.h (keys + flag bytes + staggered
// u,v,w); a hash of chunk keys gives neighbour bricks, as ChunkHashTable does.
//
// Parity pins (see tests/): the 1-D known answer of
// test/old/TestAlgorithms_Solve_Pressure.h:56-200, the uniform-flow answer of
// test/old/TestSolvePressureUmfpack.h:271-336, A symmetric, A = D.G on domain
// rows, and scipy SuperLU (a direct LU, like UMFPACK) on the assembled
// systems, committed as golden vectors.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

namespace {

constexpr int N = 8, NC = 512;

inline uint64_t pack(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
         (uint64_t)(uint32_t)(z + (1 << 20));
}

// floor division of a cell key into its chunk key, ChunkHashTable.tpp:181-196
inline int chunk_of(int k) { return k >= 0 ? k / N : (k + 1) / N - 1; }
// Modulo.h: non-negative remainder
inline int mod_n(int k) {
  int m = k % N;
  return m < 0 ? m + N : m;
}

struct Grid {
  int nb;
  const int32_t* keys;
  const uint8_t* flags;
  std::unordered_map<uint64_t, int> map;
  Grid(int nb_, const int32_t* k, const uint8_t* f) : nb(nb_), keys(k), flags(f) {
    map.reserve(nb * 2);
    for (int b = 0; b < nb; ++b) map[pack(k[3 * b], k[3 * b + 1], k[3 * b + 2])] = b;
  }
  // global cell key of (brick, offset): Offset::ToKey, Offset.h:36-46
  void cell_key(int b, int t, int g[3]) const {
    g[0] = keys[3 * b] * N + (t & 7);
    g[1] = keys[3 * b + 1] * N + ((t >> 3) & 7);
    g[2] = keys[3 * b + 2] * N + (t >> 6);
  }
  // index of the cell with global key g, or -1 if its chunk is absent
  long long index(const int g[3]) const {
    auto it = map.find(pack(chunk_of(g[0]), chunk_of(g[1]), chunk_of(g[2])));
    if (it == map.end()) return -1;
    return (long long)it->second * NC + mod_n(g[0]) + N * mod_n(g[1]) + N * N * mod_n(g[2]);
  }
  // neighbour (i in 0..2, s = +-1); -1 when invalid (Offset_Base_Neighbour.h:14-33:
  // the chunk must exist and the cell's existence bit must be set)
  long long nbr(long long c, int i, int s) const {
    int g[3];
    cell_key((int)(c / NC), (int)(c % NC), g);
    g[i] += s;
    long long n = index(g);
    if (n < 0 || !(flags[n] & 0x80)) return -1;
    return n;
  }
  bool exists(long long c) const { return flags[c] & 0x80; }
  // Data_CellType_Domain_Fluid_Without_0_Pres.h:59-62 with the material
  // predicates of ..._0_Pres.tpp:25-130: Fluid, Inflow, Interior
  bool in_domain(long long c) const {
    if (!exists(c)) return false;
    int m = flags[c] & 7;
    return m == 1 || m == 3 || m == 4;
  }
  bool cst(long long c, int i) const { return flags[c] & (0x08 << i); }
  // Data_CellType_Interface_Constant.h:46-61
  bool neumann(long long c, int i, int s) const {
    long long n = nbr(c, i, s);
    if (n < 0) return false;
    return s == 1 ? cst(n, i) : cst(c, i);
  }
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------- storage --
int orc_chunk_key(int k) { return chunk_of(k); }
// Chunk key for a general chunk size n (ChunkHashTable.tpp:181-196).
int orc_chunk_key_n(int k, int n) { return k >= 0 ? k / n : (k + 1) / n - 1; }
// Neighbour of offset `off` inside an n^dim chunk (x fastest) along direction
// i (0-based) with sign dir; returns 1 when it wraps into the neighbouring
// chunk (the table of Offset_GetNeighbour.h:12-43, restated digit-wise).
int orc_offset_neighbour(int n, int dim, int i, int dir, int off, int* res) {
  int dig[8] = {0, 0, 0, 0, 0, 0, 0, 0}, o = off;
  for (int d = 0; d < dim; ++d) {
    dig[d] = o % n;
    o /= n;
  }
  int v = dig[i] + dir;
  int wrap = (v < 0 || v >= n) ? 1 : 0;
  dig[i] = (v % n + n) % n;
  int r = 0, b = 1;
  for (int d = 0; d < dim; ++d) {
    r += dig[d] * b;
    b *= n;
  }
  *res = r;
  return wrap;
}
int orc_offset(int kx, int ky, int kz) { return mod_n(kx) + N * mod_n(ky) + N * N * mod_n(kz); }

// ------------------------------------------------------------- assembly ----
// Algorithms_Solve_Pressure::Init_Iteration (Algorithms_Solve_Pressure.h:44-157):
// rows numbered by a BFS (deque) started from each unnumbered cell in grid
// iteration order (here: brick order, then offset order); per row an ordered
// map col -> value; diagonal -= inv_h^2 per face, += inv_h^2 back for a
// Neumann face with a valid neighbour, off-diagonal +inv_h^2 for an in-domain
// neighbour across a non-Neumann face.
// off must hold ncell+1 ints, idx/val 7*ncell; row_of[ncell] gets the row or -1.
// Returns n (the number of rows).
int orc_assemble(int nb, const int32_t* keys, const uint8_t* flags, const double* inv_h, int* off, int* idx, double* val,
                 int* row_of) {
  Grid G(nb, keys, flags);
  long long ncell = (long long)nb * NC;
  for (long long c = 0; c < ncell; ++c) row_of[c] = -1;
  std::vector<int> rowstart;  // filled per row as in the reference's m_offset[lay_cur+1]
  int lay = 0;
  int pos = 0;
  std::vector<std::pair<int, std::map<int, double>>> rows;
  std::vector<std::map<int, double>> rowmap;
  std::deque<long long> next;
  for (long long c0 = 0; c0 < ncell; ++c0) {
    if (!G.exists(c0)) continue;
    if (row_of[c0] >= 0) continue;
    next.push_back(c0);
    while (!next.empty()) {
      long long c = next.front();
      next.pop_front();
      if (!G.in_domain(c)) continue;
      int cur;
      if (row_of[c] < 0) {
        row_of[c] = lay;
        cur = lay++;
      } else {
        cur = row_of[c];
      }
      std::map<int, double> m;
      double diag = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int s = -1; s <= 1; s += 2) {
          diag -= std::pow(inv_h[i], 2);
          long long n = G.nbr(c, i, s);
          if (n >= 0) {
            if (G.neumann(c, i, s)) {
              diag += std::pow(inv_h[i], 2);
            } else if (G.in_domain(n)) {
              int c2;
              if (row_of[n] < 0) {
                row_of[n] = lay;
                c2 = lay++;
                next.push_back(n);
              } else {
                c2 = row_of[n];
              }
              m.insert({c2, std::pow(inv_h[i], 2)});
            }
          }
        }
      m.insert({cur, diag});
      if ((int)rowmap.size() <= cur) rowmap.resize(cur + 1);
      rowmap[cur] = m;
    }
  }
  off[0] = 0;
  for (int r = 0; r < lay; ++r) {
    for (auto& kv : rowmap[r]) {
      idx[pos] = kv.first;
      val[pos] = kv.second;
      ++pos;
    }
    off[r + 1] = pos;
  }
  (void)rowstart;
  (void)rows;
  return lay;
}

// Policy_Divergence::Get_Divergence (Policy_Divergence.h:20-29), same operation
// order; an absent +neighbour reads speed 0.  Written for in-domain cells, 0
// elsewhere.
void orc_divergence(int nb, const int32_t* keys, const uint8_t* flags, const double* inv_h, const double* u, const double* v,
                    const double* w, double* div) {
  Grid G(nb, keys, flags);
  const double* U[3] = {u, v, w};
  long long ncell = (long long)nb * NC;
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < ncell; ++c) {
    if (!G.in_domain(c)) {
      div[c] = 0.0;
      continue;
    }
    double ret = 0.0;
    for (int i = 0; i < 3; ++i) {
      ret -= U[i][c] * inv_h[i];
      int g[3];
      G.cell_key((int)(c / NC), (int)(c % NC), g);
      g[i] += 1;
      long long n = G.index(g);
      ret += (n >= 0 ? U[i][n] : 0.0) * inv_h[i];
    }
    div[c] = ret;
  }
}

// Write-back and velocity update of Divergence_Projection
// (Algorithms_Solve_Pressure.h:184-207): p = solution on domain cells, 0
// elsewhere (caller passes p already in cell layout); for each existing cell
// and direction i, if Get_Pressure_If_Correction(c,i)
// (Policy_Pressure_If_Correction.h:33-49) then
// Set(i, u_i - Get_Gradiant(c,i)) (Policy_Gradiant.h:32-37), which skips
// const faces (Data_Grid_Speed.tpp:20-26).
void orc_gradient_update(int nb, const int32_t* keys, const uint8_t* flags, const double* inv_h, const double* p, double* u,
                         double* v, double* w) {
  Grid G(nb, keys, flags);
  double* U[3] = {u, v, w};
  long long ncell = (long long)nb * NC;
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < ncell; ++c) {
    if (!G.exists(c)) continue;
    bool dom = G.in_domain(c);
    for (int i = 0; i < 3; ++i) {
      long long n = G.nbr(c, i, -1);
      bool corr = dom || (n >= 0 && G.in_domain(n));
      if (!corr || G.cst(c, i)) continue;
      double g = p[c] * inv_h[i];
      g -= (n >= 0 ? p[n] : 0.0) * inv_h[i];
      U[i][c] = U[i][c] - g;
    }
  }
}

// Jacobi-preconditioned CG for the negative-definite CSR system A x = b
// (the UMFPACK slot's answer, Policy_Solve_Linear_Umfpack.h:56-70), FP64,
// to ||b - A x|| <= rtol ||b||.  Returns iterations, -1 if not converged.
int orc_cg_csr(int n, const int* off, const int* idx, const double* val, const double* b, double* x, double rtol,
               int maxit) {
  std::vector<double> r(n), z(n), p(n, 0.0), q(n), dg(n, 0.0);
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    for (int k = off[i]; k < off[i + 1]; ++k)
      if (idx[k] == i) dg[i] += -val[k];
    r[i] = -b[i];
  }
  double bb = 0;
  for (int i = 0; i < n; ++i) bb += r[i] * r[i];
  if (bb == 0) return 0;
  double rz = 0;
  for (int i = 0; i < n; ++i) {
    z[i] = r[i] / dg[i];
    rz += r[i] * z[i];
  }
  double beta = 0;
  for (int it = 1; it <= maxit; ++it) {
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    double pq = 0;
    for (int i = 0; i < n; ++i) {
      double t = 0;
      for (int k = off[i]; k < off[i + 1]; ++k) t += val[k] * p[idx[k]];
      q[i] = -t;
      pq += p[i] * q[i];
    }
    double a = rz / pq, rr = 0;
    for (int i = 0; i < n; ++i) {
      x[i] += a * p[i];
      r[i] -= a * q[i];
      rr += r[i] * r[i];
    }
    if (rr <= rtol * rtol * bb) return it;
    double rz2 = 0;
    for (int i = 0; i < n; ++i) {
      z[i] = r[i] / dg[i];
      rz2 += r[i] * z[i];
    }
    beta = rz2 / rz;
    rz = rz2;
  }
  return -1;
}

// Full reference projection with the CSR direct-solve stand-in:
// Init_Iteration + Divergence_Projection of Algorithms_Solve_Pressure.h.
// p (cell layout) is written; u,v,w are updated in place; returns the max
// |div u| over the domain after the update (the :209-219 post-check) through
// *maxdiv, and the CG iterations.
int orc_project(int nb, const int32_t* keys, const uint8_t* flags, const double* h, double* u, double* v, double* w, double* p,
                double rtol, double* maxdiv) {
  double inv_h[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
  long long ncell = (long long)nb * NC;
  std::vector<int> off(ncell + 1), idx(7 * ncell), row(ncell);
  std::vector<double> val(7 * ncell), div(ncell);
  int n = orc_assemble(nb, keys, flags, inv_h, off.data(), idx.data(), val.data(), row.data());
  orc_divergence(nb, keys, flags, inv_h, u, v, w, div.data());
  std::vector<double> b(n), x(n, 0.0);
  for (long long c = 0; c < ncell; ++c)
    if (row[c] >= 0) b[row[c]] = div[c];
  int it = n > 0 ? orc_cg_csr(n, off.data(), idx.data(), val.data(), b.data(), x.data(), rtol, 100000) : 0;
  for (long long c = 0; c < ncell; ++c) p[c] = row[c] >= 0 ? x[row[c]] : 0.0;
  orc_gradient_update(nb, keys, flags, inv_h, p, u, v, w);
  orc_divergence(nb, keys, flags, inv_h, u, v, w, div.data());
  double m = 0;
  for (long long c = 0; c < ncell; ++c) m = std::max(m, std::fabs(div[c]));
  if (maxdiv) *maxdiv = m;
  return it;
}

// Policy_Dirichlet_Boundary.h: GetIsAir is Air or Solid; a Dirichlet boundary
// cell is an air cell with a valid non-air neighbour.  p_air follows
// GetDirichletBoundaryPressure (:78-104) term by term, with the real
// velocity ur (u(c+e_i) read from the chunk's memory, 0 for an absent chunk).
namespace {
bool is_air(const Grid& G, long long c) { return (G.flags[c] & 7) == 0 || (G.flags[c] & 7) == 2; }
double air_pressure(const Grid& G, long long c, const double* const U[3], const double* inv_h, double nu) {
  if (!G.exists(c) || !is_air(G, c)) return 0.0;
  double ret = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int s = -1; s <= 1; s += 2) {
      long long n = G.nbr(c, i, s);
      if (n < 0 || is_air(G, n)) continue;
      double v1 = U[i][c];
      int g[3];
      G.cell_key((int)(c / NC), (int)(c % NC), g);
      g[i] += 1;
      long long np = G.index(g);
      double v2 = np >= 0 ? U[i][np] : 0.0;
      ret -= 2 * nu * (v2 - v1) * inv_h[i] * s;
    }
  return ret;
}
}  // namespace

// Algorithms_Solve_Pressure_Dirichlet::Divergence_Projection (:167-218) with
// the CSR stand-in solve: b = div - sum of p_air h_i^-2 over the Dirichlet
// boundary neighbours (:170-180), p = solution in the domain, p_air on the
// Dirichlet boundary cells, 0 elsewhere (:183-197), then the update.  ur/vr/wr:
// the real velocity for p_air (NULL: u, v, w as given).
int orc_project_dirichlet(int nb, const int32_t* keys, const uint8_t* flags, const double* h, double* u, double* v, double* w,
                          const double* ur, const double* vr, const double* wr, double nu, double* p, double rtol,
                          double* maxdiv) {
  double inv_h[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
  long long ncell = (long long)nb * NC;
  Grid G(nb, keys, flags);
  std::vector<double> real[3];
  const double* U[3] = {ur ? ur : u, vr ? vr : v, wr ? wr : w};
  for (int i = 0; i < 3; ++i) {
    real[i].assign(U[i], U[i] + ncell);
    U[i] = real[i].data();
  }
  std::vector<double> pair(ncell);
  for (long long c = 0; c < ncell; ++c) pair[c] = air_pressure(G, c, U, inv_h, nu);
  std::vector<int> off(ncell + 1), idx(7 * ncell), row(ncell);
  std::vector<double> val(7 * ncell), div(ncell);
  int n = orc_assemble(nb, keys, flags, inv_h, off.data(), idx.data(), val.data(), row.data());
  orc_divergence(nb, keys, flags, inv_h, u, v, w, div.data());
  std::vector<double> b(n), x(n, 0.0);
  for (long long c = 0; c < ncell; ++c) {
    if (row[c] < 0) continue;
    double bv = div[c];
    double ret = 0.0;
    bool any = false;
    for (int i = 0; i < 3; ++i)
      for (int s = -1; s <= 1; s += 2) {
        long long nn = G.nbr(c, i, s);
        if (nn < 0 || !is_air(G, nn)) continue;
        any = true;
        ret += pair[nn] * inv_h[i] * inv_h[i];
      }
    if (any) bv -= ret;
    b[row[c]] = bv;
  }
  int it = n > 0 ? orc_cg_csr(n, off.data(), idx.data(), val.data(), b.data(), x.data(), rtol, 100000) : 0;
  for (long long c = 0; c < ncell; ++c) p[c] = row[c] >= 0 ? x[row[c]] : pair[c];
  orc_gradient_update(nb, keys, flags, inv_h, p, u, v, w);
  orc_divergence(nb, keys, flags, inv_h, u, v, w, div.data());
  double m = 0;
  for (long long c = 0; c < ncell; ++c) m = std::max(m, std::fabs(div[c]));
  if (maxdiv) *maxdiv = m;
  return it;
}

// y = A x, matrix-free, with the assembled operator (for parity of the GPU
// operator); x is read as 0 outside the domain.
void orc_apply_A(int nb, const int32_t* keys, const uint8_t* flags, const double* inv_h, const double* x, double* y) {
  Grid G(nb, keys, flags);
  long long ncell = (long long)nb * NC;
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < ncell; ++c) {
    if (!G.in_domain(c)) {
      y[c] = 0.0;
      continue;
    }
    double diag = 0.0, acc = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int s = -1; s <= 1; s += 2) {
        diag -= inv_h[i] * inv_h[i];
        long long n = G.nbr(c, i, s);
        if (n < 0) continue;
        if (G.neumann(c, i, s)) diag += inv_h[i] * inv_h[i];
        else if (G.in_domain(n)) acc += inv_h[i] * inv_h[i] * x[n];
      }
    y[c] = diag * x[c] + acc;
  }
}

}  // extern "C"

// =================================================================== MG ====
// CPU FP64 MG-preconditioned CG on dense arrays over the bounding box, the
// same algorithm as the device solver (coarse labels: Dirichlet if any child
// Dirichlet, else unknown if any child unknown, else Neumann; rediscretised
// 7-point operator per level; damped Jacobi nu+nu; full-weighting
// restriction = (1/8) x transpose of trilinear prolongation; exact dense
// solve on the single-brick coarsest level).  It is the CPU baseline
// standing in for the reference's pyamg solve (Algorithmes_Solve_Pressure_Dirichlet_Pyamg.h:289-294,408).

namespace {

struct DLevel {
  int n[3];
  double ih2[3];
  std::vector<uint8_t> code, lab;
  std::vector<double> r, z, zt, res;
  // coarse levels: the operator as face couplings cf[k] (-x,+x,-y,+y,-z,+z;
  // 0 toward a cell that is not an unknown) and the diagonal dg
  bool coef = false;
  std::vector<double> cf[6], dg;
  long long size() const { return (long long)n[0] * n[1] * n[2]; }
  long long id(int x, int y, int z) const { return x + (long long)n[0] * (y + (long long)n[1] * z); }
};

const int DX[6] = {-1, 1, 0, 0, 0, 0}, DY[6] = {0, 0, -1, 1, 0, 0}, DZ[6] = {0, 0, 0, 0, -1, 1};

struct MG {
  std::vector<DLevel> L;
  int nc = 0;
  std::vector<long long> cellof;
  std::vector<double> inv;
  double om_pre[8] = {0}, om_post[8] = {0};    // per-sweep weights on level 0 (the device's schedule)
  double omc_pre[8] = {0}, omc_post[8] = {0};  // on the coarse levels
  int nuc = 2;
  double omega = 0.8;                        // weight of the sweep being run
  int nu = 2;
  // coarse block smoother (mac_params.coarse_inner > 1): inner sweeps of each
  // 8^3 block with its neighbours frozen, scaled by damp
  int inner = 1;
  double omi[8] = {0};
  double damp = 1.0;
  // from level 3 on (those levels have few bricks): one undamped launch of
  // nuc * inner sweeps (at most 8) per side instead of nuc launches
  int inner_d = 1;
  double omi_d[8] = {0};
  bool deep = false;  // the launch being run is a deep-level one

  void jacobi_block(DLevel& l, const std::vector<double>& r, const std::vector<double>* zin, std::vector<double>& zout) {
    const int inner = deep ? inner_d : this->inner;
    const double* omi = deep ? omi_d : this->omi;
    const double damp = deep ? 1.0 : this->damp;
    const int bx = l.n[0] / 8, by = l.n[1] / 8, bz = l.n[2] / 8;
#pragma omp parallel for schedule(static)
    for (long long B = 0; B < (long long)bx * by * bz; ++B) {
      const int X0 = 8 * (int)(B % bx), Y0 = 8 * (int)((B / bx) % by), Z0 = 8 * (int)(B / ((long long)bx * by));
      double z0[512], zc[512], zn[512];
      for (int t = 0; t < 512; ++t) z0[t] = zin ? (*zin)[l.id(X0 + (t & 7), Y0 + ((t >> 3) & 7), Z0 + (t >> 6))] : 0.0;
      std::copy(z0, z0 + 512, zc);
      for (int k = 0; k < inner; ++k) {
        for (int t = 0; t < 512; ++t) {
          const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
          const long long c = l.id(X0 + x, Y0 + y, Z0 + z);
          const unsigned cd = l.code[c];
          double out = 0;
          if (cd & 0x40) {
            double acc = 0, d = l.coef ? l.dg[c] : 0.0;
            for (int q = 0; q < 6; ++q) {
              if (!(cd & (1u << q)) && !l.coef) continue;
              const double w = l.coef ? l.cf[q][c] : l.ih2[q >> 1];
              if (!l.coef) d += w;
              const int xx = x + DX[q], yy = y + DY[q], zz = z + DZ[q];
              double v;
              if (xx >= 0 && yy >= 0 && zz >= 0 && xx < 8 && yy < 8 && zz < 8) {
                v = zc[xx + 8 * yy + 64 * zz];
              } else {
                const int gx = X0 + xx, gy = Y0 + yy, gz = Z0 + zz;
                v = (zin && gx >= 0 && gy >= 0 && gz >= 0 && gx < l.n[0] && gy < l.n[1] && gz < l.n[2]) ? (*zin)[l.id(gx, gy, gz)] : 0.0;
              }
              acc += w * v;
            }
            const double lz = d * zc[t] - acc;
            out = d > 0 ? zc[t] + omi[k] * (r[c] - lz) / d : 0;
          }
          zn[t] = out;
        }
        std::copy(zn, zn + 512, zc);
      }
      for (int t = 0; t < 512; ++t) zout[l.id(X0 + (t & 7), Y0 + ((t >> 3) & 7), Z0 + (t >> 6))] = z0[t] + damp * (zc[t] - z0[t]);
    }
  }

  double lap(const DLevel& l, const std::vector<double>& v, int x, int y, int z, double& dg) const {
    long long c = l.id(x, y, z);
    unsigned cd = l.code[c];
    double acc = 0, d = 0;
    if (l.coef) d = l.dg[c];
    for (int k = 0; k < 6; ++k) {
      if (!(cd & (1u << k)) && !l.coef) continue;
      double w = l.coef ? l.cf[k][c] : l.ih2[k >> 1];
      if (!l.coef) d += w;
      int xx = x + DX[k], yy = y + DY[k], zz = z + DZ[k];
      if (xx < 0 || yy < 0 || zz < 0 || xx >= l.n[0] || yy >= l.n[1] || zz >= l.n[2]) continue;
      acc += w * v[l.id(xx, yy, zz)];
    }
    dg = d;
    return d * v[c] - acc;
  }

  void jacobi(DLevel& l, const std::vector<double>& r, const std::vector<double>* zin, std::vector<double>& zout) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l.n[2]; ++z)
      for (int y = 0; y < l.n[1]; ++y)
        for (int x = 0; x < l.n[0]; ++x) {
          long long c = l.id(x, y, z);
          double out = 0;
          if (l.code[c] & 0x40) {
            if (zin) {
              double dg;
              double lz = lap(l, *zin, x, y, z, dg);
              out = dg > 0 ? (*zin)[c] + omega * (r[c] - lz) / dg : 0;
            } else {
              double dg = 0;
              if (l.coef) dg = l.dg[c];
              else
                for (int k = 0; k < 6; ++k)
                  if (l.code[c] & (1u << k)) dg += l.ih2[k >> 1];
              out = dg > 0 ? omega * r[c] / dg : 0;
            }
          }
          zout[c] = out;
        }
  }

  void residual(DLevel& l, const std::vector<double>& r, const std::vector<double>& zv, std::vector<double>& res) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l.n[2]; ++z)
      for (int y = 0; y < l.n[1]; ++y)
        for (int x = 0; x < l.n[0]; ++x) {
          long long c = l.id(x, y, z);
          double dg;
          res[c] = (l.code[c] & 0x40) ? r[c] - lap(l, zv, x, y, z, dg) : 0.0;
        }
  }

  void restrict_(const DLevel& f, DLevel& cl) {
    const double wt[4] = {0.25, 0.75, 0.75, 0.25};
#pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < cl.n[2]; ++Z)
      for (int Y = 0; Y < cl.n[1]; ++Y)
        for (int X = 0; X < cl.n[0]; ++X) {
          long long C = cl.id(X, Y, Z);
          double acc = 0;
          if (cl.code[C] & 0x40) {
            for (int dz = 0; dz < 4; ++dz)
              for (int dy = 0; dy < 4; ++dy)
                for (int dx = 0; dx < 4; ++dx) {
                  int x = 2 * X + dx - 1, y = 2 * Y + dy - 1, z = 2 * Z + dz - 1;
                  if (x < 0 || y < 0 || z < 0 || x >= f.n[0] || y >= f.n[1] || z >= f.n[2]) continue;
                  acc += wt[dx] * (wt[dy] * wt[dz]) * f.res[f.id(x, y, z)];
                }
            acc *= 0.125;
          }
          cl.r[C] = acc;
        }
  }

  void prolong(DLevel& f, const DLevel& cl, std::vector<double>& zf) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < f.n[2]; ++z)
      for (int y = 0; y < f.n[1]; ++y)
        for (int x = 0; x < f.n[0]; ++x) {
          long long c = f.id(x, y, z);
          if (!(f.code[c] & 0x40)) continue;
          int X0 = x >> 1, Y0 = y >> 1, Z0 = z >> 1;
          int X1 = X0 + ((x & 1) ? 1 : -1), Y1 = Y0 + ((y & 1) ? 1 : -1), Z1 = Z0 + ((z & 1) ? 1 : -1);
          auto at = [&](int X, int Y, int Z) -> double {
            if (X < 0 || Y < 0 || Z < 0 || X >= cl.n[0] || Y >= cl.n[1] || Z >= cl.n[2]) return 0.0;
            return cl.z[cl.id(X, Y, Z)];
          };
          double acc = 0;
          acc += (27.0 / 64.0) * at(X0, Y0, Z0);
          acc += (9.0 / 64.0) * (at(X1, Y0, Z0) + at(X0, Y1, Z0) + at(X0, Y0, Z1));
          acc += (3.0 / 64.0) * (at(X1, Y1, Z0) + at(X1, Y0, Z1) + at(X0, Y1, Z1));
          acc += (1.0 / 64.0) * at(X1, Y1, Z1);
          zf[c] += acc;
        }
  }

  void coarse_solve(DLevel& l) {
    std::vector<double> rl(nc);
    for (int i = 0; i < nc; ++i) rl[i] = l.r[cellof[i]];
    std::fill(l.z.begin(), l.z.end(), 0.0);
    for (int i = 0; i < nc; ++i) {
      double s = 0;
      for (int j = 0; j < nc; ++j) s += inv[(size_t)j * nc + i] * rl[j];
      l.z[cellof[i]] = s;
    }
  }

  // z0 = V(r0) at level 0 (r0 = L[0].r, result in L[0].z)
  void vcycle() {
    int nl = (int)L.size();
    if (nl == 1) {
      coarse_solve(L[0]);
      return;
    }
    for (int l = 0; l < nl - 1; ++l) {
      DLevel& lv = L[l];
      const double* pre = l ? omc_pre : om_pre;
      const int n = l ? nuc : nu;
      if (l > 0 && l < nl - 2 && inner > 1) {
        deep = l >= 3 && n > 1;
        jacobi_block(lv, lv.r, nullptr, lv.z);
        for (int k = 1; k < (deep ? 1 : n); ++k) {
          jacobi_block(lv, lv.r, &lv.z, lv.zt);
          std::swap(lv.z, lv.zt);
        }
        deep = false;
      } else {
        omega = pre[0];
        jacobi(lv, lv.r, nullptr, lv.z);
        for (int k = 1; k < n; ++k) {
          omega = pre[k];
          jacobi(lv, lv.r, &lv.z, lv.zt);
          std::swap(lv.z, lv.zt);
        }
      }
      residual(lv, lv.r, lv.z, lv.res);
      restrict_(lv, L[l + 1]);
    }
    coarse_solve(L[nl - 1]);
    for (int l = nl - 2; l >= 0; --l) {
      DLevel& lv = L[l];
      prolong(lv, L[l + 1], lv.z);
      const double* post = l ? omc_post : om_post;
      const bool blk = l > 0 && l < nl - 2 && inner > 1;
      deep = blk && l >= 3 && (l ? nuc : nu) > 1;
      const int n = deep ? 1 : (l ? nuc : nu);
      for (int k = 0; k < n; ++k) {
        if (blk) {
          jacobi_block(lv, lv.r, &lv.z, lv.zt);
        } else {
          omega = post[k];
          jacobi(lv, lv.r, &lv.z, lv.zt);
        }
        std::swap(lv.z, lv.zt);
      }
      deep = false;
    }
  }
};

}  // namespace

extern "C" {

// MG-PCG projection solve on the CPU.  Inputs as orc_project; solves
// L p = -div u (L = -A) to ||f - L p|| <= rtol ||f|| and writes p (cell layout).
// Velocity is not updated (use orc_gradient_update).  out[0] = iterations,
// out[1] = solve ms (excluding set-up), out[2] = set-up ms, out[3] = true
// relative residual, out[4] = levels.  Returns 0, or -1 on non-convergence.
// Test hook: the next orc_mgpcg call copies one level of its hierarchy out
// after the set-up: dims (n[3], dense over the level's brick bounding box),
// and, when n0 n1 n2 <= cap, the stencil codes and the coarse operator's
// NCF = 7 arrays (couplings -x,+x,-y,+y,-z,+z, then the diagonal; zeros on
// level 0), each of n0 n1 n2 values in x-fastest order.
static int g_dump_level = -1;
static int* g_dump_dims = nullptr;
static uint8_t* g_dump_code = nullptr;
static double* g_dump_cf = nullptr;
static long long g_dump_cap = 0;
void orc_mg_dump_request(int level, int* dims, uint8_t* code, double* cf, long long cap) {
  g_dump_level = level;
  g_dump_dims = dims;
  g_dump_code = code;
  g_dump_cf = cf;
  g_dump_cap = cap;
}

int orc_mgpcg(int nb, const int32_t* keys, const uint8_t* flags, const double* h, const double* u, const double* v,
              const double* w, double* p, double rtol, int maxit, int nu, double omega, int smoother, double cheb_lower,
              int nu_coarse, double cheb_lower_coarse, int coarse_inner, double coarse_damp, int nthreads, double* out) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  auto t0 = std::chrono::steady_clock::now();
  double inv_h[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
  long long ncell = (long long)nb * NC;
  int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  for (int b = 0; b < nb; ++b)
    for (int i = 0; i < 3; ++i) {
      mn[i] = std::min(mn[i], (int)keys[3 * b + i]);
      mx[i] = std::max(mx[i], (int)keys[3 * b + i]);
    }
  MG mg;
  mg.nu = nu = std::min(8, std::max(1, nu));
  mg.nuc = nu_coarse > 0 ? std::min(8, nu_coarse) : nu;
  // weights: omega for every sweep, or the reciprocal Chebyshev nodes of
  // [a, 2] (mac_amd.h, enum mac_smoother), reversed after the coarse solve
  auto sched = [&](int n, double a, double* pre, double* post) {
    for (int k = 0; k < n; ++k) {
      double om = omega;
      if (smoother == 1) om = 1.0 / (0.5 * (2.0 + a) + 0.5 * (2.0 - a) * std::cos((2.0 * k + 1.0) * M_PI / (2.0 * n)));
      pre[k] = om;
      post[n - 1 - k] = om;
    }
  };
  sched(mg.nu, cheb_lower, mg.om_pre, mg.om_post);
  sched(mg.nuc, cheb_lower_coarse, mg.omc_pre, mg.omc_post);
  mg.inner = std::min(8, std::max(1, coarse_inner));
  mg.damp = coarse_damp;
  {
    double post[8];
    sched(mg.inner, cheb_lower_coarse, mg.omi, post);
    mg.inner_d = std::min(8, mg.nuc * mg.inner);
    sched(mg.inner_d, cheb_lower_coarse, mg.omi_d, post);
  }
  // level 0 dense over the brick bounding box
  {
    DLevel l0;
    for (int i = 0; i < 3; ++i) {
      l0.n[i] = N * (mx[i] - mn[i] + 1);
      l0.ih2[i] = inv_h[i] * inv_h[i];
    }
    mg.L.push_back(l0);
  }
  int span = 0;
  for (int i = 0; i < 3; ++i) span = std::max(span, mx[i] - mn[i]);
  int nl = 1;
  while ((span >> (nl - 1)) > 0) ++nl;
  for (int l = 1; l < nl; ++l) {
    DLevel cl;
    for (int i = 0; i < 3; ++i) {
      cl.n[i] = N * (((mx[i] - mn[i]) >> l) + 1);
      cl.ih2[i] = mg.L[l - 1].ih2[i] * 0.25;
    }
    mg.L.push_back(cl);
  }
  for (auto& l : mg.L) {
    long long s = l.size();
    l.code.assign(s, 0);
    l.lab.assign(s, 0);
    l.r.assign(s, 0.0);
    l.z.assign(s, 0.0);
    l.zt.assign(s, 0.0);
    l.res.assign(s, 0.0);
  }
  // level-0 flags in dense layout, then codes/labels
  DLevel& L0 = mg.L[0];
  std::vector<uint8_t> fl(L0.size(), 0);
  std::vector<long long> dense_of(ncell);
  for (int b = 0; b < nb; ++b)
    for (int t = 0; t < NC; ++t) {
      int x = (keys[3 * b] - mn[0]) * N + (t & 7), y = (keys[3 * b + 1] - mn[1]) * N + ((t >> 3) & 7),
          z = (keys[3 * b + 2] - mn[2]) * N + (t >> 6);
      long long d = L0.id(x, y, z);
      fl[d] = flags[(size_t)b * NC + t];
      dense_of[(size_t)b * NC + t] = d;
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L0.n[2]; ++z)
    for (int y = 0; y < L0.n[1]; ++y)
      for (int x = 0; x < L0.n[0]; ++x) {
        long long c = L0.id(x, y, z);
        unsigned f = fl[c];
        bool ex = f & 0x80;
        int m = f & 7;
        bool dom = ex && (m == 1 || m == 3 || m == 4);
        unsigned cd = 0;
        if (dom) {
          cd = 0x40;
          for (int k = 0; k < 6; ++k) {
            int xx = x + DX[k], yy = y + DY[k], zz = z + DZ[k];
            unsigned nf = 0;
            if (xx >= 0 && yy >= 0 && zz >= 0 && xx < L0.n[0] && yy < L0.n[1] && zz < L0.n[2]) nf = fl[L0.id(xx, yy, zz)];
            bool nvalid = nf & 0x80;
            unsigned cbit = 0x08u << (k >> 1);
            bool cstf = (k & 1) ? (nf & cbit) : (f & cbit);
            if (!(nvalid && cstf)) cd |= 1u << k;
          }
        }
        L0.code[c] = cd;
        L0.lab[c] = dom ? 2 : ((!ex || m == 2) ? 0 : 1);
      }
  for (int l = 1; l < nl; ++l) {
    DLevel& f = mg.L[l - 1];
    DLevel& cl = mg.L[l];
#pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < cl.n[2]; ++Z)
      for (int Y = 0; Y < cl.n[1]; ++Y)
        for (int X = 0; X < cl.n[0]; ++X) {
          bool anyD = false, anyF = false;
          for (int k = 0; k < 8; ++k) {
            int x = 2 * X + (k & 1), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k >> 2);
            if (x >= f.n[0] || y >= f.n[1] || z >= f.n[2]) continue;
            int lb = f.lab[f.id(x, y, z)];
            anyD |= lb == 1;
            anyF |= lb == 2;
          }
          cl.lab[cl.id(X, Y, Z)] = anyD ? 1 : (anyF ? 2 : 0);
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < cl.n[2]; ++Z)
      for (int Y = 0; Y < cl.n[1]; ++Y)
        for (int X = 0; X < cl.n[0]; ++X) {
          long long C = cl.id(X, Y, Z);
          unsigned cd = 0;
          if (cl.lab[C] == 2) {
            cd = 0x40;
            for (int k = 0; k < 6; ++k) {
              int xx = X + DX[k], yy = Y + DY[k], zz = Z + DZ[k];
              int nlb = 0;
              if (xx >= 0 && yy >= 0 && zz >= 0 && xx < cl.n[0] && yy < cl.n[1] && zz < cl.n[2]) nlb = cl.lab[cl.id(xx, yy, zz)];
              if (nlb != 0) cd |= 1u << k;
            }
          }
          cl.code[C] = cd;
        }
    {
      // The coarse operator: the Galerkin operator P0^T A P0 / 16 of the
      // piecewise-constant aggregation P0 (a coarse cell's value on its 8
      // children), which equals the rediscretised 7-point operator away from
      // boundaries.  Coupling to a face neighbour = the sum of the fine
      // couplings across the coarse face / 16 (0 when the neighbour is not
      // a coarse unknown: its share stays in the diagonal); diagonal = (the
      // children's diagonals - the couplings between children, both ways) /
      // 16.  A Neumann face or a Dirichlet cell inside a coarse cell enters
      // through the fine couplings it removes, so the coarse grids see the
      // boundary where the fine grid has it.
      cl.coef = true;
      for (int k = 0; k < 6; ++k) cl.cf[k].assign(cl.size(), 0.0);
      cl.dg.assign(cl.size(), 0.0);
      auto fcoup = [&](int x, int y, int z, int k) -> double {
        const long long c = f.id(x, y, z);
        if (f.coef) return f.cf[k][c];
        if (!(f.code[c] & (1u << k))) return 0.0;
        const int xx = x + DX[k], yy = y + DY[k], zz = z + DZ[k];
        if (xx < 0 || yy < 0 || zz < 0 || xx >= f.n[0] || yy >= f.n[1] || zz >= f.n[2]) return 0.0;
        return (f.code[f.id(xx, yy, zz)] & 0x40) ? f.ih2[k >> 1] : 0.0;
      };
      auto fdiag = [&](int x, int y, int z) -> double {
        const long long c = f.id(x, y, z);
        if (f.coef) return f.dg[c];
        double d = 0;
        for (int k = 0; k < 6; ++k)
          if (f.code[c] & (1u << k)) d += f.ih2[k >> 1];
        return d;
      };
#pragma omp parallel for collapse(2) schedule(static)
      for (int Z = 0; Z < cl.n[2]; ++Z)
        for (int Y = 0; Y < cl.n[1]; ++Y)
          for (int X = 0; X < cl.n[0]; ++X) {
            const long long C = cl.id(X, Y, Z);
            if (!(cl.code[C] & 0x40)) continue;
            double cs[6] = {0, 0, 0, 0, 0, 0}, d = 0;
            for (int q = 0; q < 8; ++q) {
              const int a = q & 1, b = (q >> 1) & 1, e = q >> 2;
              const int x = 2 * X + a, y = 2 * Y + b, z = 2 * Z + e;
              if (x >= f.n[0] || y >= f.n[1] || z >= f.n[2]) continue;
              if (!(f.code[f.id(x, y, z)] & 0x40)) continue;
              d += fdiag(x, y, z);
              const int side[6] = {a == 0, a == 1, b == 0, b == 1, e == 0, e == 1};  // face k leaves the block
              for (int k = 0; k < 6; ++k) {
                const double w = fcoup(x, y, z, k);
                if (side[k]) cs[k] += w;
                else d -= w;
              }
            }
            cl.dg[C] = d / 16.0;
            // a cell within 1e-12 of the rediscretised operator (on every
            // face toward an unknown) takes it exactly, as the device does
            const unsigned cc = cl.code[C];
            bool std_cell = true;
            double dstd = 0.0;
            for (int k = 0; k < 6; ++k) {
              const int XX = X + DX[k], YY = Y + DY[k], ZZ = Z + DZ[k];
              const bool nb_unknown = XX >= 0 && YY >= 0 && ZZ >= 0 && XX < cl.n[0] && YY < cl.n[1] && ZZ < cl.n[2] &&
                                      (cl.code[cl.id(XX, YY, ZZ)] & 0x40);
              cl.cf[k][C] = nb_unknown ? cs[k] / 16.0 : 0.0;
              const double want = (nb_unknown && (cc & (1u << k))) ? cl.ih2[k >> 1] : 0.0;
              std_cell = std_cell && std::fabs(cl.cf[k][C] - want) <= 1e-12 * cl.ih2[k >> 1];
              if (cc & (1u << k)) dstd += cl.ih2[k >> 1];
            }
            std_cell = std_cell && std::fabs(cl.dg[C] - dstd) <= 1e-12 * dstd;
            if (std_cell) {
              for (int k = 0; k < 6; ++k) cl.cf[k][C] = cl.cf[k][C] != 0.0 ? cl.ih2[k >> 1] : 0.0;
              cl.dg[C] = dstd;
            }
          }
    }
  }
  if (g_dump_level >= 0) {
    if (g_dump_level < nl) {
      const DLevel& dl = mg.L[g_dump_level];
      for (int i = 0; i < 3; ++i) g_dump_dims[i] = dl.n[i];
      const long long sz = dl.size();
      if (sz <= g_dump_cap) {
        std::copy(dl.code.begin(), dl.code.end(), g_dump_code);
        for (int k = 0; k < 7; ++k)
          for (long long c = 0; c < sz; ++c)
            g_dump_cf[k * sz + c] = dl.coef ? (k < 6 ? dl.cf[k][c] : dl.dg[c]) : 0.0;
      }
    } else {
      g_dump_dims[0] = g_dump_dims[1] = g_dump_dims[2] = 0;
    }
    g_dump_level = -1;
  }
  // coarsest dense inverse
  {
    DLevel& lc = mg.L[nl - 1];
    std::vector<int> rowof(lc.size(), -1);
    for (long long c = 0; c < lc.size(); ++c)
      if (lc.code[c] & 0x40) {
        rowof[c] = (int)mg.cellof.size();
        mg.cellof.push_back(c);
      }
    int n = mg.nc = (int)mg.cellof.size();
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
      long long c = mg.cellof[i];
      int x = (int)(c % lc.n[0]), y = (int)((c / lc.n[0]) % lc.n[1]), z = (int)(c / ((long long)lc.n[0] * lc.n[1]));
      if (lc.coef) A[(size_t)i * n + i] = lc.dg[c];
      for (int k = 0; k < 6; ++k) {
        if (!(lc.code[c] & (1u << k)) && !lc.coef) continue;
        double wgt = lc.coef ? lc.cf[k][c] : lc.ih2[k >> 1];
        if (!lc.coef) A[(size_t)i * n + i] += wgt;
        int xx = x + DX[k], yy = y + DY[k], zz = z + DZ[k];
        if (xx < 0 || yy < 0 || zz < 0 || xx >= lc.n[0] || yy >= lc.n[1] || zz >= lc.n[2]) continue;
        int j = rowof[lc.id(xx, yy, zz)];
        if (j >= 0) A[(size_t)i * n + j] -= wgt;
      }
    }
    double maxd = 0;
    for (int i = 0; i < n; ++i) maxd = std::max(maxd, A[(size_t)i * n + i]);
    std::vector<double> G;
    for (int attempt = 0; attempt < 2; ++attempt) {
      bool ok = true;
      G = A;
      if (attempt == 1)
        for (int i = 0; i < n; ++i) G[(size_t)i * n + i] += 1e-10 * maxd;
      for (int j = 0; j < n && ok; ++j) {
        double s = G[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= G[(size_t)j * n + k] * G[(size_t)j * n + k];
        if (s <= 1e-13 * maxd) {
          ok = false;
          break;
        }
        double gjj = std::sqrt(s);
        G[(size_t)j * n + j] = gjj;
        for (int i = j + 1; i < n; ++i) {
          double t2 = G[(size_t)i * n + j];
          for (int k = 0; k < j; ++k) t2 -= G[(size_t)i * n + k] * G[(size_t)j * n + k];
          G[(size_t)i * n + j] = t2 / gjj;
        }
      }
      if (ok) break;
    }
    mg.inv.assign((size_t)n * n, 0.0);
    std::vector<double> yv(n);
    for (int col = 0; col < n; ++col) {
      for (int i = 0; i < n; ++i) {
        double s = (i == col) ? 1.0 : 0.0;
        for (int k = 0; k < i; ++k) s -= G[(size_t)i * n + k] * yv[k];
        yv[i] = s / G[(size_t)i * n + i];
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = yv[i];
        for (int k = i + 1; k < n; ++k) s -= G[(size_t)k * n + i] * mg.inv[(size_t)k * n + col];
        mg.inv[(size_t)i * n + col] = s / G[(size_t)i * n + i];
      }
    }
  }
  // rhs f = -div u (dense)
  std::vector<double> div(ncell);
  orc_divergence(nb, keys, flags, inv_h, u, v, w, div.data());
  long long s0 = L0.size();
  std::vector<double> f(s0, 0.0), x(s0, 0.0), r(s0, 0.0), d(s0, 0.0), q(s0, 0.0);
  for (long long c = 0; c < ncell; ++c) f[dense_of[c]] = -div[c];
  auto t1 = std::chrono::steady_clock::now();
  double ff = 0;
#pragma omp parallel for reduction(+ : ff)
  for (long long c = 0; c < s0; ++c) ff += f[c] * f[c];
  int it = 0;
  double tol2 = rtol * rtol * ff;
  double true_rel = 0;
  bool conv = ff == 0;
  for (int restart = 0; restart < 4 && !conv; ++restart) {
    // r = f - L x
    double rr = 0;
#pragma omp parallel for collapse(2) reduction(+ : rr) schedule(static)
    for (int z = 0; z < L0.n[2]; ++z)
      for (int y = 0; y < L0.n[1]; ++y)
        for (int xx = 0; xx < L0.n[0]; ++xx) {
          long long c = L0.id(xx, y, z);
          double rv = 0;
          if (L0.code[c] & 0x40) {
            double dg;
            rv = f[c] - mg.lap(L0, x, xx, y, z, dg);
          }
          r[c] = rv;
          rr += rv * rv;
        }
    if (rr <= tol2) {
      conv = true;
      break;
    }
    std::fill(d.begin(), d.end(), 0.0);
    L0.r = r;
    mg.vcycle();
    double rz = 0;
#pragma omp parallel for reduction(+ : rz)
    for (long long c = 0; c < s0; ++c) rz += r[c] * L0.z[c];
    double beta = 0;
    while (it < maxit) {
      const std::vector<double>& zz = L0.z;
#pragma omp parallel for schedule(static)
      for (long long c = 0; c < s0; ++c) d[c] = zz[c] + beta * d[c];
      double dq = 0;
#pragma omp parallel for collapse(2) reduction(+ : dq) schedule(static)
      for (int z = 0; z < L0.n[2]; ++z)
        for (int y = 0; y < L0.n[1]; ++y)
          for (int xx = 0; xx < L0.n[0]; ++xx) {
            long long c = L0.id(xx, y, z);
            double qv = 0;
            if (L0.code[c] & 0x40) {
              double dg;
              qv = mg.lap(L0, d, xx, y, z, dg);
            }
            q[c] = qv;
            dq += d[c] * qv;
          }
      double alpha = rz / dq;
      rr = 0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
      for (long long c = 0; c < s0; ++c) {
        x[c] += alpha * d[c];
        r[c] -= alpha * q[c];
        rr += r[c] * r[c];
      }
      ++it;
      if (rr <= tol2) break;
      L0.r = r;
      mg.vcycle();
      double rz2 = 0;
#pragma omp parallel for reduction(+ : rz2)
      for (long long c = 0; c < s0; ++c) rz2 += r[c] * L0.z[c];
      beta = rz2 / rz;
      rz = rz2;
    }
    // true residual
    double tr = 0;
#pragma omp parallel for collapse(2) reduction(+ : tr) schedule(static)
    for (int z = 0; z < L0.n[2]; ++z)
      for (int y = 0; y < L0.n[1]; ++y)
        for (int xx = 0; xx < L0.n[0]; ++xx) {
          long long c = L0.id(xx, y, z);
          if (L0.code[c] & 0x40) {
            double dg;
            double rv = f[c] - mg.lap(L0, x, xx, y, z, dg);
            tr += rv * rv;
          }
        }
    true_rel = ff > 0 ? std::sqrt(tr / ff) : 0;
    conv = tr <= tol2 * 1.0000002;
    if (it >= maxit) break;
  }
  auto t2 = std::chrono::steady_clock::now();
  for (long long c = 0; c < ncell; ++c) p[c] = x[dense_of[c]];
  if (out) {
    out[0] = it;
    out[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    out[2] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    out[3] = true_rel;
    out[4] = nl;
  }
  return conv ? 0 : -1;
}

}  // extern "C"
