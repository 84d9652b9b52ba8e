// sim_oracle.cpp -- CPU oracle for the per-timestep loop (topology, dt,
// forces, RK2, interpolation, particles).  TEST INFRASTRUCTURE ONLY (see
// mac_oracle.cpp's header): loaded by tests/ as the checker, never by the
// product.
//
// A sequential restatement of the reference algorithms on a sparse chunk map
// (std::map of chunk key -> 8^3 chunk), following the reference's own control
// flow: stack-based layer BFS (Algorithms_Update_CellType_Layer.h:22-95,
// Algorithms_Update_CellType_Layer_Depth.h:24-138 with its depth stacks),
// grid[key] inserting cells (ChunkHashTable.tpp:63-87), cell erase and chunk
// free (:133-161), per-particle sequential loops.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stack>
#include <tuple>
#include <vector>

extern "C" int orc_project(int nb, const int32_t* keys, const uint8_t* flags, const double* h, double* u, double* v, double* w,
                           double* p, double rtol, double* maxdiv);
extern "C" int orc_project_dirichlet(int nb, const int32_t* keys, const uint8_t* flags, const double* h, double* u, double* v,
                                     double* w, const double* ur, const double* vr, const double* wr, double nu, double* p,
                                     double rtol, double* maxdiv);

namespace {

struct K3 {
  int x, y, z;
  bool operator<(const K3& o) const { return std::tie(z, y, x) < std::tie(o.z, o.y, o.x); }
};

struct Chunk {
  uint8_t flags[512];
  int8_t layer[512];
  int16_t depth[512];
  double S[3][3][512];
  double A[3][512];
  double P[512];
  Chunk() {
    // the prototype cell: Air, no const, not existing, layer empty, zero fields
    memset(flags, 0, sizeof(flags));
    memset(layer, -1, sizeof(layer));
    // the reference's depth prototype is never initialised (test/Jet3.cpp:333-336);
    // a new cell's depth is overwritten before it is read, -1 is taken here
    for (int t = 0; t < 512; ++t) depth[t] = -1;
    memset(S, 0, sizeof(S));
    memset(A, 0, sizeof(A));
    memset(P, 0, sizeof(P));
  }
  bool any() const {
    for (int t = 0; t < 512; ++t)
      if (flags[t] & 0x80) return true;
    return false;
  }
};

inline int chunk_of(int k) { return k >= 0 ? k / 8 : (k + 1) / 8 - 1; }
inline int off_of(int kx, int ky, int kz) { return (kx & 7) + 8 * (ky & 7) + 64 * (kz & 7); }

inline bool is_fluid(unsigned m) { return m == 1 || m == 3 || m == 4 || m == 5; }
inline unsigned set_fluid(unsigned m) { return (m == 0 || m == 4) ? 1 : m; }
inline unsigned set_air(unsigned m) { return (m == 1 || m == 4) ? 0 : m; }

// same counter-based generator as the library's jittered inflow fill
inline double uni(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  uint64_t z = seed * 0xD1B54A32D192ED03ull + stream * 0x8CB92BA72F3D8DD7ull + a * 0x9E3779B97F4A7C15ull + (b + 1) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
inline uint64_t pack(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) | (uint64_t)(uint32_t)(z + (1 << 20));
}

struct Params {
  double h[3], g[3], nu, cfl, dt_min, dt_max;
  int layer_max, integrator, particle_rk, ppc;
  int depth_interior, depth_remove, floor_axis;  // depth_interior 0: the chain without depth
  double floor_value;
  int cube_n;      // > 0: Policy_Add_Particle_Cube lattice instead of the jitter
  double proj_nu;  // > 0: the viscous air pressures of the projection
  double inflow_speed[3], inflow_amp;  // the inflow pulsation (amp 0: off)
  int inflow_period;
  int surface_bc;  // 1: the free-surface velocity condition before every extrapolation
  uint64_t seed;
  double lo[3], hi[3];
};

// Gaussian elimination with partial pivoting of a 3x3 system [A | b]; the
// steps of the linear-solver slot's n <= 3 path (NaN on a zero pivot).
void gauss3(double A[3][4], double x[3]) {
  for (int c = 0; c < 3; ++c) {
    int piv = c;
    for (int i = c + 1; i < 3; ++i)
      if (std::fabs(A[i][c]) > std::fabs(A[piv][c])) piv = i;
    if (A[piv][c] == 0.0) {
      x[0] = x[1] = x[2] = std::nan("");
      return;
    }
    if (piv != c)
      for (int j = 0; j < 4; ++j) std::swap(A[c][j], A[piv][j]);
    for (int i = c + 1; i < 3; ++i) {
      const double m = A[i][c] / A[c][c];
      for (int j = c; j < 4; ++j) A[i][j] -= m * A[c][j];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double s = A[i][3];
    for (int j = i + 1; j < 3; ++j) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
}

// SolveMatrix (:578-778): the three air faces of a cell with air on one side
// of every axis (signs s1, s2, s3), from the 3x3 system of CreateMatrix
// (:557-576) solved by Gaussian elimination with partial pivoting (the n = 3
// path of the linear-solver slot); NaN if the right side is.  U(c, dx, dy, dz)
// is face component c of the cell at offset (dx, dy, dz) from the corner cell.
template <class UF>
void corner_system(const double h[3], int s1, int s2, int s3, UF U, double res[3]) {
  const double i1 = 1.0 / h[0], i2 = 1.0 / h[1], i3 = 1.0 / h[2];
  // b[0]
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
  // b[1]
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
  // b[2]
  double b2 = 0;
  b2 += (s1 == 1) ? i1 * U(0, 0, 0, 0) : -i1 * U(0, 1, 0, 0);
  b2 += (s2 == 1) ? i2 * U(1, 0, 0, 0) : -i2 * U(1, 0, 1, 0);
  b2 += (s3 == 1) ? i3 * U(2, 0, 0, 0) : -i3 * U(2, 0, 0, 1);
  if (std::isnan(b0) || std::isnan(b1) || std::isnan(b2)) {
    res[0] = res[1] = res[2] = std::nan("");
    return;
  }
  // CreateMatrix: m_value[k] in compressed columns of 3 (UMFPACK's layout,
  // umfpack_di_symbolic / umfpack_di_solve(UMFPACK_A, ...)): column k / 3,
  // row k % 3, so the third row is the cell's divergence
  double A[3][4];
  A[0][0] = -s1 * 2 * i1 - s1 * 0.5 * i3;     // m_value[0]
  A[0][1] = s2 * 2 * i2 + s2 * 0.5 * i3;      // m_value[3]
  A[0][2] = -s3 * 0.5 * i1 + s2 * s3 * 0.5 * i2;  // m_value[6]
  A[1][0] = -s1 * 0.5 * i3 + s1 * 0.5 * i2;   // m_value[1]
  A[1][1] = s2 * 2 * i2 + s2 * 0.5 * i2;      // m_value[4]
  A[1][2] = -s3 * 2 * i3 - s3 * 0.5 * i1;     // m_value[7]
  A[2][0] = s1 * i1;                          // m_value[2]
  A[2][1] = s2 * i2;                          // m_value[5]
  A[2][2] = s3 * i3;                          // m_value[8]
  A[0][3] = b0;
  A[1][3] = b1;
  A[2][3] = b2;
  gauss3(A, res);
}

struct Sim {
  Params P;
  double ih[3];
  std::map<K3, std::unique_ptr<Chunk>> grid;
  std::vector<std::array<double, 3>> parts;  // position slot 0
  std::vector<std::array<double, 3>> parts1, parts2;  // slots 1 and 2 (RK4 "all")
  uint64_t tick = 0;

  Chunk* chunk(int kx, int ky, int kz) {
    auto it = grid.find({chunk_of(kx), chunk_of(ky), chunk_of(kz)});
    return it == grid.end() ? nullptr : it->second.get();
  }
  // grid[key]: creates the chunk from the prototype if needed, sets the bit
  Chunk* insert(int kx, int ky, int kz) {
    K3 c{chunk_of(kx), chunk_of(ky), chunk_of(kz)};
    auto it = grid.find(c);
    if (it == grid.end()) it = grid.emplace(c, std::make_unique<Chunk>()).first;
    it->second->flags[off_of(kx, ky, kz)] |= 0x80;
    return it->second.get();
  }
  bool valid(int kx, int ky, int kz) {
    Chunk* ch = chunk(kx, ky, kz);
    return ch && (ch->flags[off_of(kx, ky, kz)] & 0x80);
  }
  // a speed read through GetNeighbour: the chunk's memory if the chunk exists
  double speed(int slot, int comp, int kx, int ky, int kz) {
    Chunk* ch = chunk(kx, ky, kz);
    return ch ? field(ch, slot, comp)[off_of(kx, ky, kz)] : 0.0;
  }
  // speed slot 0..2, or 3 = the acceleration
  static double* field(Chunk* ch, int slot, int comp) { return slot < 3 ? ch->S[slot][comp] : ch->A[comp]; }
  bool fluid(int kx, int ky, int kz) {
    Chunk* ch = chunk(kx, ky, kz);
    if (!ch) return false;
    unsigned f = ch->flags[off_of(kx, ky, kz)];
    return (f & 0x80) && is_fluid(f & 7);
  }
  int8_t layer_of(int kx, int ky, int kz) {
    Chunk* ch = chunk(kx, ky, kz);
    return ch ? ch->layer[off_of(kx, ky, kz)] : -1;
  }
  int16_t depth_of(int kx, int ky, int kz) {
    Chunk* ch = chunk(kx, ky, kz);
    return ch ? ch->depth[off_of(kx, ky, kz)] : -1;
  }
  template <typename F>
  void for_cells(F f) {
    for (auto& kv : grid)
      for (int t = 0; t < 512; ++t)
        if (kv.second->flags[t] & 0x80)
          f(kv.second.get(), t, kv.first.x * 8 + (t & 7), kv.first.y * 8 + ((t >> 3) & 7), kv.first.z * 8 + (t >> 6));
  }
  void key_of(const std::array<double, 3>& p, int k[3]) {
    for (int i = 0; i < 3; ++i) k[i] = (int)lround(p[i] * ih[i]);
  }

  // full: the step's chain; otherwise the chain inside an RK4 stage
  // (Algorithms_Create_Interior_Depth: no particles, no floor) with the keys
  // of position slot ks, every slot erased together
  void topology(bool full = true, int ks = 0) {
    const bool dm = P.depth_interior > 0;
    std::vector<std::array<double, 3>>& kparts = ks == 2 ? parts2 : parts;
    const bool multi = !full && parts1.size() == parts.size() && parts2.size() == parts.size();
    auto erase_where = [&](const std::vector<char>& keep) {
      std::vector<std::array<double, 3>>* all[3] = {&parts, &parts1, &parts2};
      for (int sl = 0; sl < (multi ? 3 : 1); ++sl) {
        std::vector<std::array<double, 3>>& v = *(sl == 0 ? &kparts : (ks == 2 ? all[sl - 1] : all[sl]));
        std::vector<std::array<double, 3>> out;
        for (size_t k = 0; k < v.size(); ++k)
          if (keep[k]) out.push_back(v[k]);
        v.swap(out);
      }
    };
    // Algorithms_Initialize_MacCell; with depth Algorithms_Initialize_Layer_And_Depth.h:16-33
    for_cells([&](Chunk* c, int t, int, int, int) {
      c->layer[t] = -1;
      const unsigned m = c->flags[t] & 7u;
      if (dm && is_fluid(m) && m != 4u) c->depth[t] = -1;
    });
    // Algorithm_DelParticle (test/Jet3.cpp:259-289): list erase, order kept
    if (full && P.floor_axis >= 0) {
      std::vector<char> keep(kparts.size());
      for (size_t k = 0; k < kparts.size(); ++k) keep[k] = !(kparts[k][P.floor_axis] < P.floor_value);
      erase_where(keep);
    }
    if (dm) {
      // Algorithms_Layer_Initial_With_Particle_Depth.h:29-46
      std::vector<char> keep(kparts.size(), 1);
      for (size_t q = 0; q < kparts.size(); ++q) {
        int k[3];
        key_of(kparts[q], k);
        Chunk* c = insert(k[0], k[1], k[2]);
        int o = off_of(k[0], k[1], k[2]);
        if ((c->flags[o] & 7u) == 4u) {
          c->layer[o] = 0;
          if (c->depth[o] > P.depth_remove) keep[q] = 0;
        } else {
          c->layer[o] = 0;
          c->flags[o] = (uint8_t)((c->flags[o] & ~7u) | set_fluid(c->flags[o] & 7u));
        }
      }
      erase_where(keep);
    } else {
      // Algorithms_Layer_Initial_With_Particle
      for (auto& p : kparts) {
        int k[3];
        key_of(p, k);
        Chunk* c = insert(k[0], k[1], k[2]);
        int o = off_of(k[0], k[1], k[2]);
        c->layer[o] = 0;
        c->flags[o] = (uint8_t)((c->flags[o] & ~7u) | set_fluid(c->flags[o] & 7u));
      }
    }
    // Algorithms_Create_Fluid_Particle with the jitter policy; with depth
    // Algorithms_Create_Fluid_Particle_Depth.h:29-44
    auto fill = [&](int kx, int ky, int kz) {
      if (!full) return;  // Algorithms_Create_Interior_Depth adds no particle
      if (P.cube_n > 0) {
        // Policy_Add_Particle_Cube::AddParticle (:29-59), x fastest
        const double dh = 1. / (2 * P.cube_n + 1);
        int tab[3] = {-P.cube_n, -P.cube_n, -P.cube_n};
        const int K[3] = {kx, ky, kz};
        while (true) {
          std::array<double, 3> r;
          for (int i = 0; i < 3; ++i) r[i] = (K[i] + tab[i] * dh) * P.h[i];
          parts.push_back(r);
          int j = 0;
          while (j < 3 && tab[j] >= P.cube_n) {
            tab[j] = -P.cube_n;
            ++j;
          }
          if (j < 3) ++tab[j];
          else break;
        }
        return;
      }
      uint64_t ck = pack(kx, ky, kz);
      for (int k = 0; k < P.ppc; ++k) {
        double ux = uni(P.seed, tick, ck, 3 * k), uy = uni(P.seed, tick, ck, 3 * k + 1), uz = uni(P.seed, tick, ck, 3 * k + 2);
        parts.push_back({((double)kx + (ux - 0.5)) * P.h[0], ((double)ky + (uy - 0.5)) * P.h[1], ((double)kz + (uz - 0.5)) * P.h[2]});
      }
    };
    for_cells([&](Chunk* c, int t, int kx, int ky, int kz) {
      unsigned f = c->flags[t];
      if ((P.ppc > 0 || P.cube_n > 0) && (f & 7u) == 3u && c->layer[t] == -1) {
        fill(kx, ky, kz);
        c->layer[t] = 0;
        c->flags[t] = (uint8_t)((f & ~7u) | set_fluid(f & 7u));
      } else if (dm && (f & 7u) == 4u && c->layer[t] == -1 && c->depth[t] <= P.depth_remove) {
        fill(kx, ky, kz);
        c->layer[t] = 0;
      } else if (dm && (f & 7u) == 4u) {
        c->layer[t] = 0;
      }
    });
    // Algorithms_Update_CellType_Layer(_Depth) (stack BFS); with depth every
    // cell's depth is emptied and the surface cells get depth 0 (:31-56)
    std::stack<std::array<int, 3>> s, s2;
    for_cells([&](Chunk* c, int t, int kx, int ky, int kz) {
      if (dm) c->depth[t] = -1;
      if (c->layer[t] != 0) return;
      bool b = false;
      for (int i = 0; i < 3 && !b; ++i)
        for (int si = -1; si <= 1; si += 2) {
          int n[3] = {kx, ky, kz};
          n[i] += si;
          if (!valid(n[0], n[1], n[2]) || layer_of(n[0], n[1], n[2]) == -1) {
            if (dm) c->depth[t] = 0;
            s.push({kx, ky, kz});
            b = true;
            break;
          }
        }
    });
    std::stack<std::array<int, 3>> sback = s;
    for (int lay = 1; lay <= P.layer_max; ++lay) {
      while (!s.empty()) {
        auto k = s.top();
        s.pop();
        for (int i = 0; i < 3; ++i)
          for (int si = -1; si <= 1; si += 2) {
            int n[3] = {k[0], k[1], k[2]};
            n[i] += si;
            if (!valid(n[0], n[1], n[2]) || layer_of(n[0], n[1], n[2]) == -1) {
              Chunk* c = insert(n[0], n[1], n[2]);
              int o = off_of(n[0], n[1], n[2]);
              c->layer[o] = (int8_t)lay;
              c->flags[o] = (uint8_t)((c->flags[o] & ~7u) | set_air(c->flags[o] & 7u));
              s2.push({n[0], n[1], n[2]});
            }
          }
      }
      std::swap(s, s2);
    }
    while (!s.empty()) {
      auto k = s.top();
      s.pop();
      for (int i = 0; i < 3; ++i) {
        int n[3] = {k[0], k[1], k[2]};
        n[i] += 1;
        if (!valid(n[0], n[1], n[2]) || layer_of(n[0], n[1], n[2]) == -1) {
          Chunk* c = insert(n[0], n[1], n[2]);
          int o = off_of(n[0], n[1], n[2]);
          c->layer[o] = (int8_t)(P.layer_max + 1);
          c->flags[o] = (uint8_t)((c->flags[o] & ~7u) | set_air(c->flags[o] & 7u));
        }
      }
    }
    // depth search inside the fluid (:96-133): Fluid below depth_interior,
    // Interior (SetInterior, unconditional) from there on
    for (int depth = 1; dm; ++depth) {
      while (!sback.empty()) {
        auto k = sback.top();
        sback.pop();
        for (int i = 0; i < 3; ++i)
          for (int si = -1; si <= 1; si += 2) {
            int n[3] = {k[0], k[1], k[2]};
            n[i] += si;
            if (!valid(n[0], n[1], n[2]) || layer_of(n[0], n[1], n[2]) != 0 || depth_of(n[0], n[1], n[2]) != -1) continue;
            Chunk* c = chunk(n[0], n[1], n[2]);
            int o = off_of(n[0], n[1], n[2]);
            c->depth[o] = (int16_t)depth;
            if (depth < P.depth_interior) c->flags[o] = (uint8_t)((c->flags[o] & ~7u) | set_fluid(c->flags[o] & 7u));
            else c->flags[o] = (uint8_t)((c->flags[o] & ~7u) | 4u);
            s2.push({n[0], n[1], n[2]});
          }
      }
      std::swap(s2, sback);
      if (sback.empty()) break;
    }
    // Algorithms_Delete_MacCell: erase layer-empty cells unless NoDelete, free empty chunks
    for (auto it = grid.begin(); it != grid.end();) {
      Chunk* c = it->second.get();
      for (int t = 0; t < 512; ++t)
        if ((c->flags[t] & 0x80) && c->layer[t] == -1 && !(c->flags[t] & 0x40)) c->flags[t] &= (uint8_t)~0x80u;
      if (!c->any()) it = grid.erase(it);
      else ++it;
    }
    if (full) tick++;
  }

  double compute_dt() {
    double mx = 0;
    for_cells([&](Chunk* c, int t, int, int, int) {
      if (!is_fluid(c->flags[t] & 7u)) return;
      double temp = 0;
      for (int i = 0; i < 3; ++i) temp += std::pow(c->S[0][i][t] * ih[i], 2);
      if (temp > mx) mx = temp;
    });
    double dt = mx > 0 ? 1 / std::sqrt(mx) * P.cfl : HUGE_VAL;
    if (dt < P.dt_min) dt = P.dt_min;
    if (dt > P.dt_max) dt = P.dt_max;
    return dt;
  }

  // Algorithms_Extrapolate::Do (Algorithms_Extrapolate.h:25-235), in its own
  // control flow: a frontier that starts at every cell of layer 1; for each
  // frontier cell A and direction i, the face between A and A+e_i (and A-e_i
  // and A) is set, when both cells have layer >= lay, to the mean of the
  // neighbouring parallel faces that touch a cell of layer < lay (and, when
  // A-e_i is missing or empty, A's own -e_i face from its +e_i and transverse
  // neighbours); cells with layer > lay reached that way form the next
  // frontier.  Non-forced Set: const faces keep their value.  A mean over no
  // face is 0/0, as in the reference.
  // Algorithm_Extrapolate_Boundary_Condition (the free-surface velocity
  // boundary condition, Algorithm_Extrapolate_Boundary_Condition.h:44-986) on
  // field `slot`, dim = 3 (Algorithm_Extrapolate_Boundary_Condition_Dim<3>)
  // or 2 (Dim<2>, the x-y plane; z faces untouched), in the defined form the
  // device follows.  The reference visits cells in hash order and every rule
  // writes faces other cells' rules read; here the rules run as phases, each
  // phase reading what the phases before it wrote:
  //   P1 apply_set_nan (:145-175) on every air cell;
  //   P2 the zero-divergence face of every cell with one air neighbour
  //      (Apply_Div_3d, :920-938, with the sum over both tangential
  //      directions where the reference's `res =` keeps only the last); in
  //      2-D Apply_Div_2d (:402-420) on the cells that have a plane partner;
  //   P3 the tangential-stress faces of the plane pairs (Apply_3d_Plane
  //      :905-918 / Apply_2d_Plane :428-448), every pair reading the state P2
  //      left; a face two pairs write takes the pair with the larger
  //      (normal axis, side);
  //   P4 the 45-degree corners (Apply_3d_Plane_45 :940-957);
  //   P5 the three-axis corners, a 3x3 solve each (SolveMatrix :578-778)
  //      reading the state before any of them writes;
  //   P6 the NaN marks of the remaining cases (Get_Is_Speed_In_Domain,
  //      Policy_Is_In_Domain_Speed.h:15-49), and in 2-D of the one-air cells
  //      without a plane partner;
  //   P7 apply_nan_extrap (:182-245) on every boundary cell, its averages
  //      read from the state before any cell writes.
  // A neighbour handle that the reference dereferences without a validity
  // check (undefined there) reads the chunk's memory here; an invalid cell
  // is never air, fluid or a boundary.  In 3-D a one-air cell's face is not
  // NaN-marked after P2: apply_nan_extrap would give it the same
  // zero-divergence value back.  Every Set is non-forced (const faces keep
  // their value).
  void surface_bc(int slot, int dim) {
    const double h[3] = {P.h[0], P.h[1], P.h[2]};
    auto F = [&](int c, int x, int y, int z) { return speed(slot, c, x, y, z); };
    auto S = [&](int c, int x, int y, int z, double v) {
      Chunk* ch = chunk(x, y, z);
      if (!ch) return;
      const int o = off_of(x, y, z);
      if (!(ch->flags[o] & 0x80)) return;
      if (!(ch->flags[o] & (0x08u << c))) field(ch, slot, c)[o] = v;
    };
    auto mat = [&](int x, int y, int z) -> int { return chunk(x, y, z)->flags[off_of(x, y, z)] & 7; };
    auto air = [&](int x, int y, int z) { return valid(x, y, z) && (mat(x, y, z) == 0 || mat(x, y, z) == 2); };
    auto fl = [&](int x, int y, int z) { return valid(x, y, z) && is_fluid(mat(x, y, z)); };
    struct Bnd {
      int n = 0, dir[6], sign[6];
    };
    // IsBoundary (:88-119): air neighbours of a valid non-air cell, axis-major
    auto bnd = [&](int x, int y, int z) {
      Bnd b;
      if (!valid(x, y, z) || air(x, y, z)) return b;
      for (int i = 0; i < dim; ++i)
        for (int s = -1; s <= 1; s += 2)
          if (air(x + s * (i == 0), y + s * (i == 1), z + s * (i == 2))) {
            b.dir[b.n] = i;
            b.sign[b.n] = s;
            ++b.n;
          }
      return b;
    };
    // Get_Is_Speed_In_Domain(N, dir): the face between N and N - e_dir
    auto in_domain = [&](int x, int y, int z, int d) {
      const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
      const int x2 = x - ex, y2 = y - ey, z2 = z - ez;
      if (!valid(x2, y2, z2)) return fl(x, y, z);
      const bool f1 = fl(x, y, z), f2 = fl(x2, y2, z2);
      if (f1 && f2) return true;
      if (!f1 && !f2) return false;
      int x3, y3, z3;
      if (f1) { x3 = x + ex; y3 = y + ey; z3 = z + ez; }
      else { x3 = x2 - ex; y3 = y2 - ey; z3 = z2 - ez; }
      if (!valid(x3, y3, z3)) return true;
      return !fl(x3, y3, z3);
    };
    std::vector<K3> cells;
    for_cells([&](Chunk*, int, int x, int y, int z) { cells.push_back({x, y, z}); });
    const double NaN = std::nan("");
    // partner of a one-air cell along axis t on side i: a valid non-air cell
    // with exactly one air neighbour, on the same axis
    auto partner = [&](int x, int y, int z, int d, int t, int i) {
      const Bnd b = bnd(x + i * (t == 0), y + i * (t == 1), z + i * (t == 2));
      return b.n == 1 && b.dir[0] == d;
    };
    auto has_partner2d = [&](int x, int y, int z, const Bnd& b) {
      const int t = b.dir[0] == 0 ? 1 : 0;
      return partner(x, y, z, b.dir[0], t, -1) || partner(x, y, z, b.dir[0], t, 1);
    };
    // P1
    for (const K3& c : cells) {
      if (!air(c.x, c.y, c.z)) continue;
      for (int i = 0; i < dim; ++i)
        for (int s = -1; s <= 1; s += 2) {
          const int x2 = c.x + s * (i == 0), y2 = c.y + s * (i == 1), z2 = c.z + s * (i == 2);
          if (valid(x2, y2, z2)) {
            if (air(x2, y2, z2)) {
              if (s == -1) S(i, c.x, c.y, c.z, NaN);
              else S(i, x2, y2, z2, NaN);
            }
          } else if (s == -1) {
            S(i, c.x, c.y, c.z, NaN);
          }
        }
    }
    // P2
    for (const K3& c : cells) {
      const Bnd b = bnd(c.x, c.y, c.z);
      if (b.n != 1) continue;
      if (dim == 2 && !has_partner2d(c.x, c.y, c.z, b)) continue;
      const int d = b.dir[0], sg = b.sign[0];
      const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
      double v = F(d, c.x, c.y, c.z);
      int tx = c.x + ex, ty = c.y + ey, tz = c.z + ez;
      if (sg == -1) {
        v = F(d, c.x + ex, c.y + ey, c.z + ez);
        tx = c.x; ty = c.y; tz = c.z;
      }
      double res;
      if (dim == 2) {
        const int t = d == 0 ? 1 : 0;
        const double v2 = F(t, c.x, c.y, c.z), v3 = F(t, c.x + (t == 0), c.y + (t == 1), c.z);
        res = v - sg * (1.0 / h[t]) * h[d] * (v3 - v2);
      } else {
        double sum = 0.0;
        for (int t = 0; t < 3; ++t) {
          if (t == d) continue;
          const double v2 = F(t, c.x, c.y, c.z), v3 = F(t, c.x + (t == 0), c.y + (t == 1), c.z + (t == 2));
          sum += (1.0 / h[t]) * (v3 - v2);
        }
        res = v - sg * sum * h[d];
      }
      S(d, tx, ty, tz, res);
    }
    // P3, gathered per target face: the face (component t) on the low t side
    // of cell X, written by the pair (A, B) = (X - s e_d - e_t, X - s e_d) of
    // one-air cells with air axis d where the pair's writing cell has side s.
    // Every pair reads the state P2 left (a face one pair writes can be the
    // air face another pair reads, on a staircase), so the values are
    // gathered first and written after.
    struct Pending {
      K3 c;
      int comp;
      double v;
    };
    std::vector<Pending> p3;
    for (const K3& X : cells) {
      for (int t = 0; t < dim; ++t) {
        bool have = false;
        double val = 0.0;
        for (int d = 0; d < dim; ++d) {
          if (d == t) continue;
          for (int s = -1; s <= 1; s += 2) {
            const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
            const int bx = X.x - s * ex, by = X.y - s * ey, bz = X.z - s * ez;
            const int ax = bx - (t == 0), ay = by - (t == 1), az = bz - (t == 2);
            const Bnd bB = bnd(bx, by, bz), bA = bnd(ax, ay, az);
            if (bB.n != 1 || bA.n != 1 || bB.dir[0] != d || bA.dir[0] != d) continue;
            // the writing cell: B (partner on its -t side) or A (on its +t side)
            if (bB.sign[0] != s && bA.sign[0] != s) continue;
            if (dim == 2 && !(has_partner2d(bx, by, bz, bB) || has_partner2d(ax, ay, az, bA))) continue;
            const int px = (s == 1) ? ex : 0, py = (s == 1) ? ey : 0, pz = (s == 1) ? ez : 0;
            const double v1 = F(d, ax + px, ay + py, az + pz), v2 = F(d, bx + px, by + py, bz + pz);
            const double v3 = F(t, bx, by, bz);
            val = -s * (1.0 / h[t]) * h[d] * (v2 - v1) + v3;
            have = true;
          }
        }
        if (have) p3.push_back({X, t, val});
      }
    }
    for (const Pending& q : p3) S(q.comp, q.c.x, q.c.y, q.c.z, q.v);
    // P4
    for (const K3& c : cells) {
      const Bnd b = bnd(c.x, c.y, c.z);
      if (b.n != 2 || b.dir[0] == b.dir[1]) continue;
      for (int k = 0; k < 2; ++k) {
        const int d = b.dir[k], sg = b.sign[k];
        const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
        if (sg == -1) S(d, c.x, c.y, c.z, F(d, c.x + ex, c.y + ey, c.z + ez));
        else S(d, c.x + ex, c.y + ey, c.z + ez, F(d, c.x, c.y, c.z));
      }
    }
    // P5
    if (dim == 3) {
      struct Corner {
        K3 c;
        double r[3];
        int s[3];
      };
      std::vector<Corner> todo;
      for (const K3& c : cells) {
        const Bnd b = bnd(c.x, c.y, c.z);
        if (b.n != 3) continue;
        int sa[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
          for (int k = 0; k < b.n; ++k)
            if (b.dir[k] == i) {
              sa[i] = b.sign[k];
              break;
            }
        if (!sa[0] || !sa[1] || !sa[2]) continue;
        Corner cr;
        cr.c = c;
        for (int i = 0; i < 3; ++i) cr.s[i] = sa[i];
        corner_solve(slot, c.x, c.y, c.z, sa[0], sa[1], sa[2], cr.r);
        todo.push_back(cr);
      }
      for (const Corner& cr : todo) {
        const K3& c = cr.c;
        for (int i = 0; i < 3; ++i) {
          if (cr.s[i] == 1) S(i, c.x + (i == 0), c.y + (i == 1), c.z + (i == 2), cr.r[i]);
          else S(i, c.x, c.y, c.z, cr.r[i]);
        }
      }
    }
    // P6
    for (const K3& c : cells) {
      const Bnd b = bnd(c.x, c.y, c.z);
      bool mark = dim == 3 ? b.n >= 4 : b.n >= 3;
      if (dim == 2 && b.n == 1 && !has_partner2d(c.x, c.y, c.z, b)) mark = true;
      if (!mark) continue;
      for (int k = 0; k < b.n; ++k) {
        const int d = b.dir[k];
        const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
        int x = c.x, y = c.y, z = c.z;
        if (b.sign[k] == 1) { x += ex; y += ey; z += ez; }
        if (!in_domain(x, y, z, d)) S(d, x, y, z, NaN);
      }
    }
    // P7
    {
      std::map<K3, std::array<double, 3>> snap;
      for (const K3& c : cells) snap[c] = {F(0, c.x, c.y, c.z), F(1, c.x, c.y, c.z), F(2, c.x, c.y, c.z)};
      auto G = [&](int comp, int x, int y, int z, double& v) {
        auto it = snap.find({x, y, z});
        if (it == snap.end()) return false;
        v = it->second[comp];
        return true;
      };
      for (const K3& c : cells) {
        if (bnd(c.x, c.y, c.z).n == 0) continue;
        int nn = 0, tdir[6], tsgn[6];
        for (int i = 0; i < dim; ++i) {
          const int ex = (i == 0), ey = (i == 1), ez = (i == 2);
          if (!valid(c.x + ex, c.y + ey, c.z + ez)) continue;
          if (std::isnan(F(i, c.x, c.y, c.z))) { tdir[nn] = i; tsgn[nn] = 0; ++nn; }
          if (std::isnan(F(i, c.x + ex, c.y + ey, c.z + ez))) { tdir[nn] = i; tsgn[nn] = 1; ++nn; }
        }
        if (nn == 0) continue;
        for (int k = 0; k < nn; ++k) {
          const int d = tdir[k];
          int x = c.x, y = c.y, z = c.z;
          if (tsgn[k] == 1) { x += (d == 0); y += (d == 1); z += (d == 2); }
          double vt = 0.0;
          int nb = 0;
          for (int d2 = 0; d2 < dim; ++d2)
            for (int s = -1; s <= 1; ++s) {
              const int ss = s == 0 ? -1 : s;  // GetNeighbour(d2, 0) is the -1 neighbour
              double a;
              if (G(d, x + ss * (d2 == 0), y + ss * (d2 == 1), z + ss * (d2 == 2), a) && !std::isnan(a)) {
                vt += a;
                ++nb;
              }
            }
          if (nb != 0) vt /= nb;
          S(d, x, y, z, vt);
        }
        double div = 0.0;
        for (int i = 0; i < dim; ++i) {
          const double v1 = F(i, c.x, c.y, c.z), v2 = F(i, c.x + (i == 0), c.y + (i == 1), c.z + (i == 2));
          div += (1.0 / h[i]) * (v2 - v1);
        }
        const double cor = div / nn;
        for (int k = 0; k < nn; ++k) {
          const int d = tdir[k];
          int x = c.x, y = c.y, z = c.z;
          if (tsgn[k] == 1) { x += (d == 0); y += (d == 1); z += (d == 2); }
          S(d, x, y, z, F(d, x, y, z) - (2 * tsgn[k] - 1) * h[d] * cor);
        }
      }
    }
  }

  // SolveMatrix (:578-778) for the corner cell (x, y, z): corner_system on
  // its neighbourhood of slot `slot`, read through the grid.
  void corner_solve(int slot, int x, int y, int z, int s1, int s2, int s3, double res[3]) {
    auto U = [&](int c, int dx, int dy, int dz) { return speed(slot, c, x + dx, y + dy, z + dz); };
    corner_system(P.h, s1, s2, s3, U, res);
  }

  // With surface_bc set, Algorithms<Algorithm_Extrapolate_Boundary_Condition,
  // Algorithms_Extrapolate> as test/Jet3.cpp:541-543,636-639 chains them.
  void extrapolate(int slot) {
    if (P.surface_bc) surface_bc(slot, 3);
    auto ok = [&](int x, int y, int z) { return valid(x, y, z) && layer_of(x, y, z) != -1; };
    auto U = [&](int comp, int x, int y, int z) { return field(chunk(x, y, z), slot, comp)[off_of(x, y, z)]; };
    auto set = [&](int x, int y, int z, int comp, double val) {
      Chunk* ch = chunk(x, y, z);
      const int o = off_of(x, y, z);
      if (!(ch->flags[o] & (0x08u << comp))) field(ch, slot, comp)[o] = val;
    };
    std::vector<K3> cur;
    for_cells([&](Chunk* c, int t, int x, int y, int z) {
      if (c->layer[t] == 1) cur.push_back({x, y, z});
    });
    int lay = 1;
    while (true) {
      std::map<K3, bool> next;
      for (const K3& A : cur) {
        const int ax = A.x, ay = A.y, az = A.z;
        for (int i = 0; i < 3; ++i) {
          const int ex = (i == 0), ey = (i == 1), ez = (i == 2);
          // +1 direction
          const int bx = ax + ex, by = ay + ey, bz = az + ez;
          if (ok(bx, by, bz) && layer_of(bx, by, bz) >= lay) {
            if (layer_of(bx, by, bz) > lay) next[{bx, by, bz}] = true;
            int n = 0;
            double val = 0;
            for (int j = 0; j < 3; ++j)
              for (int s2 = -1; s2 <= 1; s2 += 2) {
                const int jx = s2 * (j == 0), jy = s2 * (j == 1), jz = s2 * (j == 2);
                if (s2 == -1 && j == i) {
                  if (ok(ax - ex, ay - ey, az - ez) && layer_of(ax - ex, ay - ey, az - ez) < lay) {
                    ++n;
                    val += U(i, ax, ay, az);
                  }
                } else if (s2 == 1 && j == i) {
                  if (ok(bx + ex, by + ey, bz + ez) && layer_of(bx + ex, by + ey, bz + ez) < lay) {
                    ++n;
                    val += U(i, bx + ex, by + ey, bz + ez);
                  }
                } else {
                  const int x3 = bx + jx, y3 = by + jy, z3 = bz + jz, x4 = ax + jx, y4 = ay + jy, z4 = az + jz;
                  if (ok(x3, y3, z3) && ok(x4, y4, z4) && (layer_of(x3, y3, z3) < lay || layer_of(x4, y4, z4) < lay)) {
                    ++n;
                    val += U(i, x3, y3, z3);
                  }
                }
              }
            set(bx, by, bz, i, val / n);
          }
          // -1 direction
          const int cx = ax - ex, cy = ay - ey, cz = az - ez;
          if (ok(cx, cy, cz)) {
            if (layer_of(cx, cy, cz) >= lay) {
              if (layer_of(cx, cy, cz) > lay) next[{cx, cy, cz}] = true;
              int n = 0;
              double val = 0;
              for (int j = 0; j < 3; ++j)
                for (int s2 = -1; s2 <= 1; s2 += 2) {
                  const int jx = s2 * (j == 0), jy = s2 * (j == 1), jz = s2 * (j == 2);
                  if (s2 == 1 && j == i) {
                    if (ok(ax + ex, ay + ey, az + ez) && layer_of(ax + ex, ay + ey, az + ez) < lay) {
                      ++n;
                      val += U(j, ax + ex, ay + ey, az + ez);
                    }
                  } else if (s2 == -1 && j == i) {
                    if (ok(cx - ex, cy - ey, cz - ez) && layer_of(cx - ex, cy - ey, cz - ez) < lay) {
                      ++n;
                      val += U(j, cx, cy, cz);
                    }
                  } else {
                    const int x3 = cx + jx, y3 = cy + jy, z3 = cz + jz, x4 = ax + jx, y4 = ay + jy, z4 = az + jz;
                    if (ok(x3, y3, z3) && ok(x4, y4, z4) && (layer_of(x3, y3, z3) < lay || layer_of(x4, y4, z4) < lay)) {
                      ++n;
                      val += U(i, x4, y4, z4);
                    }
                  }
                }
              set(ax, ay, az, i, val / n);
            }
          } else {
            int n = 0;
            double val = 0;
            for (int j = 0; j < 3; ++j)
              for (int s2 = -1; s2 <= 1; s2 += 2) {
                const int jx = s2 * (j == 0), jy = s2 * (j == 1), jz = s2 * (j == 2);
                if (s2 == 1 && j == i) {
                  if (ok(ax + ex, ay + ey, az + ez) && layer_of(ax + ex, ay + ey, az + ez) < lay) {
                    ++n;
                    val += U(j, ax + ex, ay + ey, az + ez);
                  }
                } else if (s2 == -1 && j == i) {
                } else {
                  const int x4 = ax + jx, y4 = ay + jy, z4 = az + jz;
                  if (ok(x4, y4, z4)) {
                    const int x3 = x4 - ex, y3 = y4 - ey, z3 = z4 - ez;
                    if (ok(x3, y3, z3) && (layer_of(x3, y3, z3) < lay || layer_of(x4, y4, z4) < lay)) {
                      ++n;
                      val += U(i, x4, y4, z4);
                    }
                  }
                }
              }
            set(ax, ay, az, i, val / n);
          }
        }
      }
      cur.clear();
      for (auto& kv : next) cur.push_back(kv.first);
      ++lay;
      if (cur.empty()) break;
    }
  }

  // Acceleration SetZero(true), then Algorithms<Gravity, Viscosity, Convection>
  // (TestJetDEau_Basic.cpp:365); rk4: no zeroing (the RK4 "all" driver
  // zeroes) and Jet3's order Gravity, Convection, Viscosity (:596-599)
  void forces(int slot, bool rk4 = false) {
    if (!rk4)
      for_cells([&](Chunk* c, int t, int, int, int) {
        for (int i = 0; i < 3; ++i) c->A[i][t] = 0.0;
      });
    // Policy_Gravity::ApplyGravity
    for (int i = 0; i < 3; ++i) {
      if (P.g[i] == 0.0) continue;
      for_cells([&](Chunk* c, int t, int kx, int ky, int kz) {
        bool need = is_fluid(c->flags[t] & 7u);
        if (!need && c->layer[t] != -1) {
          int n[3] = {kx, ky, kz};
          n[i] -= 1;
          need = valid(n[0], n[1], n[2]) && fluid(n[0], n[1], n[2]);
        }
        if (need && !(c->flags[t] & (0x08u << i))) c->A[i][t] = c->A[i][t] + P.g[i];
      });
    }
    auto viscosity = [&]() {
    // Algorithms_Viscosity
    if (P.nu != 0.0) {
      for_cells([&](Chunk* c, int t, int kx, int ky, int kz) {
        bool app = is_fluid(c->flags[t] & 7u);
        if (!app && c->layer[t] != -1)
          for (int i = 0; i < 3; ++i) {
            int n[3] = {kx, ky, kz};
            n[i] -= 1;
            if (valid(n[0], n[1], n[2]) && fluid(n[0], n[1], n[2])) app = true;
          }
        if (!app) return;
        double lap[3];
        for (int i = 0; i < 3; ++i) {
          double temp = 0;
          for (int j = 0; j < 3; ++j) {
            int p1[3] = {kx, ky, kz}, m1[3] = {kx, ky, kz};
            p1[j] += 1;
            m1[j] -= 1;
            temp -= 2 * c->S[slot][i][t] * std::pow(ih[j], 2);
            temp += speed(slot, i, p1[0], p1[1], p1[2]) * std::pow(ih[j], 2);
            temp += speed(slot, i, m1[0], m1[1], m1[2]) * std::pow(ih[j], 2);
          }
          lap[i] = temp;
        }
        for (int i = 0; i < 3; ++i)
          if (!(c->flags[t] & (0x08u << i))) c->A[i][t] = c->A[i][t] + lap[i] * P.nu;
      });
    }
    };
    auto convection = [&]() {
    // Algorithms_Convection with Policy_Upwind_1_Order
    for_cells([&](Chunk* c, int t, int kx, int ky, int kz) {
      int K[3] = {kx, ky, kz};
      for (int i = 0; i < 3; ++i) {
        bool app = is_fluid(c->flags[t] & 7u);
        if (!app && c->layer[t] != -1) {
          int n[3] = {kx, ky, kz};
          n[i] -= 1;
          app = valid(n[0], n[1], n[2]) && fluid(n[0], n[1], n[2]);
        }
        if (!app) continue;
        double temp = 0;
        for (int j = 0; j < 3; ++j) {
          double U;
          if (i == j) {
            U = c->S[slot][j][t];
          } else {
            int a[3] = {K[0], K[1], K[2]}, b[3] = {K[0], K[1], K[2]}, d[3] = {K[0], K[1], K[2]};
            a[j] += 1;
            b[i] -= 1;
            d[i] -= 1;
            d[j] += 1;
            U = 0.25 * (c->S[slot][j][t] + speed(slot, j, a[0], a[1], a[2]) + speed(slot, j, b[0], b[1], b[2]) +
                        speed(slot, j, d[0], d[1], d[2]));
          }
          double U1 = c->S[slot][i][t], U2;
          int n[3] = {K[0], K[1], K[2]};
          if (U > 0) {
            n[j] -= 1;
            U2 = speed(slot, i, n[0], n[1], n[2]);
            U *= (U1 - U2) * ih[j];
          } else {
            n[j] += 1;
            U2 = speed(slot, i, n[0], n[1], n[2]);
            U *= (U2 - U1) * ih[j];
          }
          temp -= U;
        }
        if (!(c->flags[t] & (0x08u << i))) c->A[i][t] = c->A[i][t] + temp;
      }
    });
    };
    if (rk4) {
      convection();
      viscosity();
    } else {
      viscosity();
      convection();
    }
  }

  // projection of field `slot` (3: the acceleration, Algorithms_Solve_Pressure_Force);
  // with proj_nu > 0 the viscous air pressures read speed slot `real`
  void project(int slot, double* maxdiv, int real = -1) {
    std::vector<int32_t> keys;
    std::vector<uint8_t> fl;
    std::vector<double> u, v, w, p, ur, vr, wr;
    std::vector<Chunk*> chs;
    const int rs = real < 0 ? slot : real;
    for (auto& kv : grid) {
      keys.push_back(kv.first.x);
      keys.push_back(kv.first.y);
      keys.push_back(kv.first.z);
      chs.push_back(kv.second.get());
    }
    int nb = (int)chs.size();
    fl.resize((size_t)nb * 512);
    u.resize((size_t)nb * 512);
    v.resize(u.size());
    w.resize(u.size());
    p.resize(u.size());
    for (int b = 0; b < nb; ++b) {
      memcpy(&fl[(size_t)b * 512], chs[b]->flags, 512);
      memcpy(&u[(size_t)b * 512], field(chs[b], slot, 0), 512 * 8);
      memcpy(&v[(size_t)b * 512], field(chs[b], slot, 1), 512 * 8);
      memcpy(&w[(size_t)b * 512], field(chs[b], slot, 2), 512 * 8);
    }
    if (nb > 0 && P.proj_nu > 0) {
      ur.resize(u.size());
      vr.resize(u.size());
      wr.resize(u.size());
      for (int b = 0; b < nb; ++b) {
        memcpy(&ur[(size_t)b * 512], field(chs[b], rs, 0), 512 * 8);
        memcpy(&vr[(size_t)b * 512], field(chs[b], rs, 1), 512 * 8);
        memcpy(&wr[(size_t)b * 512], field(chs[b], rs, 2), 512 * 8);
      }
      orc_project_dirichlet(nb, keys.data(), fl.data(), P.h, u.data(), v.data(), w.data(), ur.data(), vr.data(), wr.data(), P.proj_nu,
                            p.data(), 1e-14, maxdiv);
    } else if (nb > 0) {
      orc_project(nb, keys.data(), fl.data(), P.h, u.data(), v.data(), w.data(), p.data(), 1e-14, maxdiv);
    }
    for (int b = 0; b < nb; ++b) {
      memcpy(field(chs[b], slot, 0), &u[(size_t)b * 512], 512 * 8);
      memcpy(field(chs[b], slot, 1), &v[(size_t)b * 512], 512 * 8);
      memcpy(field(chs[b], slot, 2), &w[(size_t)b * 512], 512 * 8);
      memcpy(chs[b]->P, &p[(size_t)b * 512], 512 * 8);
    }
  }

  void advance(double dt, double* maxdiv) {
    bool tvd = P.integrator == 1;
    forces(0);
    for_cells([&](Chunk* c, int t, int, int, int) {
      for (int i = 0; i < 3; ++i)
        c->S[1][i][t] = tvd ? c->S[0][i][t] + c->A[i][t] * dt : c->S[0][i][t] + c->A[i][t] * dt * 0.5;
      for (int i = 0; i < 3; ++i) c->A[i][t] = 0.0;
    });
    project(1, maxdiv);
    forces(1);
    for_cells([&](Chunk* c, int t, int, int, int) {
      for (int i = 0; i < 3; ++i) {
        if (c->flags[t] & (0x08u << i)) continue;
        c->S[0][i][t] = tvd ? 0.5 * c->S[0][i][t] + 0.5 * c->S[1][i][t] + 0.5 * c->A[i][t] * dt : c->S[0][i][t] + c->A[i][t] * dt;
      }
    });
    project(0, maxdiv);
  }

  // Policy_Speed_Interpolation_Linear_Symmetric::Get_Speed, recursive as in the reference
  double impl(const double d2[3], int i, int kx, int ky, int kz, int k, int slot) {
    int n[3] = {kx, ky, kz};
    n[i] += 1;
    double w0 = (i != k) ? (1 - d2[i]) : (0.5 - d2[i]);
    double w1 = (i != k) ? d2[i] : (0.5 + d2[i]);
    if (i < 2) return w0 * impl(d2, i + 1, kx, ky, kz, k, slot) + w1 * impl(d2, i + 1, n[0], n[1], n[2], k, slot);
    return w0 * speed(slot, k, kx, ky, kz) + w1 * speed(slot, k, n[0], n[1], n[2]);
  }
  std::array<double, 3> get_speed(const std::array<double, 3>& pos, int slot = 0) {
    double d[3];
    int key0[3];
    for (int i = 0; i < 3; ++i) {
      d[i] = pos[i] * ih[i];
      key0[i] = (int)lround(d[i]);
      d[i] -= key0[i];
    }
    std::array<double, 3> ret;
    for (int k = 0; k < 3; ++k) {
      int key[3];
      double d2[3];
      for (int j = 0; j < 3; ++j) {
        if (j != k && d[j] < 0) {
          key[j] = key0[j] - 1;
          d2[j] = 1 + d[j];
        } else {
          key[j] = key0[j];
          d2[j] = d[j];
        }
      }
      ret[k] = impl(d2, 0, key[0], key[1], key[2], k, slot);
    }
    return ret;
  }

  // Algorithms_RungeKutta_all::Do (:35-156) with Policy::Do of test/Jet3.cpp:
  // per stage the cell-type chain of m_alg_init2 (keys of slot 0, then 2),
  // extrapolation of the stage's speed slot, gravity / convection /
  // viscosity added to the acceleration, its projection (air pressures from
  // the stage's speed) and extrapolation, then the grid and particle loops in
  // the reference's order and operand order.
  void rk4all(double dt, double* maxdiv) {
    auto zero = [&]() {
      for_cells([&](Chunk* c, int t, int, int, int) {
        for (int i = 0; i < 3; ++i) c->A[i][t] = 0.0;
      });
    };
    zero();
    parts1 = parts;
    parts2 = parts;
    for (int stage = 1; stage <= 4; ++stage) {
      const int ks = stage == 1 ? 0 : 2;
      topology(false, ks);
      extrapolate(ks);
      forces(ks, true);
      project(3, maxdiv, ks);
      extrapolate(3);
      for_cells([&](Chunk* c, int t, int, int, int) {
        for (int i = 0; i < 3; ++i) {
          const double a = c->A[i][t];
          switch (stage) {
            case 1:
              c->S[2][i][t] = c->S[0][i][t] + a * dt * 0.5;
              c->S[1][i][t] = c->S[0][i][t] + a * dt * 1. / 6.;
              break;
            case 2:
              c->S[2][i][t] = c->S[0][i][t] + a * dt * 0.5;
              c->S[1][i][t] = c->S[1][i][t] + a * dt * 1. / 3.;
              break;
            case 3:
              c->S[2][i][t] = c->S[0][i][t] + a * dt;
              c->S[1][i][t] = c->S[1][i][t] + a * dt * 1. / 3.;
              break;
            default:
              if (!(c->flags[t] & (0x08u << i))) c->S[0][i][t] = c->S[1][i][t] + a * dt * 1. / 6.;
              break;
          }
        }
      });
      for (size_t q = 0; q < parts.size(); ++q) {
        std::array<double, 3>&p0 = parts[q], &p1 = parts1[q], &p2 = parts2[q];
        if (stage == 1) {
          const std::array<double, 3> v = get_speed(p0, ks);
          for (int k = 0; k < 3; ++k) {
            p2[k] = p0[k] + dt * 0.5 * v[k];
            p1[k] = p0[k] + dt * 1.0 / 6.0 * v[k];
          }
        } else if (stage == 4) {
          const std::array<double, 3> v = get_speed(p2, ks);
          for (int k = 0; k < 3; ++k) p0[k] = p1[k] + dt * 1.0 / 6.0 * v[k];
        } else {
          std::array<double, 3> v = get_speed(p2, ks);
          for (int k = 0; k < 3; ++k) p2[k] = p0[k] + (stage == 2 ? dt * 0.5 : dt) * v[k];
          v = get_speed(p2, ks);
          for (int k = 0; k < 3; ++k) p1[k] = p1[k] + dt * 1.0 / 3.0 * v[k];
        }
      }
      if (stage < 4) zero();
    }
  }

  // The seeded inflow pulsation (no counterpart in the reference; it stands
  // where test/Jet.cpp:188-256 perturbs its inflow after the time step): the const
  // faces of existing Inflow cells get inflow_speed * (1 + amp sin(2 pi n /
  // period + phi)), n = tick, phi = 2 pi U(seed).
  void pulse_inflow() {
    if (P.inflow_amp == 0.0) return;
    const double phi = 2.0 * M_PI * uni(P.seed, 0x1f10, 0, 0);
    const double f = 1.0 + P.inflow_amp * std::sin(2.0 * M_PI * (double)tick / (double)P.inflow_period + phi);
    for (auto& kv : grid) {
      Chunk& ch = *kv.second;
      for (int t = 0; t < 512; ++t) {
        const unsigned fl = ch.flags[t];
        if (!(fl & 0x80) || (fl & 7u) != 3u) continue;
        for (int i = 0; i < 3; ++i)
          if (fl & (0x08u << i)) ch.S[0][i][t] = P.inflow_speed[i] * f;
      }
    }
  }
  // particles leaving [lo, hi) are deleted (slot 0)
  void box_delete() {
    if (P.lo[0] == P.hi[0] && P.lo[1] == P.hi[1] && P.lo[2] == P.hi[2]) return;
    std::vector<std::array<double, 3>> out;
    for (auto& r : parts)
      if (r[0] >= P.lo[0] && r[0] < P.hi[0] && r[1] >= P.lo[1] && r[1] < P.hi[1] && r[2] >= P.lo[2] && r[2] < P.hi[2]) out.push_back(r);
    parts.swap(out);
  }

  void move(double dt) {
    std::vector<std::array<double, 3>> out;
    bool box = !(P.lo[0] == P.hi[0] && P.lo[1] == P.hi[1] && P.lo[2] == P.hi[2]);
    for (auto& pos : parts) {
      std::array<double, 3> res = pos, k, q, v;
      if (P.particle_rk == 4) {
        v = get_speed(pos);
        for (int j = 0; j < 3; ++j) { k[j] = dt * v[j]; res[j] += (1. / 6.) * k[j]; q[j] = pos[j] + k[j] * 0.5; }
        v = get_speed(q);
        for (int j = 0; j < 3; ++j) { k[j] = dt * v[j]; res[j] += (1. / 3.) * k[j]; q[j] = pos[j] + k[j] * 0.5; }
        v = get_speed(q);
        for (int j = 0; j < 3; ++j) { k[j] = dt * v[j]; res[j] += (1. / 3.) * k[j]; q[j] = pos[j] + k[j]; }
        v = get_speed(q);
        for (int j = 0; j < 3; ++j) { k[j] = dt * v[j]; res[j] += (1. / 6.) * k[j]; }
      } else {
        v = get_speed(pos);
        for (int j = 0; j < 3; ++j) { k[j] = dt * v[j]; q[j] = pos[j] + k[j] * 0.5; }
        v = get_speed(q);
        for (int j = 0; j < 3; ++j) res[j] = pos[j] + dt * v[j];
      }
      if (box && !(res[0] >= P.lo[0] && res[0] < P.hi[0] && res[1] >= P.lo[1] && res[1] < P.hi[1] && res[2] >= P.lo[2] && res[2] < P.hi[2]))
        continue;
      out.push_back(res);
    }
    parts.swap(out);
  }
};

}  // namespace

extern "C" {

// prm: h[3], g[3], nu, cfl, dt_min, dt_max, lo[3], hi[3] (16 doubles);
// iprm: layer_max, integrator, particle_rk, ppc, depth_interior, depth_remove,
// floor_axis, fill_cube_n, inflow_period, surface_bc; prm[16] = floor_value, prm[17] =
// projection viscosity, prm[18..20] = inflow_speed, prm[21] = inflow_amp; seed.
void* orc_sim_create(const double* prm, const int* iprm, unsigned long long seed) {
  Sim* s = new Sim();
  for (int i = 0; i < 3; ++i) {
    s->P.h[i] = prm[i];
    s->P.g[i] = prm[3 + i];
    s->P.lo[i] = prm[10 + i];
    s->P.hi[i] = prm[13 + i];
    s->ih[i] = 1.0 / prm[i];
  }
  s->P.nu = prm[6];
  s->P.cfl = prm[7];
  s->P.dt_min = prm[8];
  s->P.dt_max = prm[9];
  s->P.layer_max = iprm[0];
  s->P.integrator = iprm[1];
  s->P.particle_rk = iprm[2];
  s->P.ppc = iprm[3];
  s->P.depth_interior = iprm[4];
  s->P.depth_remove = iprm[5];
  s->P.floor_axis = iprm[6];
  s->P.floor_value = prm[16];
  s->P.proj_nu = prm[17];
  s->P.cube_n = iprm[7];
  s->P.inflow_period = iprm[8];
  s->P.surface_bc = iprm[9];
  for (int i = 0; i < 3; ++i) s->P.inflow_speed[i] = prm[18 + i];
  s->P.inflow_amp = prm[21];
  s->P.seed = seed;
  return s;
}

void orc_sim_destroy(void* p) { delete (Sim*)p; }

void orc_sim_upload(void* p, int nb, const int32_t* keys, const uint8_t* flags, const int8_t* layer, const double* u, const double* v,
                    const double* w, int np, const double* pxyz) {
  Sim* s = (Sim*)p;
  s->grid.clear();
  for (int b = 0; b < nb; ++b) {
    auto c = std::make_unique<Chunk>();
    memcpy(c->flags, &flags[(size_t)b * 512], 512);
    if (layer) memcpy(c->layer, &layer[(size_t)b * 512], 512);
    const double* src[3] = {u, v, w};
    for (int i = 0; i < 3; ++i)
      if (src[i]) {
        memcpy(c->S[0][i], &src[i][(size_t)b * 512], 512 * 8);
        memcpy(c->S[1][i], &src[i][(size_t)b * 512], 512 * 8);
        memcpy(c->S[2][i], &src[i][(size_t)b * 512], 512 * 8);
      }
    s->grid.emplace(K3{keys[3 * b], keys[3 * b + 1], keys[3 * b + 2]}, std::move(c));
  }
  s->parts.clear();
  for (int k = 0; k < np; ++k) s->parts.push_back({pxyz[3 * k], pxyz[3 * k + 1], pxyz[3 * k + 2]});
}

unsigned long long orc_sim_get_tick(void* p) { return ((Sim*)p)->tick; }
void orc_sim_set_tick(void* p, unsigned long long t) { ((Sim*)p)->tick = t; }

void orc_sim_counts(void* p, int* nb, int* np) {
  Sim* s = (Sim*)p;
  *nb = (int)s->grid.size();
  *np = (int)s->parts.size();
}

// Bricks in ascending (z, y, x) chunk-key order, like the library.
void orc_sim_download(void* p, int32_t* keys, uint8_t* flags, int8_t* layer, double* u, double* v, double* w, double* ax, double* ay,
                      double* az, double* pr, double* pxyz) {
  Sim* s = (Sim*)p;
  int b = 0;
  for (auto& kv : s->grid) {
    Chunk* c = kv.second.get();
    if (keys) {
      keys[3 * b] = kv.first.x;
      keys[3 * b + 1] = kv.first.y;
      keys[3 * b + 2] = kv.first.z;
    }
    if (flags) memcpy(&flags[(size_t)b * 512], c->flags, 512);
    if (layer) memcpy(&layer[(size_t)b * 512], c->layer, 512);
    double* dst[7] = {u, v, w, ax, ay, az, pr};
    const double* src[7] = {c->S[0][0], c->S[0][1], c->S[0][2], c->A[0], c->A[1], c->A[2], c->P};
    for (int i = 0; i < 7; ++i)
      if (dst[i]) memcpy(&dst[i][(size_t)b * 512], src[i], 512 * 8);
    ++b;
  }
  if (pxyz)
    for (size_t k = 0; k < s->parts.size(); ++k)
      for (int i = 0; i < 3; ++i) pxyz[3 * k + i] = s->parts[k][i];
}

// per-cell depth in download order
void orc_sim_get_depth(void* p, int16_t* depth) {
  size_t b = 0;
  for (auto& kv : ((Sim*)p)->grid) memcpy(&depth[512 * b++], kv.second->depth, sizeof(kv.second->depth));
}
void orc_sim_set_depth(void* p, const int16_t* depth) {
  size_t b = 0;
  for (auto& kv : ((Sim*)p)->grid) memcpy(kv.second->depth, &depth[512 * b++], sizeof(kv.second->depth));
}
void orc_sim_topology(void* p) { ((Sim*)p)->topology(); }
double orc_sim_compute_dt(void* p) { return ((Sim*)p)->compute_dt(); }
void orc_sim_forces(void* p, int slot) { ((Sim*)p)->forces(slot); }
void orc_sim_advance(void* p, double dt, double* maxdiv) { ((Sim*)p)->advance(dt, maxdiv); }
void orc_sim_move(void* p, double dt) { ((Sim*)p)->move(dt); }
void orc_sim_extrapolate(void* p, int slot) { ((Sim*)p)->extrapolate(slot); }
void orc_sim_surface_bc(void* p, int slot, int dim) { ((Sim*)p)->surface_bc(slot, dim); }
// The corner system alone on a 3x3x3 neighbourhood: nb81[c * 27 + (dx + 1) +
// 3 (dy + 1) + 9 (dz + 1)] is face component c of the cell at (dx, dy, dz).
void orc_corner_solve(const double* h, int s1, int s2, int s3, const double* nb81, double* res) {
  auto U = [&](int c, int dx, int dy, int dz) { return nb81[c * 27 + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)]; };
  corner_system(h, s1, s2, s3, U, res);
}
void orc_sim_interpolate(void* p, int n, const double* pos, double* vel) {
  Sim* s = (Sim*)p;
  for (int k = 0; k < n; ++k) {
    auto v = s->get_speed({pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]});
    for (int i = 0; i < 3; ++i) vel[3 * k + i] = v[i];
  }
}
// One step in the library's order: topology, dt, advance, move.
double orc_sim_step(void* p, double* maxdiv) {
  Sim* s = (Sim*)p;
  s->topology();
  double dt = s->compute_dt();
  s->pulse_inflow();
  if (s->P.integrator == 2) {
    // test/Jet3.cpp's init chain ends with the extrapolation of speed slot 0
    s->extrapolate(0);
    s->rk4all(dt, maxdiv);
    s->box_delete();
  } else {
    s->advance(dt, maxdiv);
    s->move(dt);
  }
  return dt;
}

}  // extern "C"
