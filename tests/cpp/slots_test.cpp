// slots_test.cpp -- drives the two template slots of include/mac_amd_slots.hpp
// on a stand-in grid that offers exactly the calls the reference's
// Algorithms_Solve_Pressure makes on its grid (iterator key()/data(),
// CellType_GetRef().GetIsInDomain(), Speed_GetRef().Get/Set/Get_Const,
// Pressure_GetRef().Pressure_Set, m_h.GetRef_Inv().Get).
//
//   slots_test IN OUT
// IN (little-endian): int32 nb; int32 keys[3nb]; u8 flags[512nb];
//   f64 u,v,w[512nb]; f64 h[3]; int32 n; int32 off[n+1]; int32 nnz;
//   int32 idx[nnz]; f64 val[nnz]; f64 b[n]; f64 viscosity; int32 force;
//   when force != 0: f64 ax,ay,az[512nb]
// OUT: f64 p,u,v,w[512nb] (slot B, brick layout: the projected field, i.e.
//   the acceleration when force != 0), f64 x[n] (slot A), int32 iters
// force != 0 runs slot B the way Algorithms_Solve_Pressure_Force does
// (src/Algorithms_Solve_Pressure_Force.h:16-24): Speed_GetRef() is
// redirected to the acceleration around the three calls, while
// Speed_GetRef(-1, false) still returns the speed (Data_Chunk_Speed.h:44-55).
// Without a GPU both slots print the library's error and the program still
// exits 0, as the reference's solver policy does.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "mac_amd_slots.hpp"

namespace mock {

struct Key {
  int k[3];
  int Get(int i) const { return k[i - 1]; }
};
struct Vec {
  double v[3];
  double Get(int i) const { return v[i - 1]; }
};
struct Speed {
  double s[3];
  bool c[3];
  double Get(int i) const { return s[i - 1]; }
  void Set(int i, double x) {  // non-forced Set: const faces keep their value
    if (!c[i - 1]) s[i - 1] = x;
  }
  bool Get_Const(int i) const { return c[i - 1]; }
};
struct CellType {
  int mat;
  bool GetIsInDomain() const { return mat == 1 || mat == 3 || mat == 4; }
  bool GetIsSolid() const { return mat == 2; }
  bool GetIs0_Pres() const { return mat == 5; }
};
struct PressureValue {
  double v;
  explicit PressureValue(double x) : v(x) {}
};
struct Pressure {
  typedef PressureValue type_pressure;
  double p = -1.0;
  void Pressure_Set(PressureValue x) { p = x.v; }
};
struct Cell {
  Key key;
  CellType ct;
  Speed sp;
  Speed acc;  // the acceleration slot, same const bits
  Pressure pr;
};
bool accel_instead_speed = false;  // Data_Chunk_Speed::baccel_instead_speed
struct Handle {  // what the reference's iterator data() returns: a handle by value
  Cell* c;
  CellType& CellType_GetRef() { return c->ct; }
  Speed& Speed_GetRef(int i = -1, bool b = true) {
    (void)i;
    return (accel_instead_speed && b) ? c->acc : c->sp;
  }
  Pressure& Pressure_GetRef() { return c->pr; }
};
struct Spacing {
  Vec inv;
  const Vec& GetRef_Inv() const { return inv; }
};
struct Grid {
  std::vector<Cell> cells;
  Spacing m_h;
  struct iterator {
    std::vector<Cell>* v;
    size_t i;
    iterator& operator++() {
      ++i;
      return *this;
    }
    bool operator!=(const iterator& o) const { return i != o.i; }
    Key key() const { return (*v)[i].key; }
    Handle data() const { return Handle{&(*v)[i]}; }
  };
  iterator begin() { return iterator{&cells, 0}; }
  iterator end() { return iterator{&cells, cells.size()}; }
};
struct Data {
  Grid* g;
  Grid& GetGridData() { return *g; }
};
struct DataRef {
  Data m_data;
};
struct NoPolicy {};

}  // namespace mock

template <class T>
static bool rd(std::ifstream& f, T* p, size_t n) {
  return (bool)f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: slots_test IN OUT\n";
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  int32_t nb = 0;
  if (!rd(in, &nb, 1) || nb <= 0) return 2;
  std::vector<int32_t> keys(3 * nb);
  std::vector<uint8_t> flags(512 * (size_t)nb);
  std::vector<double> u(512 * (size_t)nb), v(u.size()), w(u.size());
  double h[3];
  int32_t n = 0, nnz = 0;
  if (!rd(in, keys.data(), keys.size()) || !rd(in, flags.data(), flags.size()) || !rd(in, u.data(), u.size()) ||
      !rd(in, v.data(), v.size()) || !rd(in, w.data(), w.size()) || !rd(in, h, 3) || !rd(in, &n, 1))
    return 2;
  std::vector<int32_t> off(n + 1);
  if (!rd(in, off.data(), off.size()) || !rd(in, &nnz, 1)) return 2;
  std::vector<int32_t> idx(nnz);
  std::vector<double> val(nnz), b(n), x(n, 0.0);
  if (!rd(in, idx.data(), idx.size()) || !rd(in, val.data(), val.size()) || !rd(in, b.data(), b.size())) return 2;
  double viscosity = 0.0;
  int32_t force = 0;
  if (!rd(in, &viscosity, 1) || !rd(in, &force, 1)) return 2;
  std::vector<double> ax(u.size(), 0.0), ay(u.size(), 0.0), az(u.size(), 0.0);
  if (force && (!rd(in, ax.data(), ax.size()) || !rd(in, ay.data(), ay.size()) || !rd(in, az.data(), az.size())))
    return 2;

  // the grid: every existing cell, in brick order (any order works)
  mock::Grid grid;
  for (int i = 0; i < 3; ++i) grid.m_h.inv.v[i] = 1.0 / h[i];
  for (int bk = 0; bk < nb; ++bk)
    for (int t = 0; t < 512; ++t) {
      const size_t s = (size_t)bk * 512 + t;
      const uint8_t f = flags[s];
      if (!(f & MAC_FLAG_EXISTS)) continue;
      mock::Cell c;
      c.key = {{8 * keys[3 * bk] + (t & 7), 8 * keys[3 * bk + 1] + ((t >> 3) & 7), 8 * keys[3 * bk + 2] + (t >> 6)}};
      c.ct.mat = f & MAC_FLAG_MAT_MASK;
      c.sp = {{u[s], v[s], w[s]}, {(f & MAC_FLAG_CONST_X) != 0, (f & MAC_FLAG_CONST_Y) != 0, (f & MAC_FLAG_CONST_Z) != 0}};
      c.acc = {{ax[s], ay[s], az[s]}, {c.sp.c[0], c.sp.c[1], c.sp.c[2]}};
      grid.cells.push_back(c);
    }
  mock::DataRef data{mock::Data{&grid}};

  // slot B
  int32_t iters = -1;
  {
    Algorithms_Solve_Pressure_HIP<mock::DataRef, mock::NoPolicy> alg(data, mock::NoPolicy(), MAC_FP64, 1e-12, viscosity);
    mock::accel_instead_speed = force != 0;
    alg.Init_Iteration();
    alg.Divergence_Projection();
    alg.End_Iteration();
    mock::accel_instead_speed = false;
    iters = alg.last_stats.iters;
  }
  // slot A on the reference-assembled system
  {
    Policy_Solve_Linear_HIP<double> pol(1e-13);
    pol.Solve_Linear_FactorizeMatrice(n, off.data(), idx.data(), val.data());
    pol.Solve_Linear(n, off.data(), idx.data(), val.data(), b.data(), x.data());
    pol.Solve_Linear_Clean();
  }

  // back to brick layout
  std::vector<double> p(u.size(), 0.0), uo(u.size(), 0.0), vo(u.size(), 0.0), wo(u.size(), 0.0);
  std::map<std::tuple<int, int, int>, int> bidx;
  for (int bk = 0; bk < nb; ++bk) bidx[std::make_tuple(keys[3 * bk], keys[3 * bk + 1], keys[3 * bk + 2])] = bk;
  for (const mock::Cell& c : grid.cells) {
    using mac_amd_detail::chunk_of;
    using mac_amd_detail::offset_of;
    const int bk = bidx[std::make_tuple(chunk_of(c.key.k[0]), chunk_of(c.key.k[1]), chunk_of(c.key.k[2]))];
    const size_t s = (size_t)bk * 512 + offset_of(c.key.k[0]) + 8 * offset_of(c.key.k[1]) + 64 * offset_of(c.key.k[2]);
    const mock::Speed& out_field = force ? c.acc : c.sp;
    p[s] = c.pr.p;
    uo[s] = out_field.s[0];
    vo[s] = out_field.s[1];
    wo[s] = out_field.s[2];
  }
  std::ofstream out(argv[2], std::ios::binary);
  for (auto* a : {&p, &uo, &vo, &wo}) out.write(reinterpret_cast<const char*>(a->data()), sizeof(double) * a->size());
  out.write(reinterpret_cast<const char*>(x.data()), sizeof(double) * x.size());
  out.write(reinterpret_cast<const char*>(&iters), sizeof(iters));
  std::cout << "slots_test: " << grid.cells.size() << " cells, " << nb << " bricks, n = " << n << ", iters " << iters << std::endl;
  return 0;
}
