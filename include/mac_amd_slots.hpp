// mac_amd_slots.hpp -- the reference's two pressure-projection template slots
// on top of the C-ABI of mac_amd.h (header-only C++; link libmac_amd.so).
//
//   Policy_Solve_Linear_HIP<Data>
//       drop-in for Policy_Solve_Linear_Umfpack<Data>
//       (src/Policy_Solve_Linear_Umfpack.h:7-71): same three calls, same
//       arrays (caller-owned CSR, b in, res out), n == 0 a no-op, errors
//       printed and execution continued (:15-21).
//
//   Algorithms_Solve_Pressure_HIP<DataType, Policy>
//       drop-in for Algorithms_Solve_Pressure<DataType, Policy>
//       (src/Algorithms_Solve_Pressure.h:11-224): constructor (data, policy),
//       Init_Iteration / Divergence_Projection / End_Iteration.  Init_Iteration
//       gathers the grid's cells into 8^3 bricks (keys, flag bytes) and
//       uploads them; Divergence_Projection uploads the speeds, runs the
//       device projection and writes pressure (0 off the domain, :184-197)
//       and corrected speeds back through the grid's own Set (which skips
//       const faces, :198-207).  The grid is only read through the calls the
//       reference's algorithm itself makes: iterator key()/data(),
//       CellType_GetRef().GetIsInDomain(), Speed_GetRef().Get/Set/Get_Const,
//       Pressure_GetRef().Pressure_Set, m_h.GetRef_Inv().Get(i).  A cell type
//       with GetIsSolid() marks solid cells (they only shape the coarse
//       multigrid levels) and one with GetIs0_Pres() marks _0_Pres cells
//       (not air for the viscous air pressures); without them every
//       non-domain cell is air.
#ifndef MAC_AMD_SLOTS_HPP
#define MAC_AMD_SLOTS_HPP

#include <cstdint>
#include <iostream>
#include <map>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "mac_amd.h"

namespace mac_amd_detail {

inline int chunk_of(int k) { return k >= 0 ? k / 8 : (k + 1) / 8 - 1; }  // ChunkHashTable.tpp:181-196
inline int offset_of(int k) { return k - 8 * chunk_of(k); }

template <class C, class = void>
struct has_solid : std::false_type {};
template <class C>
struct has_solid<C, decltype((void)std::declval<C&>().GetIsSolid())> : std::true_type {};

template <class C>
bool is_solid(C& c, std::true_type) { return c.GetIsSolid(); }
template <class C>
bool is_solid(C&, std::false_type) { return false; }

// GetIs0_Pres (Data_CellType_Implemented_0_Pres.h:84): a _0_Pres cell is
// neither in the domain nor air, so the viscous variant gives it no air
// pressure (Policy_Dirichlet_Boundary.h:84-109 tests GetIsAir)
template <class C, class = void>
struct has_0_pres : std::false_type {};
template <class C>
struct has_0_pres<C, decltype((void)std::declval<C&>().GetIs0_Pres())> : std::true_type {};

template <class C>
bool is_0_pres(C& c, std::true_type) { return c.GetIs0_Pres(); }
template <class C>
bool is_0_pres(C&, std::false_type) { return false; }

// Pressure_Set with the pressure object's own value type when it names one
// (the reference wraps the value: type_pressure(res[i]), Algorithms_Solve_Pressure.h:189)
template <class P, class = void>
struct has_value_type : std::false_type {};
template <class P>
struct has_value_type<P, decltype((void)std::declval<typename P::type_pressure>())> : std::true_type {};

template <class P>
void pressure_set(P&& pr, double v, std::true_type) {
  pr.Pressure_Set(typename std::decay<P>::type::type_pressure(v));
}
template <class P>
void pressure_set(P&& pr, double v, std::false_type) {
  pr.Pressure_Set(v);
}

// Speed_GetRef(-1, false): the speed slot itself even while
// Data_Chunk_Speed::baccel_instead_speed redirects Speed_GetRef() to the
// acceleration (Data_Chunk_Speed.h:44-55, Algorithms_Solve_Pressure_Force.h:16-24);
// a cell type without that overload has only the one field.
template <class D, class = void>
struct has_real_speed : std::false_type {};
template <class D>
struct has_real_speed<D, decltype((void)std::declval<D&>().Speed_GetRef(-1, false))> : std::true_type {};

template <class D>
auto real_speed(D& d, std::true_type) -> decltype(d.Speed_GetRef(-1, false)) { return d.Speed_GetRef(-1, false); }
template <class D>
auto real_speed(D& d, std::false_type) -> decltype(d.Speed_GetRef()) { return d.Speed_GetRef(); }

inline void report(int status, const char* message) {
  if (status != MAC_OK) std::cout << "mac_amd error or warning :" << status << " in " << message << ": " << mac_last_error() << std::endl;
}

}  // namespace mac_amd_detail

// ------------------------------------------------------------- slot A ----
template <typename Data>
class Policy_Solve_Linear_HIP {
  static_assert(std::is_same<Data, double>::value, "the device solver works in double");
  mac_ctx* m_ctx = nullptr;
  double m_rtol;

 public:
  typedef Policy_Solve_Linear_HIP<Data> type_linear_solver;
  explicit Policy_Solve_Linear_HIP(double rtol = 1e-13) : m_rtol(rtol) {}
  // policies are copied into the algorithms that mix them in: each copy owns
  // its own context
  Policy_Solve_Linear_HIP(const Policy_Solve_Linear_HIP& o) : m_ctx(nullptr), m_rtol(o.m_rtol) {}
  Policy_Solve_Linear_HIP& operator=(const Policy_Solve_Linear_HIP& o) {
    m_rtol = o.m_rtol;
    return *this;
  }
  ~Policy_Solve_Linear_HIP() {
    if (m_ctx) mac_destroy(m_ctx);
  }
  void Solve_Linear_FactorizeMatrice(int n, int* offset, int* indice, Data* value) {
    (void)offset;
    (void)indice;
    (void)value;
    if (n != 0 && !m_ctx) mac_amd_detail::report(mac_create(&m_ctx, nullptr), " create");
  }
  void Solve_Linear_Clean() {}
  void Solve_Linear(int n, int* offset, int* indice, Data* value, Data* b, Data* res) {
    if (n == 0) return;
    if (!m_ctx) mac_amd_detail::report(mac_create(&m_ctx, nullptr), " create");
    if (!m_ctx) return;
    mac_amd_detail::report(mac_solve_csr(m_ctx, n, offset, indice, value, b, res, m_rtol), " solve");
  }
};

// ------------------------------------------------------------- slot B ----
template <typename DataType, typename Policy>
class Algorithms_Solve_Pressure_HIP : public Policy {
  typedef typename std::remove_reference<decltype(std::declval<DataType&>().m_data.GetGridData())>::type type_grid;
  typedef decltype(std::declval<type_grid&>().begin()) iterator;
  type_grid& m_grid;
  mac_params m_prm;
  mac_ctx* m_ctx = nullptr;
  std::vector<int32_t> m_keys;            // 3 per brick
  std::vector<uint8_t> m_flags;           // 512 per brick
  std::vector<long long> m_cell_slot;     // grid iteration order -> brick slot (b * 512 + t)
  std::vector<double> m_u, m_v, m_w, m_p, m_ur, m_vr, m_wr;

 public:
  mac_stats last_stats{};

  // viscosity > 0: the viscous variant, Algorithms_Solve_Pressure_Dirichlet
  // (pass the grid's m_viscosity).  Its air pressures are formed from the real
  // speed, Speed_GetRef(-1, false) (Policy_Dirichlet_Boundary.h:100-102), which
  // Divergence_Projection uploads beside the projected field, so the adapter
  // also works under Algorithms_Solve_Pressure_Force (test/Jet3.cpp:626,633),
  // where the projected field is the acceleration.
  Algorithms_Solve_Pressure_HIP(DataType data, const Policy& pol, int precision = MAC_FP64, double rtol = 1e-8,
                                double viscosity = 0.0)
      : Policy(pol), m_grid(data.m_data.GetGridData()) {
    mac_default_params(&m_prm);
    for (int i = 0; i < 3; ++i) m_prm.h[i] = 1.0 / m_grid.m_h.GetRef_Inv().Get(i + 1);
    m_prm.rtol = rtol;
    m_prm.precision = precision;
    m_prm.viscosity = viscosity;
  }
  Algorithms_Solve_Pressure_HIP(const Algorithms_Solve_Pressure_HIP&) = delete;
  Algorithms_Solve_Pressure_HIP& operator=(const Algorithms_Solve_Pressure_HIP&) = delete;
  ~Algorithms_Solve_Pressure_HIP() {
    if (m_ctx) mac_destroy(m_ctx);
  }

  // Bricks and flag bytes of the current cells (the reference assembles its
  // CSR here, Algorithms_Solve_Pressure.h:44-157).
  void Init_Iteration() {
    using mac_amd_detail::chunk_of;
    using mac_amd_detail::offset_of;
    std::map<std::tuple<int, int, int>, int> bricks;  // (z, y, x) order
    std::vector<std::tuple<int, int, int>> cells;
    for (iterator it = m_grid.begin(); it != m_grid.end(); ++it) {
      const auto k = it.key();
      cells.emplace_back(k.Get(1), k.Get(2), k.Get(3));
      bricks.emplace(std::make_tuple(chunk_of(k.Get(3)), chunk_of(k.Get(2)), chunk_of(k.Get(1))), 0);
    }
    int nb = 0;
    m_keys.clear();
    for (auto& kv : bricks) {
      kv.second = nb++;
      m_keys.push_back(std::get<2>(kv.first));
      m_keys.push_back(std::get<1>(kv.first));
      m_keys.push_back(std::get<0>(kv.first));
    }
    m_flags.assign((size_t)nb * 512, 0);
    m_cell_slot.clear();
    size_t ci = 0;
    for (iterator it = m_grid.begin(); it != m_grid.end(); ++it, ++ci) {
      const int x = std::get<0>(cells[ci]), y = std::get<1>(cells[ci]), z = std::get<2>(cells[ci]);
      const int b = bricks[std::make_tuple(chunk_of(z), chunk_of(y), chunk_of(x))];
      const long long slot = (long long)b * 512 + offset_of(x) + 8 * offset_of(y) + 64 * offset_of(z);
      m_cell_slot.push_back(slot);
      auto d = it.data();
      auto&& ct = d.CellType_GetRef();
      typedef typename std::decay<decltype(ct)>::type type_ct;
      uint8_t mat = MAC_AIR;
      if (ct.GetIsInDomain()) mat = MAC_FLUID;
      else if (mac_amd_detail::is_solid(ct, mac_amd_detail::has_solid<type_ct>()))
        mat = MAC_SOLID;
      else if (mac_amd_detail::is_0_pres(ct, mac_amd_detail::has_0_pres<type_ct>()))
        mat = MAC_0_PRES;
      uint8_t f = (uint8_t)(MAC_FLAG_EXISTS | mat);
      auto&& sp = d.Speed_GetRef();
      if (sp.Get_Const(1)) f |= MAC_FLAG_CONST_X;
      if (sp.Get_Const(2)) f |= MAC_FLAG_CONST_Y;
      if (sp.Get_Const(3)) f |= MAC_FLAG_CONST_Z;
      m_flags[(size_t)slot] = f;
    }
    if (!m_ctx) mac_amd_detail::report(mac_create(&m_ctx, &m_prm), " create");
    if (!m_ctx) return;
    mac_amd_detail::report(mac_upload_bricks(m_ctx, nb, m_keys.data(), m_flags.data(), nullptr, nullptr, nullptr), " upload");
  }

  void End_Iteration() {}

  void Divergence_Projection() {
    if (!m_ctx || m_keys.empty()) return;
    const size_t n = m_flags.size();
    m_u.assign(n, 0.0);
    m_v.assign(n, 0.0);
    m_w.assign(n, 0.0);
    m_p.assign(n, 0.0);
    const bool viscous = m_prm.viscosity > 0.0;
    if (viscous) {
      m_ur.assign(n, 0.0);
      m_vr.assign(n, 0.0);
      m_wr.assign(n, 0.0);
    }
    size_t ci = 0;
    for (iterator it = m_grid.begin(); it != m_grid.end(); ++it, ++ci) {
      auto d = it.data();
      auto&& sp = d.Speed_GetRef();
      const long long s = m_cell_slot[ci];
      m_u[s] = sp.Get(1);
      m_v[s] = sp.Get(2);
      m_w[s] = sp.Get(3);
      if (viscous) {
        auto&& rs = mac_amd_detail::real_speed(d, mac_amd_detail::has_real_speed<decltype(d)>());
        m_ur[s] = rs.Get(1);
        m_vr[s] = rs.Get(2);
        m_wr[s] = rs.Get(3);
      }
    }
    mac_amd_detail::report(mac_set_velocity(m_ctx, m_u.data(), m_v.data(), m_w.data()), " upload speed");
    if (viscous)
      mac_amd_detail::report(mac_set_real_velocity(m_ctx, m_ur.data(), m_vr.data(), m_wr.data()), " upload real speed");
    mac_amd_detail::report(mac_project(m_ctx, 0, &last_stats), " project");
    mac_amd_detail::report(mac_get_pressure(m_ctx, m_p.data()), " download pressure");
    mac_amd_detail::report(mac_get_velocity(m_ctx, m_u.data(), m_v.data(), m_w.data()), " download speed");
    ci = 0;
    for (iterator it = m_grid.begin(); it != m_grid.end(); ++it, ++ci) {
      const long long s = m_cell_slot[ci];
      auto d = it.data();
      auto&& pr = d.Pressure_GetRef();
      mac_amd_detail::pressure_set(pr, m_p[s], mac_amd_detail::has_value_type<typename std::decay<decltype(pr)>::type>());
      auto&& sp = d.Speed_GetRef();
      const double nv[3] = {m_u[s], m_v[s], m_w[s]};
      for (int i = 1; i <= 3; ++i)
        if (sp.Get(i) != nv[i - 1]) sp.Set(i, nv[i - 1]);
    }
  }
};

#endif
