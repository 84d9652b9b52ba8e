// mac_project.hip -- context, brick hierarchy and the C-ABI of libmac_amd.so
// for the free-surface pressure projection (reference:
// Algorithms_Solve_Pressure.h, Policy_Solve_Linear_Umfpack.h).
//
// Host code here only builds the brick structure (keys -> ids, neighbour,
// child and parent tables: a few tens of thousands of entries) and drives the
// kernels of project_kernels.h on the context's stream.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <tuple>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <unordered_map>
#include <vector>

#include "../../include/mac_amd.h"
#include "mac_internal.h"
#include "project_kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return fail(MAC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != 0) return rc_; \
  } while (0)

inline uint64_t pack_key(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
         (uint64_t)(uint32_t)(z + (1 << 20));
}

// Event-timing classes: each is one kernel at one level range, so that its
// average launch time can be matched against a profiler's per-kernel figure.
enum KClass {
  K_DIV, K_PMATVEC, K_UPDATE, K_JACOBI2_0, K_JACOBI_0, K_RESID0, K_PROLONG0, K_RESTRICT1,
  K_C_JACOBI2, K_C_JACOBI, K_C_RESID, K_C_RESTRICT, K_C_PROLONG, K_CSOLVE, K_GRAD, K_MAXDIV, K_INIT, K_FIN, K_TAIL,
  K_PROLONG_JACOBI0, K_RESTRICT_BLOCK1, K_C_RESTRICT_BLOCK, K_C_PROLONG_BLOCK, K_N
};
const char* kNames[K_N] = {"divergence",    "cg_pupdate_matvec", "cg_update",       "jacobi2_l0",     "jacobi_l0",
                           "residual_l0",   "prolong_l0",        "restrict_l1",     "coarse_jacobi2", "coarse_jacobi",
                           "coarse_residual", "coarse_restrict", "coarse_prolong",  "coarse_solve",   "gradient_update",
                           "max_div",       "resid_init",        "finalize",        "coarse_tail",    "prolong_jacobi_l0",
                           "restrict_block_l1", "coarse_restrict_block", "coarse_prolong_block"};

inline int grid_of(int nact) {
  int g = (nact + WPB - 1) / WPB;
  return g < 1 ? 1 : (g > GRID_CAP ? GRID_CAP : g);
}

struct Level {
  int nb = 0;
  std::vector<int> keys;  // 3*nb, shifted so that every key is >= 0
  std::vector<int> h_nbr, h_child, h_parent, h_oct;
  int *nbr = nullptr, *child = nullptr, *parent = nullptr, *oct = nullptr, *act = nullptr;
  int *meta = nullptr;   // per active brick: id + 6 face neighbours (MSTRIDE)
  int *pmeta = nullptr;  // per active brick: prolongation sources from level+1 (PSTRIDE)
  int *rmeta = nullptr;  // per active brick: restriction sources from level-1 (RSTRIDE)
  std::vector<int> h_act;
  int nact = 0;
  long long ndom = 0;
  uint8_t *code = nullptr, *lab = nullptr;
  double ih2[3] = {1, 1, 1};
  // coarse levels: the operator's couplings and diagonal (NCF arrays of
  // nb * MAC_BC values), FP64 and FP32
  double* cf64 = nullptr;
  float* cf32 = nullptr;
  size_t ncell() const { return (size_t)nb * MAC_BC; }
  template <typename T>
  const T* cf() const {
    if constexpr (sizeof(T) == 4) return (const T*)cf32;
    else return (const T*)cf64;
  }
  template <typename T>
  Ih<T> ihc() const {
    return {(T)ih2[0], (T)ih2[1], (T)ih2[2]};
  }
  void *r = nullptr, *z = nullptr, *zt = nullptr, *res = nullptr;
};

}  // namespace

struct mac_ctx {
  mac_params prm;
  double om_pre[8], om_post[8];  // per-sweep Jacobi weights before / after the coarse correction, level 0
  double omc_pre[8], omc_post[8];  // the same on the coarse levels
  BlockW blockw;                   // inner sweeps of the coarse block smoother
  BlockW blockw_deep;              // the single launch per side on the levels from BLOCK_DEEP on
  // 1, 2 or 4 CG iterations captured as a graph, per V-cycle precision (0
  // FP64, 1 FP32) and per parity of the search-direction buffers
  hipGraphExec_t gexec[2][2][3] = {};
  int solves_on_grid = 0;  // projections since the last build (graphs pay off from the third)
  const void* gz[2] = {nullptr, nullptr};  // the V-cycle output of the captured iterations
  int last_iters[2] = {0, 0};              // iterations of the last converged solve, unguided / guided
  hipStream_t st = nullptr;
  uint8_t* pin = nullptr;  // pinned staging for the hierarchy's host tables
  size_t pin_cap = 0;
  int dev = 0;
  int nb0 = 0;
  std::vector<int> keys0;
  std::vector<Level> L;
  uint8_t* flags = nullptr;
  double *u = nullptr, *v = nullptr, *w = nullptr, *u0 = nullptr, *v0 = nullptr, *w0 = nullptr;
  double *x = nullptr, *f = nullptr, *r = nullptr, *d[2] = {nullptr, nullptr}, *q = nullptr, *tmpv = nullptr;
  int* glist = nullptr;
  int nglist = 0;
  // viscous air pressure (viscosity > 0): p_air per cell, and the velocity it
  // is formed from (nullptr: the field being projected)
  double* pD = nullptr;
  const double *ur = nullptr, *vr = nullptr, *wr = nullptr;
  // device copy of a real velocity uploaded by mac_set_real_velocity (owned,
  // part of the current grid's allocations)
  double* ur_own[3] = {nullptr, nullptr, nullptr};
  // slot A (mac_solve_csr): device workspace and pinned state block
  void* csr_ws = nullptr;
  size_t csr_ws_bytes = 0;
  void* csr_pin = nullptr;
  int* dkeys0 = nullptr;  // level-0 shifted chunk keys on the device
  int ext0[3] = {0, 0, 0};
  // coarsest level exact solve
  int nc = 0;
  int *rowof = nullptr, *cellof = nullptr;
  double* cinv64 = nullptr;
  float* cinv32 = nullptr;
  double* partials = nullptr;
  Scal* s = nullptr;
  Scal* hs = nullptr;  // pinned
  bool have_grid = false;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  struct Pend {
    int cls;
    hipEvent_t a, b;
    double bytes;
    long long seq;
  };
  std::vector<Pend> pend;
  std::vector<std::pair<long long, long long>> noop;  // seq ranges of launches that exited at once
  long long seq = 0;
  double kt_ms[K_N] = {0};
  long long kt_n[K_N] = {0};
  double kt_bytes[K_N] = {0};
  // device allocations of the current grid, and those of the grid before,
  // kept for reuse by the next build (a time step rebuilds every step)
  std::vector<std::pair<void*, size_t>> allocs;
  std::multimap<size_t, void*> pool;
};

namespace {

void drop_graphs(mac_ctx* c) {
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 3; ++k) {
        if (c->gexec[i][j][k]) hipGraphExecDestroy(c->gexec[i][j][k]);
        c->gexec[i][j][k] = nullptr;
      }
    c->gz[i] = nullptr;
  }
}

void free_all(mac_ctx* c, bool keep_for_reuse = false) {
  drop_graphs(c);
  c->solves_on_grid = 0;
  for (auto& a : c->allocs) {
    if (keep_for_reuse) c->pool.emplace(a.second, a.first);
    else hipFree(a.first);
  }
  c->allocs.clear();
  if (!keep_for_reuse) {
    for (auto& kv : c->pool) hipFree(kv.second);
    c->pool.clear();
  }
  c->pD = nullptr;
  if (c->ur_own[0] && c->ur == c->ur_own[0]) c->ur = c->vr = c->wr = nullptr;
  c->ur_own[0] = c->ur_own[1] = c->ur_own[2] = nullptr;
  c->L.clear();
  c->have_grid = false;
}

// Device array of n elements: a block of the last grid's of at least that
// size and at most twice it, else a new one.  Contents are undefined.
template <typename T>
int dalloc(mac_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  const size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
  auto it = c->pool.lower_bound(bytes);
  if (it != c->pool.end() && it->first <= 2 * bytes) {
    *p = (T*)it->second;
    c->allocs.emplace_back(it->second, it->first);
    c->pool.erase(it);
    return 0;
  }
  HIPCHK(hipMalloc((void**)p, bytes));
  c->allocs.emplace_back((void*)*p, bytes);
  return 0;
}
// Blocks of the last grid that this one did not take go back to the device.
void drain_pool(mac_ctx* c) {
  for (auto& kv : c->pool) hipFree(kv.second);
  c->pool.clear();
}

template <typename T>
int upload(mac_ctx* c, T** p, const std::vector<T>& h) {
  CHK(dalloc(c, p, h.size()));
  if (!h.empty()) HIPCHK(hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->st));
  return 0;
}

// Host tables bound for the device, sent in one transfer: packed (256-byte
// aligned) into the context's pinned staging buffer and copied into one
// device block, whose parts the destinations then point at.  The staging
// buffer is rewritten only after the stream has been synchronised since the
// previous transfer.
struct Uploads {
  struct Item {
    void** dst;
    size_t off;
  };
  std::vector<Item> items;
  std::vector<uint8_t> host;
  template <typename T>
  void add(T** dst, const std::vector<T>& v) {
    const size_t off = host.size();
    const size_t bytes = v.size() * sizeof(T);
    host.resize(off + ((std::max<size_t>(bytes, 1) + 255) & ~(size_t)255), 0);
    if (bytes) memcpy(host.data() + off, v.data(), bytes);
    items.push_back({(void**)dst, off});
  }
};

int flush_uploads(mac_ctx* c, Uploads& u) {
  if (u.items.empty()) return 0;
  const size_t n = u.host.size();
  if (c->pin_cap < n) {
    if (c->pin) hipHostFree(c->pin);
    c->pin = nullptr;
    c->pin_cap = 0;
    const size_t cap = n + n / 2;
    HIPCHK(hipHostMalloc((void**)&c->pin, cap, hipHostMallocDefault));
    c->pin_cap = cap;
  }
  memcpy(c->pin, u.host.data(), n);
  uint8_t* block;
  CHK(dalloc(c, &block, n));
  HIPCHK(hipMemcpyAsync(block, c->pin, n, hipMemcpyHostToDevice, c->st));
  for (auto& it : u.items) *it.dst = block + it.off;
  u.items.clear();
  u.host.clear();
  return 0;
}

// ---- kernel launch bookkeeping (optional event timing) ----
// An empty launch: what an event pair measures around a kernel that does nothing.
__global__ void k_empty() {}

struct TimedLaunch {
  mac_ctx* c;
  int cls;
  double bytes;
  hipEvent_t a = nullptr, b = nullptr;
  long long myseq;
  bool ext;  // the launch records a and b itself (launch() below)
  TimedLaunch(mac_ctx* c_, int cls_, double bytes_, bool ext_ = true)
      : c(c_), cls(cls_), bytes(bytes_), myseq(c_->seq++), ext(ext_) {
    if (!c->timing) return;
    while (c->evused + 2 > c->evpool.size()) {
      hipEvent_t e;
      hipEventCreate(&e);
      c->evpool.push_back(e);
    }
    a = c->evpool[c->evused++];
    b = c->evpool[c->evused++];
    if (!ext) hipEventRecord(a, c->st);
  }
  ~TimedLaunch() {
    if (!c->timing) return;
    if (!ext) hipEventRecord(b, c->st);
    c->pend.push_back({cls, a, b, bytes, myseq});
  }
};

// Launches kern on the context's stream.  In timing mode the dispatch itself
// records the TimedLaunch's events (hipExtLaunchKernel), so they bracket the
// kernel's own execution, as a kernel-trace profile's timestamps do, without
// the cost of two event records around the launch.
template <typename... KArgs, typename... Args>
void launch(const TimedLaunch& tl, void (*kern)(KArgs...), dim3 grid, dim3 block, Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  std::tuple<KArgs...> t{static_cast<KArgs>(args)...};
  void* p[sizeof...(KArgs)];
  std::apply(
      [&](auto&... e) {
        int i = 0;
        ((p[i++] = (void*)&e), ...);
      },
      t);
  if (tl.a) (void)hipExtLaunchKernel((const void*)kern, grid, block, p, 0, tl.c->st, tl.a, tl.b, 0);
  else (void)hipLaunchKernel((const void*)kern, grid, block, p, 0, tl.c->st);
}

// One-block reduction of the partials of the previous kernel plus the CG
// scalar step (see k_finalize); returns the launch's sequence number.
// to_host: the launch also writes the state to the pinned mirror (poll_scal).
// x = 0 on every cell of the level-0 bricks
void clear_x(mac_ctx* c) {
  const size_t n16 = (size_t)c->nb0 * MAC_BC * sizeof(double) / 16;
  const unsigned blocks = (unsigned)std::min<size_t>((n16 + 511) / 512, 65535);
  k_zero16<<<blocks, 256, 0, c->st>>>((uint4*)c->x, n16);
}

long long finalize(mac_ctx* c, int nparts, int op, bool to_host = false) {
  TimedLaunch tl(c, K_FIN, 0.0, false);
  k_finalize<<<1, 1024, 0, c->st>>>(nparts, c->partials, c->s, op, c->prm.rtol * c->prm.rtol, to_host ? c->hs : nullptr);
  return tl.myseq;
}

void flush_timing(mac_ctx* c) {
  if (!c->timing) {
    c->noop.clear();
    return;
  }
  for (auto& p : c->pend) {
    bool noop = false;  // launched after convergence: the kernel exited at once
    for (auto& rg : c->noop) noop |= (p.seq >= rg.first && p.seq <= rg.second);
    if (noop) continue;
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    c->kt_ms[p.cls] += ms;
    c->kt_n[p.cls] += 1;
    c->kt_bytes[p.cls] += p.bytes;
  }
  c->pend.clear();
  c->noop.clear();
  c->evused = 0;
}

// ---- hierarchy construction ----

// Brick index by key (keys >= 0): a dense array over the keys' bounding box
// when that box is at most 64 times the brick count, else a hash map.
struct KeyIndex {
  int ext[3] = {0, 0, 0};
  bool dense_ok = false;
  std::vector<int> dense;
  std::unordered_map<uint64_t, int> map;
  void init(const int* keys, int nb) {
    ext[0] = ext[1] = ext[2] = 1;
    for (int b = 0; b < nb; ++b)
      for (int i = 0; i < 3; ++i) ext[i] = std::max(ext[i], keys[3 * b + i] + 1);
    const double n = (double)ext[0] * ext[1] * ext[2];
    dense_ok = n <= 64.0 * std::max(nb, 1024);
    if (dense_ok) {
      dense.assign((size_t)n, -1);
    } else {
      dense.clear();
      map.clear();
      map.reserve(2 * (size_t)nb);
    }
  }
  // stores id at the key unless one is there; returns the id found there, or -1
  int insert(int x, int y, int z, int id) {
    if (dense_ok) {
      int& v = dense[(size_t)x + (size_t)ext[0] * ((size_t)y + (size_t)ext[1] * z)];
      const int was = v;
      if (was < 0) v = id;
      return was;
    }
    auto r = map.emplace(pack_key(x, y, z), id);
    return r.second ? -1 : r.first->second;
  }
  int find(int x, int y, int z) const {
    if (dense_ok) {
      if (x < 0 || y < 0 || z < 0 || x >= ext[0] || y >= ext[1] || z >= ext[2]) return -1;
      return dense[(size_t)x + (size_t)ext[0] * ((size_t)y + (size_t)ext[1] * z)];
    }
    auto it = map.find(pack_key(x, y, z));
    return it == map.end() ? -1 : it->second;
  }
};

void build_nbr(Level& lv) {
  KeyIndex idx;
  idx.init(lv.keys.data(), lv.nb);
  for (int b = 0; b < lv.nb; ++b) idx.insert(lv.keys[3 * b], lv.keys[3 * b + 1], lv.keys[3 * b + 2], b);
  lv.h_nbr.assign((size_t)lv.nb * 27, -1);
  for (int b = 0; b < lv.nb; ++b) {
    const int kx = lv.keys[3 * b], ky = lv.keys[3 * b + 1], kz = lv.keys[3 * b + 2];
    int* nb = &lv.h_nbr[(size_t)b * 27];
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) nb[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)] = idx.find(kx + ox, ky + oy, kz + oz);
  }
}

// Dense inverse of the coarsest operator (Cholesky; a tiny diagonal shift if
// a component has no Dirichlet contact, which keeps the preconditioner SPD).
// cf (coarse levels): the operator's couplings and diagonal (NCF arrays of
// MAC_BC values); nullptr on level 0, whose operator is the codes and ih2.
void dense_inverse(const std::vector<uint8_t>& code, const double ih2[3], const double* cf, std::vector<int>& rowof,
                   std::vector<int>& cellof, std::vector<double>& inv) {
  rowof.assign(MAC_BC, -1);
  cellof.clear();
  for (int t = 0; t < MAC_BC; ++t)
    if (code[t] & CODE_DOM) {
      rowof[t] = (int)cellof.size();
      cellof.push_back(t);
    }
  int n = (int)cellof.size();
  inv.assign((size_t)n * n, 0.0);
  if (n == 0) return;
  std::vector<double> A((size_t)n * n, 0.0);
  const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
  for (int i = 0; i < n; ++i) {
    int t = cellof[i];
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    if (cf) A[(size_t)i * n + i] = cf[(size_t)6 * MAC_BC + t];
    for (int k = 0; k < 6; ++k) {
      if (!cf && !(code[t] & (1u << k))) continue;
      double w = cf ? cf[(size_t)k * MAC_BC + t] : ih2[k >> 1];
      if (!cf) A[(size_t)i * n + i] += w;
      int xx = x + dx[k], yy = y + dy[k], zz = z + dz[k];
      if (xx < 0 || yy < 0 || zz < 0 || xx > 7 || yy > 7 || zz > 7) continue;
      int j = rowof[xx + 8 * yy + 64 * zz];
      if (j >= 0) A[(size_t)i * n + j] -= w;
    }
  }
  double maxd = 0;
  for (int i = 0; i < n; ++i) maxd = std::max(maxd, A[(size_t)i * n + i]);
  // Cholesky A = G G^T
  std::vector<double> G = A;
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
  // inverse column by column: solve G y = e_c, G^T x = y
  std::vector<double> y(n);
  for (int col = 0; col < n; ++col) {
    for (int i = 0; i < n; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= G[(size_t)i * n + k] * y[k];
      y[i] = s / G[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= G[(size_t)k * n + i] * inv[(size_t)k * n + col];
      inv[(size_t)i * n + col] = s / G[(size_t)i * n + i];
    }
  }
  // symmetrise (rounding)
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double a = 0.5 * (inv[(size_t)i * n + j] + inv[(size_t)j * n + i]);
      inv[(size_t)i * n + j] = inv[(size_t)j * n + i] = a;
    }
}

int build_grid(mac_ctx* c, int nbricks, const int32_t* keys3, const uint8_t* flags) {
  free_all(c, true);
  c->nb0 = nbricks;
  c->keys0.assign(keys3, keys3 + 3 * (size_t)nbricks);
  // range check + shift
  int mn[3] = {1 << 30, 1 << 30, 1 << 30};
  for (int b = 0; b < nbricks; ++b)
    for (int i = 0; i < 3; ++i) {
      if (keys3[3 * b + i] < -(1 << 19) || keys3[3 * b + i] >= (1 << 19)) return fail(MAC_E_ARG, "chunk key out of range");
      mn[i] = std::min(mn[i], (int)keys3[3 * b + i]);
    }
  const mac_params& P = c->prm;
  double inv_h[3];
  for (int i = 0; i < 3; ++i) inv_h[i] = 1.0 / P.h[i];

  c->L.clear();
  c->L.emplace_back();
  {
    Level& l0 = c->L[0];
    l0.nb = nbricks;
    l0.keys.resize(3 * (size_t)nbricks);
    for (int b = 0; b < nbricks; ++b)
      for (int i = 0; i < 3; ++i) l0.keys[3 * b + i] = keys3[3 * b + i] - mn[i];
    for (int i = 0; i < 3; ++i) l0.ih2[i] = inv_h[i] * inv_h[i];
    KeyIndex seen;
    seen.init(l0.keys.data(), nbricks);
    for (int b = 0; b < nbricks; ++b)
      if (seen.insert(l0.keys[3 * b], l0.keys[3 * b + 1], l0.keys[3 * b + 2], b) >= 0)
        return fail(MAC_E_ARG, "duplicate chunk key");
  }
  // coarse levels until a single brick remains
  while (c->L.back().nb > 1) {
    Level& f = c->L.back();
    Level cl;
    KeyIndex map;
    {
      std::vector<int> ck(f.keys.size());
      for (size_t i = 0; i < ck.size(); ++i) ck[i] = f.keys[i] >> 1;
      map.init(ck.data(), f.nb);
    }
    f.h_parent.assign(f.nb, -1);
    f.h_oct.assign(f.nb, 0);
    cl.h_child.reserve((size_t)f.nb * 8);
    for (int b = 0; b < f.nb; ++b) {
      int kx = f.keys[3 * b] >> 1, ky = f.keys[3 * b + 1] >> 1, kz = f.keys[3 * b + 2] >> 1;
      int id = map.insert(kx, ky, kz, cl.nb);
      if (id < 0) {
        id = cl.nb++;
        cl.keys.push_back(kx);
        cl.keys.push_back(ky);
        cl.keys.push_back(kz);
        cl.h_child.resize((size_t)cl.nb * 8, -1);
      }
      int o = (f.keys[3 * b] & 1) + 2 * (f.keys[3 * b + 1] & 1) + 4 * (f.keys[3 * b + 2] & 1);
      cl.h_child[(size_t)id * 8 + o] = b;
      f.h_parent[b] = id;
      f.h_oct[b] = o;
    }
    for (int i = 0; i < 3; ++i) cl.ih2[i] = f.ih2[i] * 0.25;
    c->L.push_back(std::move(cl));
  }
  for (auto& lv : c->L) build_nbr(lv);

  // device structures
  Uploads up;
  for (size_t l = 0; l < c->L.size(); ++l) {
    Level& lv = c->L[l];
    up.add(&lv.nbr, lv.h_nbr);
    if (!lv.h_child.empty()) up.add(&lv.child, lv.h_child);
    if (!lv.h_parent.empty()) {
      up.add(&lv.parent, lv.h_parent);
      up.add(&lv.oct, lv.h_oct);
    }
    CHK(dalloc(c, &lv.code, (size_t)lv.nb * MAC_BC));
    CHK(dalloc(c, &lv.lab, (size_t)lv.nb * MAC_BC));
  }
  CHK(flush_uploads(c, up));
  size_t n0 = (size_t)nbricks * MAC_BC;
  CHK(dalloc(c, &c->flags, n0));
  HIPCHK(hipMemcpyAsync(c->flags, flags, n0, hipMemcpyDefault, c->st));
  Level& l0 = c->L[0];
  k_code0<<<nbricks, KTHREADS, 0, c->st>>>(nbricks, l0.nbr, c->flags, l0.code, l0.lab);
  HIPCHK(hipGetLastError());
  for (size_t l = 1; l < c->L.size(); ++l) {
    Level& lv = c->L[l];
    k_coarsen_lab<<<lv.nb, KTHREADS, 0, c->st>>>(lv.child, c->L[l - 1].lab, lv.lab);
    k_code_c<<<lv.nb, KTHREADS, 0, c->st>>>(lv.nbr, lv.lab, lv.code);
  }
  // the coarse operators, level by level from the one above
  for (size_t l = 1; l < c->L.size(); ++l) {
    Level& lv = c->L[l];
    const Level& f = c->L[l - 1];
    CHK(dalloc(c, &lv.cf64, NCF * lv.ncell()));
    CHK(dalloc(c, &lv.cf32, NCF * lv.ncell()));
    k_coarse_coef<<<lv.nb, MAC_BC, 0, c->st>>>(f.code, f.nbr, l > 1 ? f.cf64 : nullptr, f.ncell(), f.ih2[0], f.ih2[1], f.ih2[2],
                                                lv.child, lv.code, lv.nbr, lv.cf64, lv.cf32, lv.ncell());
  }
  HIPCHK(hipGetLastError());
  // active bricks per level: every level counted, then one read back
  std::vector<size_t> coff(c->L.size() + 1, 0);
  for (size_t l = 0; l < c->L.size(); ++l) coff[l + 1] = coff[l] + (size_t)c->L[l].nb;
  int* dcnt;
  CHK(dalloc(c, &dcnt, coff.back()));
  for (size_t l = 0; l < c->L.size(); ++l)
    k_brick_count<<<c->L[l].nb, KTHREADS, 0, c->st>>>(c->L[l].nb, c->L[l].code, dcnt + coff[l]);
  HIPCHK(hipGetLastError());
  std::vector<int> allcnt(coff.back());
  std::vector<uint8_t> topcode(MAC_BC);  // the coarsest level's codes and operator, for its exact solve
  std::vector<double> topcf;
  HIPCHK(hipMemcpyAsync(allcnt.data(), dcnt, coff.back() * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(topcode.data(), c->L.back().code, MAC_BC, hipMemcpyDeviceToHost, c->st));
  if (c->L.size() > 1) {
    const Level& lt = c->L.back();  // a single brick
    topcf.resize((size_t)NCF * MAC_BC);
    for (int q = 0; q < NCF; ++q)
      HIPCHK(hipMemcpyAsync(topcf.data() + (size_t)q * MAC_BC, lt.cf64 + (size_t)q * lt.ncell(), MAC_BC * sizeof(double),
                            hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  for (size_t l = 0; l < c->L.size(); ++l) {
    Level& lv = c->L[l];
    const int* cnt = allcnt.data() + coff[l];
    std::vector<int> act;
    lv.ndom = 0;
    for (int b = 0; b < lv.nb; ++b)
      if (cnt[b] > 0) {
        act.push_back(b);
        lv.ndom += cnt[b];
      }
    lv.nact = (int)act.size();
    lv.h_act = act;
    up.add(&lv.act, act);
    if (l == 0) {
      // bricks whose faces the gradient update may touch: active, or with an
      // active -x/-y/-z neighbour
      std::vector<int> gl;
      for (int b = 0; b < lv.nb; ++b) {
        bool use = cnt[b] > 0;
        for (int k : {12, 10, 4}) {
          int nbb = lv.h_nbr[(size_t)b * 27 + k];
          if (nbb >= 0 && cnt[nbb] > 0) use = true;
        }
        if (use) gl.push_back(b);
      }
      c->nglist = (int)gl.size();
      up.add(&c->glist, gl);
      up.add(&c->dkeys0, c->L[0].keys);
      for (int i = 0; i < 3; ++i) c->ext0[i] = 0;
      for (int b = 0; b < c->nb0; ++b)
        for (int i = 0; i < 3; ++i) c->ext0[i] = std::max(c->ext0[i], c->L[0].keys[3 * b + i] + 1);
    }
  }
  // flattened per-active-brick metadata
  for (size_t l = 0; l < c->L.size(); ++l) {
    Level& lv = c->L[l];
    std::vector<int> meta((size_t)std::max(lv.nact, 1) * MSTRIDE, -1);
    for (int i = 0; i < lv.nact; ++i) {
      int b = lv.h_act[i];
      const int* nb = &lv.h_nbr[(size_t)b * 27];
      int* m = &meta[(size_t)i * MSTRIDE];
      m[0] = b;
      m[1] = nb[12];
      m[2] = nb[14];
      m[3] = nb[10];
      m[4] = nb[16];
      m[5] = nb[4];
      m[6] = nb[22];
      m[7] = 0;
    }
    up.add(&lv.meta, meta);
    if (l + 1 < c->L.size()) {
      Level& cl = c->L[l + 1];
      std::vector<int> pm((size_t)std::max(lv.nact, 1) * PSTRIDE, -1);
      for (int i = 0; i < lv.nact; ++i) {
        int B = lv.h_act[i];
        int P = lv.h_parent[B], o = lv.h_oct[B];
        int ex = (o & 1) ? 1 : -1, ey = (o & 2) ? 1 : -1, ez = (o & 4) ? 1 : -1;
        int* m = &pm[(size_t)i * PSTRIDE];
        m[0] = B;
        m[1] = o;
        for (int k = 0; k < 8; ++k) {
          int sx = (k & 1) ? ex : 0, sy = (k & 2) ? ey : 0, sz = (k & 4) ? ez : 0;
          m[2 + k] = cl.h_nbr[(size_t)P * 27 + (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)];
        }
        const int* nb = &lv.h_nbr[(size_t)B * 27];  // face neighbours, the order of meta
        m[10] = nb[12];
        m[11] = nb[14];
        m[12] = nb[10];
        m[13] = nb[16];
        m[14] = nb[4];
        m[15] = nb[22];
      }
      up.add(&lv.pmeta, pm);
      std::vector<int> rm((size_t)std::max(cl.nact, 1) * RSTRIDE, -1);
      for (int i = 0; i < cl.nact; ++i) {
        int C = cl.h_act[i];
        int* m = &rm[(size_t)i * RSTRIDE];
        m[0] = C;
        m[1] = m[2] = m[3] = 0;
        for (int t = 0; t < 64; ++t) {
          int rx = (t & 3) - 1, ry = ((t >> 2) & 3) - 1, rz = (t >> 4) - 1;
          int cn = cl.h_nbr[(size_t)C * 27 + ((rx >> 1) + 1) + 3 * ((ry >> 1) + 1) + 9 * ((rz >> 1) + 1)];
          m[4 + t] = cn >= 0 ? cl.h_child[(size_t)cn * 8 + (rx & 1) + 2 * (ry & 1) + 4 * (rz & 1)] : -1;
        }
      }
      up.add(&cl.rmeta, rm);
    }
  }
  // coarsest exact solve
  {
    Level& lc = c->L.back();
    std::vector<int> rowof, cellof;
    std::vector<double> inv;
    dense_inverse(topcode, lc.ih2, topcf.empty() ? nullptr : topcf.data(), rowof, cellof, inv);
    c->nc = (int)cellof.size();
    std::vector<float> inv32(inv.begin(), inv.end());
    up.add(&c->rowof, rowof);
    up.add(&c->cellof, cellof);
    up.add(&c->cinv64, inv);
    up.add(&c->cinv32, inv32);
    CHK(flush_uploads(c, up));
  }
  // vectors
  CHK(dalloc(c, &c->u, n0));
  CHK(dalloc(c, &c->v, n0));
  CHK(dalloc(c, &c->w, n0));
  CHK(dalloc(c, &c->u0, n0));
  CHK(dalloc(c, &c->v0, n0));
  CHK(dalloc(c, &c->w0, n0));
  CHK(dalloc(c, &c->x, n0));
  CHK(dalloc(c, &c->f, n0));
  CHK(dalloc(c, &c->r, n0));
  CHK(dalloc(c, &c->d[0], n0));
  CHK(dalloc(c, &c->d[1], n0));
  CHK(dalloc(c, &c->q, n0));
  CHK(dalloc(c, &c->tmpv, n0));
  // every vector starts at zero (zero outside the domain is assumed by the
  // kernels; reused blocks hold old data): one launch clears them all
  ZeroList zl;
  zl.k = 0;
  auto clear = [&](double* p, size_t n) -> int {
    if (zl.k == ZeroList::CAP) {
      k_zero_list<<<dim3(256, zl.k), KTHREADS, 0, c->st>>>(zl);
      zl.k = 0;
    }
    zl.p[zl.k] = p;
    zl.n[zl.k++] = n;
    return 0;
  };
  for (double* p : {c->u, c->v, c->w, c->x, c->f, c->r, c->d[0], c->d[1], c->q, c->tmpv}) clear(p, n0);
  for (size_t l = 0; l < c->L.size(); ++l) {
    Level& lv = c->L[l];
    size_t n = (size_t)lv.nb * MAC_BC;
    double *a, *b, *cc, *dd;
    CHK(dalloc(c, &a, n));
    CHK(dalloc(c, &b, n));
    CHK(dalloc(c, &cc, n));
    CHK(dalloc(c, &dd, n));
    for (double* p : {a, b, cc, dd}) clear(p, n);
    lv.r = a;
    lv.z = b;
    lv.zt = cc;
    lv.res = dd;
  }
  if (zl.k > 0) k_zero_list<<<dim3(256, zl.k), KTHREADS, 0, c->st>>>(zl);
  HIPCHK(hipGetLastError());
  CHK(dalloc(c, &c->partials, std::max(nbricks, 1024)));
  CHK(dalloc(c, &c->s, 1));
  HIPCHK(hipMemsetAsync(c->s, 0, sizeof(Scal), c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  drain_pool(c);
  c->have_grid = true;
  return 0;
}

int check_params(const mac_params* p) {
  if (p->nu < 1 || p->nu > 8 || p->max_iter < 1 || !(p->rtol > 0)) return fail(MAC_E_ARG, "bad solver parameters");
  if (p->smoother != MAC_SMOOTH_JACOBI && p->smoother != MAC_SMOOTH_CHEBYSHEV) return fail(MAC_E_ARG, "unknown smoother");
  if (!(p->viscosity >= 0)) return fail(MAC_E_ARG, "viscosity must be >= 0");
  if (p->coarse_inner < 1 || p->coarse_inner > 8) return fail(MAC_E_ARG, "coarse_inner must lie in 1..8");
  if (!(p->coarse_damp > 0 && p->coarse_damp <= 1)) return fail(MAC_E_ARG, "coarse_damp must lie in (0, 1]");
  if (p->smoother == MAC_SMOOTH_CHEBYSHEV && !(p->cheb_lower > 0 && p->cheb_lower < 2 && p->cheb_lower_coarse > 0 &&
                                                p->cheb_lower_coarse < 2))
    return fail(MAC_E_ARG, "cheb_lower and cheb_lower_coarse must lie in (0, 2)");
  if (p->nu_coarse < 0 || p->nu_coarse > 8) return fail(MAC_E_ARG, "nu_coarse must lie in 0..8");
  if (p->smoother == MAC_SMOOTH_JACOBI && !(p->omega > 0 && p->omega < 2)) return fail(MAC_E_ARG, "omega must lie in (0, 2)");
  return 0;
}

// Sweep weights from the parameters: nu copies of omega, or the reciprocals
// of the Chebyshev nodes of [a, 2] (the polynomial of degree nu with value 1
// at 0 that is smallest on [a, 2]); the sweeps after the coarse correction
// take them in reverse order.
void schedule(const mac_params& p, int nu, double a, double* pre, double* post) {
  const double b = 2.0;
  for (int k = 0; k < nu; ++k) {
    double om = p.omega;
    if (p.smoother == MAC_SMOOTH_CHEBYSHEV) om = 1.0 / (0.5 * (b + a) + 0.5 * (b - a) * std::cos((2.0 * k + 1.0) * M_PI / (2.0 * nu)));
    pre[k] = om;
    post[nu - 1 - k] = om;
  }
}

inline int nu_of(const mac_params& p, int level) {
  const int nu = (level > 0 && p.nu_coarse > 0) ? p.nu_coarse : p.nu;
  return std::min(8, std::max(1, nu));
}

void set_schedule(mac_ctx* c) {
  schedule(c->prm, nu_of(c->prm, 0), c->prm.cheb_lower, c->om_pre, c->om_post);
  schedule(c->prm, nu_of(c->prm, 1), c->prm.cheb_lower_coarse, c->omc_pre, c->omc_post);
  double pre[8], post[8];
  const int m = std::min(8, std::max(1, c->prm.coarse_inner));
  schedule(c->prm, m, c->prm.cheb_lower_coarse, pre, post);
  c->blockw.m = m;
  c->blockw.damp = c->prm.coarse_damp;
  for (int k = 0; k < 8; ++k) {
    c->blockw.om64[k] = k < m ? pre[k] : 0.0;
    c->blockw.om[k] = (float)c->blockw.om64[k];
  }
  // from BLOCK_DEEP on, the nu_coarse launches of m sweeps become one
  // undamped launch of nu_coarse * m sweeps (at most 8)
  const int md = std::min(8, nu_of(c->prm, 1) * m);
  schedule(c->prm, md, c->prm.cheb_lower_coarse, pre, post);
  c->blockw_deep.m = md;
  c->blockw_deep.damp = 1.0;
  for (int k = 0; k < 8; ++k) {
    c->blockw_deep.om64[k] = k < md ? pre[k] : 0.0;
    c->blockw_deep.om[k] = (float)c->blockw_deep.om64[k];
  }
}

// ---- the V-cycle ----

inline long long lvl_dom(const mac_ctx* c, int l) { return c->L[l].ndom; }

// The first coarse level on which the block smoother runs as one launch per
// side (those levels have few bricks, so a launch is mostly latency).
#define BLOCK_DEEP 3

// Damped-Jacobi sweep at level l (level 0 reads the CG residual in FP64).
template <typename TZ>
void sweep(mac_ctx* c, int l, const TZ* zin, TZ* zout, int want_dot, double omega) {
  Level& lv = c->L[l];
  const TZ ih[3] = {(TZ)lv.ih2[0], (TZ)lv.ih2[1], (TZ)lv.ih2[2]};
  const TZ om = (TZ)omega;
  const double sz = sizeof(TZ), nd = (double)lv.ndom;
  const double sr = (l == 0) ? 8.0 : sz;
  {
  // (the coarse levels' coefficient arrays, read near boundaries only, are
  // not counted)
  TimedLaunch tl(c, l == 0 ? K_JACOBI_0 : K_C_JACOBI, nd * (sr + (zin ? 2 : 1) * sz + 1));
  if (l == 0)
    launch(tl, k_jacobi<double, TZ>, grid_of(lv.nact), KTHREADS, lv.nact, lv.meta, lv.code, c->r, zin, zout, ih[0], ih[1],
                                                                   ih[2], om, want_dot, c->partials, c->s);
  else
    launch(tl, k_jacobi_coef<TZ>, lv.nact, KTHREADS, lv.nact, lv.meta, lv.code, (const TZ*)lv.r, zin, zout, lv.cf<TZ>(),
                                                       lv.ncell(), lv.ihc<TZ>(), om, c->s);
  }
  // coarse sweeps never take the dot product (the bottom level is level 0 then)
  if (want_dot) finalize(c, grid_of(lv.nact), OP_RZ);
}

// z = V(r) at level 0 (r = the CG residual).  Levels below the first level
// without unknowns are not visited: the bottom visited level then just
// smooths (nu sweeps down, nu up), which keeps the cycle symmetric.  Every
// cycle uses the same buffers in the same roles (each level starts from its
// z / zt pair and ends where the sweep count puts it), so a captured cycle
// replays correctly.
template <typename TZ>
int vcycle(mac_ctx* c, TZ** zout) {
  const int nl = (int)c->L.size();
  const double sz = sizeof(TZ);
  int nlv = nl;
  for (int l = 1; l < nl; ++l)
    if (c->L[l].nact == 0) {
      nlv = l;
      break;
    }
  const int bot = nlv - 1;
  // the last two levels in one launch: the top brick and its children level
  const bool tail = c->prm.coarse_tail && bot == nl - 1 && nl >= 3 && c->L[nl - 1].nb == 1;
  const int tlev = tail ? nl - 2 : bot;
  std::vector<TZ*> zfin(nl, nullptr);
  std::vector<char> first_done(nl, 0);
  // the block smoother runs on the coarse levels above the last two (those
  // smooth point-wise, fused or not)
  auto is_block = [&](int l) { return l > 0 && l < nl - 2 && c->prm.coarse_inner > 1; };
  // down
  for (int l = 0; l < tlev; ++l) {
    Level& lv = c->L[l];
    const int nu = nu_of(c->prm, l);
    const double* opre = l ? c->omc_pre : c->om_pre;
    TZ ih[3] = {(TZ)lv.ih2[0], (TZ)lv.ih2[1], (TZ)lv.ih2[2]};
    TZ* z = (TZ*)lv.z;
    TZ* zt = (TZ*)lv.zt;
    TZ* res = (TZ*)lv.res;
    double nd = (double)lvl_dom(c, l);
    double sr = (l == 0) ? 8.0 : sz;
    const bool block = is_block(l);
    if (block) {
      const bool deep = l >= BLOCK_DEEP && nu > 1;
      // the first step from zero ran inside the restriction into this level
      for (int k = first_done[l] ? 1 : 0; k < (deep ? 1 : nu); ++k) {
        TimedLaunch tl(c, K_C_JACOBI, nd * ((k ? 3 : 2) * sz + 1));  // r, zin, zout, code
        launch(tl, k_jacobi_block<TZ, TZ>, lv.nact, KTHREADS, lv.nact, lv.meta, lv.code, (const TZ*)lv.r, k ? z : nullptr,
                                                                 k ? zt : z, lv.cf<TZ>(), lv.ncell(), lv.ihc<TZ>(),
                                                                 deep ? c->blockw_deep : c->blockw, c->s);
        if (k) std::swap(z, zt);
      }
    } else if (nu >= 2) {
      // the first two sweeps from zero in one pass
      if (l == 0) {
        TimedLaunch tl(c, K_JACOBI2_0, nd * (sr + sz + 1));
        launch(tl, k_jacobi2<double, TZ>, grid_of(lv.nact), KTHREADS, lv.nact, lv.meta, lv.code, c->r, z, ih[0], ih[1],
                                                                         ih[2], (TZ)opre[0], (TZ)opre[1], c->s);
      } else {
        // a coarse level: the two sweeps as two launches (zt from zero, then z)
        sweep<TZ>(c, l, nullptr, zt, 0, opre[0]);
        sweep<TZ>(c, l, zt, z, 0, opre[1]);
      }
    } else {
      sweep<TZ>(c, l, nullptr, z, 0, opre[0]);
    }
    for (int k = 2; k < nu && !block; ++k) {
      sweep<TZ>(c, l, z, zt, 0, opre[k]);
      std::swap(z, zt);
    }
    Level& cl = c->L[l + 1];
    {
      TimedLaunch tl(c, l == 0 ? K_RESID0 : K_C_RESID, nd * (sr + 2 * sz + 1));
      if (l == 0)
        launch(tl, k_residual<double, TZ>, grid_of(lv.nact), KTHREADS, lv.nact, lv.meta, lv.code, c->r, z, res, ih[0], ih[1],
                                                                         ih[2], c->s);
      else
        launch(tl, k_residual_coef<TZ>, lv.nact, KTHREADS, lv.nact, lv.meta, lv.code, (const TZ*)lv.r, z, res, lv.cf<TZ>(),
                                                             lv.ncell(), lv.ihc<TZ>(), c->s);
    }
    {
      // into a block level visited below: the restriction also takes that
      // level's first smoothing step from zero (brick-local)
      const bool fuse = c->prm.fuse_coarse && l + 1 < tlev && is_block(l + 1);
      const bool cdeep = l + 1 >= BLOCK_DEEP && nu_of(c->prm, l + 1) > 1;
      const double cd = (double)cl.ndom;
      TZ* zsm = fuse ? (TZ*)cl.z : nullptr;
      const BlockW& bw = cdeep ? c->blockw_deep : c->blockw;
      {
        TimedLaunch tl(c, fuse ? (l == 0 ? K_RESTRICT_BLOCK1 : K_C_RESTRICT_BLOCK) : (l == 0 ? K_RESTRICT1 : K_C_RESTRICT),
                       nd * sz + cd * sz + (fuse ? cd * (sz + 1) : 0.0));
        launch(tl, k_restrict_sep<TZ>, cl.nact, KTHREADS, cl.rmeta, cl.code, res, (TZ*)cl.r, c->s, zsm, cl.cf<TZ>(), cl.ncell(),
                                                            cl.ihc<TZ>(), bw);
      }
      first_done[l + 1] = fuse;
    }
    zfin[l] = z;  // pre-smoothed iterate, finished on the way up
  }
  if (tail) {
    Level& lt = c->L[tlev];
    const Level& top = c->L[nl - 1];
    TailArgs a;
    for (int o = 0; o < 8; ++o) a.child[o] = top.h_child[o];
    a.nu = nu_of(c->prm, tlev);
    a.nc = c->nc;
    for (int k = 0; k < 3; ++k) a.ih2[k] = lt.ih2[k];
    for (int k = 0; k < 8; ++k) {
      a.om_pre[k] = c->omc_pre[k];
      a.om_post[k] = c->omc_post[k];
    }
    const TZ* cinv = std::is_same<TZ, float>::value ? (const TZ*)c->cinv32 : (const TZ*)c->cinv64;
    {
      TimedLaunch tl(c, K_TAIL, 0.0);
      launch(tl, k_vcycle_tail<TZ>, 1, TAIL_THREADS, a, lt.code, lt.cf<TZ>(), lt.ncell(), (const TZ*)lt.r, (TZ*)lt.z, top.code,
                                                       c->rowof, c->cellof, cinv, c->s);
    }
    zfin[tlev] = (TZ*)lt.z;
  } else {
    Level& lb = c->L[bot];
    const int want_dot = (bot == 0) ? 1 : 0;
    if (bot == nl - 1) {
      const TZ* cinv = std::is_same<TZ, float>::value ? (const TZ*)c->cinv32 : (const TZ*)c->cinv64;
      TimedLaunch* tl = new TimedLaunch(c, K_CSOLVE, 0.0, false);
      if (bot == 0)
        k_coarse_solve<double, TZ><<<(MAC_BC + 63) / 64, 64, 0, c->st>>>(c->nc, c->rowof, c->cellof, cinv, c->r, (TZ*)lb.z,
                                                                          want_dot, c->partials, c->s);
      else
        k_coarse_solve<TZ, TZ><<<(MAC_BC + 63) / 64, 64, 0, c->st>>>(c->nc, c->rowof, c->cellof, cinv, (const TZ*)lb.r, (TZ*)lb.z,
                                                                      want_dot, c->partials, c->s);
      delete tl;
      if (want_dot) finalize(c, (MAC_BC + 63) / 64, OP_RZ);
      zfin[bot] = (TZ*)lb.z;
    } else {
      TZ* z = (TZ*)lb.z;
      TZ* zt = (TZ*)lb.zt;
      const int nu = nu_of(c->prm, bot);
      const double* opre = bot ? c->omc_pre : c->om_pre;
      const double* opost = bot ? c->omc_post : c->om_post;
      for (int k = 0; k < 2 * nu; ++k) {
        sweep<TZ>(c, bot, k ? z : nullptr, k ? zt : z, (k == 2 * nu - 1) ? want_dot : 0, k < nu ? opre[k] : opost[k - nu]);
        if (k) std::swap(z, zt);
      }
      zfin[bot] = z;
    }
  }
  // up
  for (int l = tlev - 1; l >= 0; --l) {
    Level& lv = c->L[l];
    Level& cl = c->L[l + 1];
    TZ* z = zfin[l];
    TZ* zt = (z == (TZ*)lv.z) ? (TZ*)lv.zt : (TZ*)lv.z;
    double nd = (double)lvl_dom(c, l);
    const bool blk = is_block(l);
    const bool deep = blk && l >= BLOCK_DEEP && nu_of(c->prm, l) > 1;
    const int nu = deep ? 1 : nu_of(c->prm, l);
    const double* opost = l ? c->omc_post : c->om_post;
    const TZ ih[3] = {(TZ)lv.ih2[0], (TZ)lv.ih2[1], (TZ)lv.ih2[2]};
    int k0 = 0;
    // level 0: the slab form in both precisions (the FP64 column form needs
    // 200 registers and ran slower than the two launches)
#ifndef MAC_FUSE_L0
#define MAC_FUSE_L0 1
#endif
#ifndef MAC_PJ_SLAB
#define MAC_PJ_SLAB 1
#endif
    if (c->prm.fuse_coarse && (blk || (MAC_FUSE_L0 && l == 0 && (sizeof(TZ) == 4 || MAC_PJ_SLAB)))) {
      // prolongation fused with the first post-smoothing step: zt = S(z + P zc)
      if (blk) {
        TimedLaunch tl(c, K_C_PROLONG_BLOCK, nd * (3 * sz + 2) + (double)cl.ndom * sz);
        launch(tl, k_prolong_block<TZ>, lv.nact, KTHREADS, lv.nact, lv.pmeta, lv.code, (const TZ*)lv.r, zfin[l + 1], z, zt,
                                                              lv.cf<TZ>(), lv.ncell(), lv.ihc<TZ>(),
                                                              deep ? c->blockw_deep : c->blockw, c->s);
      } else {
        const int want_dot = nu == 1 ? 1 : 0;
        TimedLaunch tl(c, K_PROLONG_JACOBI0, nd * (8.0 + 2 * sz + 2) + (double)cl.ndom * sz);
        if (MAC_PJ_SLAB) {
          launch(tl, k_prolong_jacobi_slab<TZ>, lv.nact, KTHREADS, lv.nact, lv.pmeta, lv.code, c->r, zfin[l + 1], z, zt, ih[0],
                                                                      ih[1], ih[2], (TZ)opost[0], want_dot, c->partials, c->s);
          if (want_dot) finalize(c, lv.nact, OP_RZ);
        } else {
          launch(tl, k_prolong_jacobi<TZ>, grid_of(lv.nact), KTHREADS, lv.nact, lv.pmeta, lv.code, c->r, zfin[l + 1], z, zt,
                                                                          ih[0], ih[1], ih[2], (TZ)opost[0], want_dot, c->partials,
                                                                          c->s);
          if (want_dot) finalize(c, grid_of(lv.nact), OP_RZ);
        }
      }
      std::swap(z, zt);
      k0 = 1;
    } else {
      TimedLaunch tl(c, l == 0 ? K_PROLONG0 : K_C_PROLONG, nd * (2 * sz + 1) + (double)cl.ndom * sz);
      launch(tl, k_prolong_col<TZ>, grid_of(lv.nact), KTHREADS, lv.nact, lv.pmeta, lv.code, zfin[l + 1], z, c->s);
    }
    for (int k = k0; k < nu; ++k) {
      if (blk) {
        TimedLaunch tl(c, K_C_JACOBI, nd * (3 * sz + 1));
        launch(tl, k_jacobi_block<TZ, TZ>, lv.nact, KTHREADS, lv.nact, lv.meta, lv.code, (const TZ*)lv.r, z, zt, lv.cf<TZ>(),
                                                                 lv.ncell(), lv.ihc<TZ>(), deep ? c->blockw_deep : c->blockw,
                                                                 c->s);
      } else {
        sweep<TZ>(c, l, z, zt, (l == 0 && k == nu - 1) ? 1 : 0, opost[k]);
      }
      std::swap(z, zt);
    }
    zfin[l] = z;
  }
  *zout = zfin[0];
  return 0;
}

int read_scal(mac_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->hs, c->s, sizeof(Scal), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// The state after a finalize launched with to_host: waiting is enough.
int poll_scal(mac_ctx* c) {
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// One CG iteration on the residual r: z = V(r) (the cycle's last sweep gives
// r.z and beta), d_new = z + beta d_old, q = A d_new, alpha, x += alpha d_new,
// r -= alpha q and the convergence test.  d[cur] holds d_old and d[cur ^ 1]
// receives d_new.  *last gets the sequence number of the test's launch.
template <typename TZ>
int cg_iteration(mac_ctx* c, int cur, TZ** zout, long long* last) {
  Level& l0 = c->L[0];
  const double* ih = l0.ih2;
  const double nd = (double)l0.ndom;
  CHK(vcycle<TZ>(c, zout));
  // the search direction is stored in the cycle's precision (FP32 in the
  // mixed method; the d buffers hold n0 doubles, so either type fits)
  using TD = TZ;
  const double sd = sizeof(TD);
  {
    TimedLaunch tl(c, K_PMATVEC, nd * (9 + sizeof(TZ) + 2 * sd));  // z, d_old, code; d_new, q
    launch(tl, k_pmatvec<TZ, TD>, grid_of(l0.nact), KTHREADS, l0.nact, l0.meta, l0.code, *zout, (const TD*)c->d[cur],
                                                                (TD*)c->d[cur ^ 1], c->q, ih[0], ih[1], ih[2], c->partials,
                                                                c->s);
  }
  finalize(c, grid_of(l0.nact), OP_DQ);
  {
    TimedLaunch tl(c, K_UPDATE, nd * (41 + sd));  // code, x, r, d, q; x, r
    launch(tl, k_update<TD>, grid_of(l0.nact), KTHREADS, l0.nact, l0.meta, l0.code, c->x, c->r, (const TD*)c->d[cur ^ 1], c->q,
                                                           c->partials, c->s);
  }
  *last = finalize(c, grid_of(l0.nact), OP_RR, true);
  return 0;
}

// 2^k captured iterations starting at parity cur (captured on first use):
// a launch boundary between graphs costs several microseconds, so a batch
// runs as few graphs as its size allows.
template <typename TZ>
int iteration_graph(mac_ctx* c, int cur, int k, hipGraphExec_t* out) {
  const int gi = std::is_same<TZ, float>::value ? 1 : 0;
  hipGraphExec_t& ge = c->gexec[gi][cur][k];
  if (!ge) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
    int rc = 0;
    for (int i = 0, p = cur; i < (1 << k) && !rc; ++i, p ^= 1) {
      TZ* z = nullptr;
      long long last;
      rc = cg_iteration<TZ>(c, p, &z, &last);
      if (!rc && c->gz[gi] && c->gz[gi] != (const void*)z) rc = fail(MAC_E_STATE, "V-cycle output buffer moved between cycles");
      if (!rc) c->gz[gi] = (const void*)z;
    }
    if (rc) {
      hipStreamEndCapture(c->st, &g);
      if (g) hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(hipStreamEndCapture(c->st, &g));
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      ge = nullptr;
      return fail(MAC_E_HIP, (std::string("hipGraphInstantiate: ") + hipGetErrorString(e)).c_str());
    }
  }
  *out = ge;
  return 0;
}

// ||f - L x||^2 into the state's tmp (and the host mirror); r is kept.
void true_residual(mac_ctx* c) {
  Level& l0 = c->L[0];
  {
    TimedLaunch tl(c, K_INIT, (double)l0.ndom * 17);
    launch(tl, k_resid0, grid_of(l0.nact), KTHREADS, l0.nact, l0.meta, l0.code, c->f, c->x, c->r, l0.ih2[0], l0.ih2[1],
                                                       l0.ih2[2], 0, c->partials);
  }
  finalize(c, grid_of(l0.nact), OP_TRUE, true);
}

template <typename TZ>
int pcg(mac_ctx* c, bool use_guess, bool r_ready, mac_stats* st) {
  Level& l0 = c->L[0];
  const double* ih = l0.ih2;
  size_t n0 = (size_t)c->nb0 * MAC_BC;
  if (!use_guess) clear_x(c);
  const int check = std::max(1, c->prm.check_every);
  const double nd = (double)l0.ndom;
  int restarts = 0;
  bool fellback = false;
  int wasted = 0;  // iterations of a dropped guided start
  int iter_now = 0;  // the device's iteration count as of the last poll
  bool zero_x = !use_guess;  // this (re)start begins from x = 0
  // A guided solve (use_guess) falls back to the zero start when its first
  // residual is no better than the zero start's (or not finite), when it has
  // not converged after stall_iters iterations, or when it ends without
  // converging (max_iter, breakdown).  The zero start gets the full max_iter.
  auto fall_back = [&]() -> int {
    fellback = true;
    if (st) st->fallback_used = 1;
    wasted = c->hs->iter;
    c->hs->iter = 0;
    c->hs->done = 0;
    iter_now = 0;
    zero_x = true;
    HIPCHK(hipMemcpyAsync(&c->s->iter, &c->hs->iter, 2 * sizeof(int), hipMemcpyHostToDevice, c->st));
    clear_x(c);
    return 0;
  };
  for (;;) {
    int cur = 0;  // d[0] holds d_old of the first iteration, not read (beta = 0)
    // (re)start: r = f - L x (the first zero start has r = f and ||r||^2 from
    // the divergence pass)
    if (r_ready) {
      r_ready = false;
    } else {
      {
        TimedLaunch tl(c, K_INIT, nd * 25);
        launch(tl, k_resid0, grid_of(l0.nact), KTHREADS, l0.nact, l0.meta, l0.code, c->f, c->x, c->r, ih[0], ih[1], ih[2], 1,
                                                           c->partials);
      }
      finalize(c, grid_of(l0.nact), OP_RR0, true);
    }
    const bool guided = use_guess && !fellback;
    if (guided) {
      CHK(poll_scal(c));
      if (!(c->hs->rr <= c->hs->ff)) {
        CHK(fall_back());
        continue;
      }
    }
    // Iterations run in batches between host polls of the convergence flag:
    // `check` before the first poll, then as many as the residual decay so
    // far predicts are still needed, so that the last batch ends at the
    // converging iteration (no iteration launched past it, no cycle after
    // it).  From the third solve on a grid (kernel timing off) each
    // iteration is one captured graph launch.
    const bool graphed = !c->timing && c->solves_on_grid >= 2;
    // the first batch: as many iterations as the last solve of this kind
    // took to converge (repeated solves converge alike), else check_every
    int batch = (restarts == 0 && c->last_iters[guided]) ? c->last_iters[guided] : check;
    bool expect = restarts == 0 && c->last_iters[guided] > 0;  // the batch is predicted to converge
    bool have_true = false;
    // the decay rate is measured from this (re)start: its ||r||^2 is known
    // for a guided start (read above) and for a zero start (= ||f||^2, read
    // with the first poll); after a restart from a nonzero x, from the first poll
    const int base_it = iter_now;
    double base_rr = guided ? c->hs->rr : -1.0;
    for (;;) {
      if (guided && c->prm.stall_iters > 0 && batch > c->prm.stall_iters - iter_now) {
        batch = std::max(1, c->prm.stall_iters - iter_now);
        expect = false;
      }
      std::vector<long long> ends;
      const long long first = c->seq;
      for (int k = 0; k < batch;) {
        if (graphed) {
          const int lg = batch - k >= 4 ? 2 : batch - k >= 2 ? 1 : 0;
          hipGraphExec_t ge;
          CHK(iteration_graph<TZ>(c, cur, lg, &ge));
          HIPCHK(hipGraphLaunch(ge, c->st));
          k += 1 << lg;
          if (lg == 0) cur ^= 1;
        } else {
          TZ* z = nullptr;
          long long last;
          CHK(cg_iteration<TZ>(c, cur, &z, &last));
          ends.push_back(last);
          ++k;
          cur ^= 1;
        }
      }
      // a batch predicted to converge takes the true residual of its result
      // along (the residual kernel runs whatever the flag says), so that a
      // converged batch needs no second round trip
      const long long tr_first = c->seq;
      if (expect) true_residual(c);
      const long long tr_last = c->seq - 1;
      HIPCHK(hipGetLastError());
      CHK(poll_scal(c));
      const int ran = c->hs->iter - iter_now;
      iter_now = c->hs->iter;
      if (c->hs->done) {
        // launches after the converging iteration's test exited at once
        if (!graphed && ran < (int)ends.size()) c->noop.push_back({ran > 0 ? ends[ran - 1] + 1 : first, tr_first - 1});
        have_true = expect;
        break;
      }
      if (expect) c->noop.push_back({tr_first, tr_last});  // not used: timed as nothing
      if (guided && c->prm.stall_iters > 0 && iter_now >= c->prm.stall_iters) break;
      if (base_rr < 0.0 && zero_x) base_rr = c->hs->ff;
      batch = check;
      if (base_rr > 0.0 && iter_now > base_it && c->hs->rr > 0.0 && c->hs->rr < base_rr && c->hs->tol2 > 0.0) {
        // iterations still needed at the mean decay rate so far
        const double rate = std::log(c->hs->rr / base_rr) / (iter_now - base_it);  // < 0
        const double need = std::log(c->hs->tol2 / c->hs->rr) / rate;
        // the mean rate since the start overestimates the later rate of an
        // MG-PCG (its first iterations gain the most), so need is rarely too
        // high; a quarter iteration of slack still avoids a batch too many
        batch = std::max(1, (int)std::ceil(need - 0.25));
        expect = true;
      } else {
        expect = false;
        if (base_rr < 0.0) base_rr = c->hs->rr;
      }
    }
    if (use_guess && !fellback && c->hs->done != 1) {
      CHK(fall_back());
      continue;
    }
    int done = c->hs->done;
    if (!have_true) {
      true_residual(c);
      CHK(poll_scal(c));
    }
    double ff = c->hs->ff;
    double tr = ff > 0 ? std::sqrt(c->hs->tmp / ff) : 0.0;
    if (st) {
      st->iters = c->hs->iter + wasted;
      st->rel_resid = ff > 0 ? std::sqrt(c->hs->rr / ff) : 0.0;
      st->true_rel_resid = tr;
    }
    if (done == 1 && tr > c->prm.rtol * 1.0000001 && restarts < 3 && c->hs->iter < c->prm.max_iter) {
      // recursive residual drifted from the true one: restart from x
      ++restarts;
      zero_x = false;
      c->hs->done = 0;
      HIPCHK(hipMemcpyAsync(&c->s->done, &c->hs->done, sizeof(int), hipMemcpyHostToDevice, c->st));
      continue;
    }
    if (st) st->restarts = restarts;
    if (done == 1 && restarts == 0) c->last_iters[use_guess && !fellback] = c->hs->iter;
    if (done == 2) return fail(MAC_E_NOCONV, "PCG reached max_iter before rtol");
    if (done == 3) return fail(MAC_E_NOCONV, "PCG breakdown (d.q <= 0)");
    return 0;
  }
}

int ensure(mac_ctx* c) {
  if (!c) return fail(MAC_E_ARG, "null context");
  if (!c->have_grid) return fail(MAC_E_STATE, "no grid uploaded");
  return 0;
}

}  // namespace

// ============================================================== C-ABI ====

extern "C" {


const char* mac_last_error(void) { return g_err.c_str(); }

void mac_default_params(mac_params* p) {
  p->h[0] = p->h[1] = p->h[2] = 1.0;
  p->rtol = 1e-8;
  p->max_iter = 500;
  p->precision = MAC_FP64;
  p->nu = 2;
  p->omega = 0.9;
  p->check_every = 4;
  p->stall_iters = 0;
  p->smoother = MAC_SMOOTH_CHEBYSHEV;
  p->cheb_lower = 0.25;
  p->nu_coarse = 2;
  p->cheb_lower_coarse = 0.05;
  p->coarse_tail = 1;
  p->viscosity = 0.0;
  p->coarse_inner = 4;
  p->coarse_damp = 0.8;
  p->fuse_coarse = 1;
}

int mac_create(mac_ctx** out, const mac_params* prm) {
  if (!out) return fail(MAC_E_ARG, "null out");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(MAC_E_NODEV, "no HIP device visible");
  if (prm && check_params(prm) != 0) return MAC_E_ARG;
  mac_ctx* c = new mac_ctx();
  if (prm) c->prm = *prm;
  else mac_default_params(&c->prm);
  set_schedule(c);
  hipGetDevice(&c->dev);
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(MAC_E_HIP, "hipStreamCreate failed");
  }
  if (hipHostMalloc((void**)&c->hs, sizeof(Scal), hipHostMallocDefault) != hipSuccess) {
    hipStreamDestroy(c->st);
    delete c;
    return fail(MAC_E_HIP, "hipHostMalloc failed");
  }
  memset(c->hs, 0, sizeof(Scal));
  *out = c;
  return 0;
}

void mac_destroy(mac_ctx* c) {
  if (!c) return;
  hipStreamSynchronize(c->st);
  free_all(c);
  for (auto e : c->evpool) hipEventDestroy(e);
  if (c->hs) hipHostFree(c->hs);
  if (c->pin) hipHostFree(c->pin);
  if (c->csr_ws) hipFree(c->csr_ws);
  if (c->csr_pin) hipHostFree(c->csr_pin);
  hipStreamDestroy(c->st);
  delete c;
}

int mac_set_params(mac_ctx* c, const mac_params* prm) {
  if (!c || !prm) return fail(MAC_E_ARG, "null argument");
  bool hchange = c->have_grid && (prm->h[0] != c->prm.h[0] || prm->h[1] != c->prm.h[1] || prm->h[2] != c->prm.h[2]);
  if (hchange) return fail(MAC_E_STATE, "h cannot change after the grid is uploaded");
  CHK(check_params(prm));
  c->prm = *prm;
  set_schedule(c);
  drop_graphs(c);  // weights, tolerance and sweep counts are baked into them
  return 0;
}

int mac_upload_bricks(mac_ctx* c, int nbricks, const int32_t* keys3, const uint8_t* flags, const double* u, const double* v,
                      const double* w) {
  if (!c || nbricks <= 0 || !keys3 || !flags) return fail(MAC_E_ARG, "bad upload arguments");
  for (int i = 0; i < 3; ++i)
    if (!(c->prm.h[i] > 0)) return fail(MAC_E_ARG, "h must be positive");
  CHK(build_grid(c, nbricks, keys3, flags));
  if (u && v && w) CHK(mac_set_velocity(c, u, v, w));
  return 0;
}

int mac_set_velocity(mac_ctx* c, const double* u, const double* v, const double* w) {
  CHK(ensure(c));
  size_t n = (size_t)c->nb0 * MAC_BC * sizeof(double);
  HIPCHK(hipMemcpyAsync(c->u, u, n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->v, v, n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->w, w, n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_set_real_velocity(mac_ctx* c, const double* u, const double* v, const double* w) {
  CHK(ensure(c));
  if (!u && !v && !w) {
    if (c->ur_own[0] && c->ur == c->ur_own[0]) c->ur = c->vr = c->wr = nullptr;
    return 0;
  }
  if (!u || !v || !w) return fail(MAC_E_ARG, "mac_set_real_velocity: give all three components or none");
  const size_t n = (size_t)c->nb0 * MAC_BC;
  for (int i = 0; i < 3; ++i)
    if (!c->ur_own[i]) CHK(dalloc(c, &c->ur_own[i], n));
  const double* src[3] = {u, v, w};
  for (int i = 0; i < 3; ++i)
    HIPCHK(hipMemcpyAsync(c->ur_own[i], src[i], n * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->ur = c->ur_own[0];
  c->vr = c->ur_own[1];
  c->wr = c->ur_own[2];
  return 0;
}

int mac_get_velocity(mac_ctx* c, double* u, double* v, double* w) {
  CHK(ensure(c));
  size_t n = (size_t)c->nb0 * MAC_BC * sizeof(double);
  HIPCHK(hipMemcpyAsync(u, c->u, n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(v, c->v, n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(w, c->w, n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_get_pressure(mac_ctx* c, double* p) {
  CHK(ensure(c));
  HIPCHK(hipMemcpyAsync(p, c->x, (size_t)c->nb0 * MAC_BC * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_set_pressure(mac_ctx* c, const double* p) {
  CHK(ensure(c));
  HIPCHK(hipMemcpyAsync(c->x, p, (size_t)c->nb0 * MAC_BC * sizeof(double), hipMemcpyHostToDevice, c->st));
  k_mask<<<c->nb0, KTHREADS, 0, c->st>>>(c->nb0, c->L[0].code, c->x);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_save_velocity(mac_ctx* c) {
  CHK(ensure(c));
  size_t n = (size_t)c->nb0 * MAC_BC * sizeof(double);
  HIPCHK(hipMemcpyAsync(c->u0, c->u, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->v0, c->v, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->w0, c->w, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_restore_velocity(mac_ctx* c) {
  CHK(ensure(c));
  size_t n = (size_t)c->nb0 * MAC_BC * sizeof(double);
  HIPCHK(hipMemcpyAsync(c->u, c->u0, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->v, c->v0, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->w, c->w0, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_project(mac_ctx* c, int guess, mac_stats* st) {
  CHK(ensure(c));
  auto t0 = std::chrono::steady_clock::now();
  Level& l0 = c->L[0];
  double inv_h[3];
  for (int i = 0; i < 3; ++i) inv_h[i] = 1.0 / c->prm.h[i];
  if (st) {
    memset(st, 0, sizeof(*st));
    st->levels = (int)c->L.size();
    st->n_unknowns = l0.ndom;
    st->n_bricks = c->nb0;
    st->n_active_bricks = l0.nact;
  }
  Scal init;
  memset(&init, 0, sizeof(init));
  init.max_iter = c->prm.max_iter;
  HIPCHK(hipMemcpyAsync(c->s, &init, sizeof(Scal), hipMemcpyHostToDevice, c->st));
  c->seq = 0;
  const double nd = (double)l0.ndom;
  const bool visc = c->prm.viscosity > 0;
  // a zero start without the viscous right-hand side term takes r = f from
  // the divergence pass (no residual pass of its own)
  const bool r_ready = !visc && !guess;
  if (l0.nact > 0) {
    TimedLaunch tl(c, K_DIV, nd * (r_ready ? 41 : 33));
    launch(tl, k_div, grid_of(l0.nact), KTHREADS, l0.nact, l0.meta, l0.code, c->u, c->v, c->w, inv_h[0], inv_h[1], inv_h[2],
                                                    c->f, nullptr, r_ready ? c->r : nullptr, c->partials);
  }
  if (l0.nact > 0) finalize(c, grid_of(l0.nact), r_ready ? OP_FF_R0 : OP_FF);
  if (visc) {
    if (!c->pD) CHK(dalloc(c, &c->pD, (size_t)c->nb0 * MAC_BC));
    k_air_pressure<<<c->nb0, KTHREADS, 0, c->st>>>(l0.nbr, c->flags, c->ur ? c->ur : c->u, c->vr ? c->vr : c->v,
                                                   c->wr ? c->wr : c->w, inv_h[0], inv_h[1], inv_h[2], c->prm.viscosity, c->pD);
    if (l0.nact > 0) {
      k_dirichlet_rhs<<<grid_of(l0.nact), KTHREADS, 0, c->st>>>(l0.nact, l0.meta, l0.nbr, c->flags, l0.code, c->pD, inv_h[0],
                                                                inv_h[1], inv_h[2], c->f, c->partials);
      finalize(c, grid_of(l0.nact), OP_FF);
    }
    // a guess must not carry the last solve's air pressures into the operator
    if (guess) k_mask<<<c->nb0, KTHREADS, 0, c->st>>>(c->nb0, l0.code, c->x);
  }
  HIPCHK(hipGetLastError());
  auto ts0 = std::chrono::steady_clock::now();
  int rc = 0;
  if (l0.nact > 0) {
    if (c->prm.precision == MAC_MIXED) rc = pcg<float>(c, guess != 0, r_ready, st);
    else rc = pcg<double>(c, guess != 0, r_ready, st);
    c->solves_on_grid += 1;
  } else {
    clear_x(c);
  }
  auto ts1 = std::chrono::steady_clock::now();
  if (rc != 0 && rc != MAC_E_NOCONV) return rc;
  if (visc) k_dirichlet_store<<<c->nb0, KTHREADS, 0, c->st>>>(c->nb0, l0.code, c->pD, c->x);
  if (c->nglist > 0) {
    TimedLaunch tl(c, K_GRAD, nd * 57);
    launch(tl, k_gradient, grid_of(c->nglist), KTHREADS, c->nglist, c->glist, l0.nbr, c->flags, l0.code, c->x, c->u, c->v,
                                                            c->w, inv_h[0], inv_h[1], inv_h[2]);
  }
  if (l0.nact > 0) {
    TimedLaunch tl(c, K_MAXDIV, nd * 25);
    launch(tl, k_maxdiv, grid_of(l0.nact), KTHREADS, l0.nact, l0.meta, l0.code, c->u, c->v, c->w, inv_h[0], inv_h[1], inv_h[2],
                                                       c->partials);
  }
  if (l0.nact > 0) finalize(c, grid_of(l0.nact), OP_MAX);
  HIPCHK(hipGetLastError());
  CHK(read_scal(c));
  auto t1 = std::chrono::steady_clock::now();
  flush_timing(c);
  if (st) {
    st->max_div = c->hs->maxdiv;
    st->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    st->ms_solve = std::chrono::duration<double, std::milli>(ts1 - ts0).count();
  }
  return rc;
}

int mac_divergence(mac_ctx* c, double* div_out) {
  CHK(ensure(c));
  Level& l0 = c->L[0];
  double inv_h[3];
  for (int i = 0; i < 3; ++i) inv_h[i] = 1.0 / c->prm.h[i];
  size_t n0 = (size_t)c->nb0 * MAC_BC;
  HIPCHK(hipMemsetAsync(c->tmpv, 0, n0 * sizeof(double), c->st));
  HIPCHK(hipMemsetAsync(c->s, 0, sizeof(Scal), c->st));
  if (l0.nact > 0)
    k_div<<<grid_of(l0.nact), KTHREADS, 0, c->st>>>(l0.nact, l0.meta, l0.code, c->u, c->v, c->w, inv_h[0], inv_h[1], inv_h[2],
                                                    c->f, c->tmpv, nullptr, c->partials);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(div_out, c->tmpv, n0 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_apply_operator(mac_ctx* c, const double* xh, double* yh) {
  CHK(ensure(c));
  Level& l0 = c->L[0];
  size_t n0 = (size_t)c->nb0 * MAC_BC;
  HIPCHK(hipMemcpyAsync(c->q, xh, n0 * sizeof(double), hipMemcpyHostToDevice, c->st));
  k_mask<<<c->nb0, KTHREADS, 0, c->st>>>(c->nb0, l0.code, c->q);
  HIPCHK(hipMemsetAsync(c->tmpv, 0, n0 * sizeof(double), c->st));
  if (l0.nact > 0) k_apply_A<<<l0.nact, KTHREADS, 0, c->st>>>(l0.act, l0.nbr, l0.code, c->q, c->tmpv, l0.ih2[0], l0.ih2[1], l0.ih2[2]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(yh, c->tmpv, n0 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemsetAsync(c->q, 0, n0 * sizeof(double), c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_gradient_update(mac_ctx* c) {
  CHK(ensure(c));
  Level& l0 = c->L[0];
  double inv_h[3];
  for (int i = 0; i < 3; ++i) inv_h[i] = 1.0 / c->prm.h[i];
  if (c->nglist > 0)
    k_gradient<<<grid_of(c->nglist), KTHREADS, 0, c->st>>>(c->nglist, c->glist, l0.nbr, c->flags, l0.code, c->x, c->u, c->v,
                                                            c->w, inv_h[0], inv_h[1], inv_h[2]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int mac_max_divergence(mac_ctx* c, double* out) {
  CHK(ensure(c));
  Level& l0 = c->L[0];
  double inv_h[3];
  for (int i = 0; i < 3; ++i) inv_h[i] = 1.0 / c->prm.h[i];
  HIPCHK(hipMemsetAsync(c->s, 0, sizeof(Scal), c->st));
  if (l0.nact > 0)
  {
    k_maxdiv<<<grid_of(l0.nact), KTHREADS, 0, c->st>>>(l0.nact, l0.meta, l0.code, c->u, c->v, c->w, inv_h[0], inv_h[1], inv_h[2],
                                                       c->partials);
    finalize(c, grid_of(l0.nact), OP_MAX);
  }
  HIPCHK(hipGetLastError());
  CHK(read_scal(c));
  *out = c->hs->maxdiv;
  return 0;
}

MAC_HIDDEN int mac_internal_set_error(int code, const char* msg) { return fail(code, msg); }
MAC_HIDDEN hipStream_t mac_internal_stream(mac_ctx* c) { return c->st; }
MAC_HIDDEN int mac_internal_device(mac_ctx* c) { return c->dev; }
MAC_HIDDEN int mac_internal_csr_workspace(mac_ctx* c, size_t bytes, void** ws, void** pinned) {
  if (!c->csr_pin) HIPCHK(hipHostMalloc(&c->csr_pin, 256, hipHostMallocDefault));
  if (bytes > c->csr_ws_bytes) {
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->csr_ws) hipFree(c->csr_ws);
    c->csr_ws = nullptr;
    c->csr_ws_bytes = 0;
    HIPCHK(hipMalloc(&c->csr_ws, bytes));
    c->csr_ws_bytes = bytes;
  }
  *ws = c->csr_ws;
  *pinned = c->csr_pin;
  return 0;
}
MAC_HIDDEN void mac_internal_set_real_velocity(mac_ctx* c, const double* u, const double* v, const double* w) {
  c->ur = u;
  c->vr = v;
  c->wr = w;
}

MAC_HIDDEN int mac_internal_build(mac_ctx* c, int nbricks, const int32_t* keys3, const uint8_t* flags) {
  if (!c || nbricks <= 0) return fail(MAC_E_ARG, "bad build arguments");
  return build_grid(c, nbricks, keys3, flags);
}

MAC_HIDDEN int mac_internal_view(mac_ctx* c, mac_grid_view* v) {
  CHK(ensure(c));
  v->st = c->st;
  v->nb = c->nb0;
  for (int i = 0; i < 3; ++i) v->ext[i] = c->ext0[i];
  v->dkeys = c->dkeys0;
  v->flags = c->flags;
  v->code = c->L[0].code;
  v->x = c->x;
  v->f = c->f;
  v->ff = &c->s->ff;
  v->ndom = c->L[0].ndom;
  v->ih2m = (c->L[0].ih2[0] + c->L[0].ih2[1] + c->L[0].ih2[2]) / 3.0;
  return 0;
}

MAC_HIDDEN int mac_internal_rhs(mac_ctx* c) {
  CHK(ensure(c));
  Level& l0 = c->L[0];
  double inv_h[3];
  for (int i = 0; i < 3; ++i) inv_h[i] = 1.0 / c->prm.h[i];
  HIPCHK(hipMemsetAsync(c->s, 0, sizeof(Scal), c->st));
  if (l0.nact > 0) {
    k_div<<<grid_of(l0.nact), KTHREADS, 0, c->st>>>(l0.nact, l0.meta, l0.code, c->u, c->v, c->w, inv_h[0], inv_h[1], inv_h[2],
                                                    c->f, nullptr, nullptr, c->partials);
    finalize(c, grid_of(l0.nact), OP_FF);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

MAC_HIDDEN int mac_internal_load(mac_ctx* c, const double* u, const double* v, const double* w) {
  CHK(ensure(c));
  size_t n = (size_t)c->nb0 * MAC_BC * sizeof(double);
  HIPCHK(hipMemcpyAsync(c->u, u, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->v, v, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->w, w, n, hipMemcpyDeviceToDevice, c->st));
  return 0;
}

MAC_HIDDEN int mac_internal_project(mac_ctx* c, double* u, double* v, double* w, double* p, int guess, mac_stats* st) {
  CHK(ensure(c));
  size_t n = (size_t)c->nb0 * MAC_BC * sizeof(double);
  CHK(mac_internal_load(c, u, v, w));
  if (guess == 1) {
    HIPCHK(hipMemcpyAsync(c->x, p, n, hipMemcpyDeviceToDevice, c->st));
    k_mask<<<c->nb0, KTHREADS, 0, c->st>>>(c->nb0, c->L[0].code, c->x);
  }
  int rc = mac_project(c, guess, st);
  if (rc != 0 && rc != MAC_E_NOCONV) return rc;
  HIPCHK(hipMemcpyAsync(u, c->u, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(v, c->v, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(w, c->w, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(p, c->x, n, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return rc;
}

int mac_enable_kernel_timing(mac_ctx* c, int on) {
  if (!c) return fail(MAC_E_ARG, "null context");
  c->timing = on != 0;
  return 0;
}

int mac_kernel_times(mac_ctx* c, int cap, const char** names, double* total_ms, long long* launches, double* bytes) {
  if (!c) return fail(MAC_E_ARG, "null context");
  for (int i = 0; i < K_N && i < cap; ++i) {
    if (names) names[i] = kNames[i];
    if (total_ms) total_ms[i] = c->kt_ms[i];
    if (launches) launches[i] = c->kt_n[i];
    if (bytes) bytes[i] = c->kt_bytes[i];
  }
  return K_N;
}

int mac_timing_overhead(mac_ctx* c, int n, double* us) {
  if (!c || n <= 0 || !us) return fail(MAC_E_ARG, "bad timing-overhead arguments");
  std::vector<hipEvent_t> ev(2 * (size_t)n);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int i = 0; i < n; ++i) {
    HIPCHK(hipEventRecord(ev[2 * i], c->st));
    k_empty<<<1, 64, 0, c->st>>>();
    HIPCHK(hipEventRecord(ev[2 * i + 1], c->st));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  double tot = 0.0;
  for (int i = 0; i < n; ++i) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
    tot += ms;
  }
  for (auto& e : ev) hipEventDestroy(e);
  *us = tot * 1e3 / n;
  return 0;
}

int mac_reset_kernel_times(mac_ctx* c) {
  if (!c) return fail(MAC_E_ARG, "null context");
  for (int i = 0; i < K_N; ++i) {
    c->kt_ms[i] = 0;
    c->kt_n[i] = 0;
    c->kt_bytes[i] = 0;
  }
  return 0;
}

int mac_sync(mac_ctx* c) {
  if (!c) return fail(MAC_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

}  // extern "C"
