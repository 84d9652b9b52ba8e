// mac_sim.hip -- the per-timestep MAC-grid loop on a sparse device brick
// grid: host-side brick bookkeeping and the C-ABI (mac_sim_*).
//
// The host owns the authoritative list of chunk keys, kept in ascending
// (z, y, x) order so that brick ids -- and with them every reduction order --
// are reproducible.  All cell and particle data live on the device; the only
// host traffic per step is the list of bricks to create (keys of particle
// bricks that are missing, plus per-brick flags) and the per-brick cell
// counts after the delete pass.  Lookups from particles go through a device
// hash of chunk keys (sim_kernels.h), stencils through the 27-neighbour table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <set>
#include <string>
#include <tuple>
#include <iterator>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "mac_internal.h"
#include "sim_kernels.h"
#include "surface_bc.h"

namespace {

#define SCHK(expr)                                                                               \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return mac_internal_set_error(MAC_E_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)
#define SRC(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != 0) return rc_; \
  } while (0)

struct Key3 {
  int x, y, z;
  bool operator<(const Key3& o) const { return std::tie(z, y, x) < std::tie(o.z, o.y, o.x); }
  bool operator==(const Key3& o) const { return x == o.x && y == o.y && z == o.z; }
};

inline int ceil_pow2(int v) {
  int c = 1;
  while (c < v) c <<= 1;
  return c;
}

}  // namespace

struct mac_sim {
  mac_sim_params prm;
  hipStream_t st = nullptr;
  mac_ctx* proj = nullptr;
  std::vector<Key3> keys;
  int nb = 0;
  // device brick arrays (capacity capb)
  int* d_keys = nullptr;
  uint8_t* flags = nullptr;
  int8_t* layer = nullptr;
  int16_t* depth = nullptr;  // Data_Grid_Depth_Empty: -1 = empty
  // speed slots 0..2 (Speed_GetRef(i) of the reference's 3-slot chunks; slot 2
  // is used by the RK4 "all" integrator only)
  double* S[3][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  double* A[3] = {nullptr, nullptr, nullptr};
  double* P = nullptr;
  double* Pprev = nullptr;  // the pressure the last projection started from
  int* nbr = nullptr;
  uint64_t* hk = nullptr;
  int* hv = nullptr;
  int hcap = 0;
  // particles (SoA, capacity capp) and the second buffer for compaction
  double *px = nullptr, *py = nullptr, *pz = nullptr, *qx = nullptr, *qy = nullptr, *qz = nullptr;
  // position slots 1 and 2 of the RK4 "all" integrator (Data_Particle m_pos[1..2])
  double* P1[3] = {nullptr, nullptr, nullptr};
  double* P2[3] = {nullptr, nullptr, nullptr};
  int* d_dest = nullptr;   // per particle: kept (1) or erased (0) by an erasing pass; dropped-particle holes
  int* d_holes = nullptr;  // the erased places being refilled
  int np = 0, capp = 0;
  // scratch
  int* d_int = nullptr;  // [0] particle count, [1] missing count, [2] overflow, [3] kept / dropped, [4..5] compaction
  int* d_any = nullptr;  // per-round change flags of the depth search
  int* d_missing = nullptr;
  int* d_missing2 = nullptr;  // absent neighbour bricks of flagged bricks (keys, each once)
  int maxmissing2 = 0;
  uint64_t* d_cset = nullptr;  // their de-duplication set (2 * maxmissing2 slots)
  int maxmissing = 0;
  int* d_bflag = nullptr;
  int* d_from = nullptr;
  double* d_part = nullptr;
  unsigned long long tick = 0;
  bool grid_ready = false;
  // allocation sizes: the per-cell arrays above hold cur_cap bricks; a second
  // set of the same arrays (spare, spare_cap bricks) is the target of the
  // next relayout, after which the two swap, so a changing brick set costs
  // no allocation until it outgrows them
  int cur_cap = 0, spare_cap = 0;
  struct Cells {
    uint8_t* flags = nullptr;
    int8_t* layer = nullptr;
    int16_t* depth = nullptr;
    double* S[3][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    double* A[3] = {nullptr, nullptr, nullptr};
    double* P = nullptr;
    double* Pprev = nullptr;
  } spare;
  int keys_cap = 0, hash_cap = 0;
  uint8_t* d_front = nullptr;  // the extrapolation's two frontier masks
  size_t front_cap = 0;
  // the free-surface condition's staging (3 values and 3 pending bytes per
  // cell, the bytes zero between calls) and per-brick activity
  double* bc_tmp = nullptr;
  uint8_t* bc_pend = nullptr;
  uint8_t* bc_cls = nullptr;   // per cell class bytes
  uint8_t* bc_brick = nullptr;  // [0, cap): summary, [cap, 2 cap): active
  size_t bc_cap = 0;            // bricks
  // starting iterate of the projections: 0 zero, 1 last pressure, 2 learned
  int guess_mode = 0;
  mac_unet* net = nullptr;
  float out_scale = 1.f;
};

namespace {

mac_sim::Cells cur_cells(const mac_sim* s) {
  mac_sim::Cells c;
  c.flags = s->flags;
  c.layer = s->layer;
  c.depth = s->depth;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) c.S[a][i] = s->S[a][i];
  for (int i = 0; i < 3; ++i) c.A[i] = s->A[i];
  c.P = s->P;
  c.Pprev = s->Pprev;
  return c;
}
void set_cur_cells(mac_sim* s, const mac_sim::Cells& c) {
  s->flags = c.flags;
  s->layer = c.layer;
  s->depth = c.depth;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) s->S[a][i] = c.S[a][i];
  for (int i = 0; i < 3; ++i) s->A[i] = c.A[i];
  s->P = c.P;
  s->Pprev = c.Pprev;
}
void free_cells(mac_sim::Cells& c) {
  hipFree(c.flags);
  hipFree(c.layer);
  hipFree(c.depth);
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) hipFree(c.S[a][i]);
  for (int i = 0; i < 3; ++i) hipFree(c.A[i]);
  hipFree(c.P);
  hipFree(c.Pprev);
  c = mac_sim::Cells();
}
int alloc_cells(mac_sim::Cells& c, int cap) {
  const size_t cells = (size_t)std::max(cap, 1) * MAC_BC;
  SCHK(hipMalloc(&c.flags, cells));
  SCHK(hipMalloc(&c.layer, cells));
  SCHK(hipMalloc(&c.depth, cells * sizeof(int16_t)));
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) SCHK(hipMalloc(&c.S[a][i], cells * sizeof(double)));
  for (int i = 0; i < 3; ++i) SCHK(hipMalloc(&c.A[i], cells * sizeof(double)));
  SCHK(hipMalloc(&c.P, cells * sizeof(double)));
  SCHK(hipMalloc(&c.Pprev, cells * sizeof(double)));
  return 0;
}

void sim_free_bricks(mac_sim* s) {
  mac_sim::Cells c = cur_cells(s);
  free_cells(c);
  set_cur_cells(s, c);
  free_cells(s->spare);
  s->cur_cap = s->spare_cap = 0;
  hipFree(s->d_keys);
  hipFree(s->nbr);
  hipFree(s->d_bflag);
  hipFree(s->d_from);
  hipFree(s->d_part);
  s->d_keys = s->nbr = s->d_bflag = s->d_from = nullptr;
  s->d_part = nullptr;
  s->keys_cap = 0;
  hipFree(s->hk);
  hipFree(s->hv);
  s->hk = nullptr;
  s->hv = nullptr;
  s->hash_cap = 0;
}

// Replace the brick set by `nk` (sorted, unique).  Bricks present before keep
// their data; new ones start as the prototype.  Rebuilds the hash and the
// neighbour table.  The cells are relaid into the spare arrays, which then
// become the current ones.
int sim_set_bricks(mac_sim* s, const std::vector<Key3>& nk) {
  if ((int)nk.size() > s->prm.max_bricks) return mac_internal_set_error(MAC_E_ARG, "brick pool exhausted (max_bricks)");
  const int nbn = (int)nk.size();
  std::vector<int> from(std::max(nbn, 1), -1);
  {
    size_t i = 0, j = 0;
    while (j < nk.size()) {
      while (i < s->keys.size() && s->keys[i] < nk[j]) ++i;
      from[j] = (i < s->keys.size() && s->keys[i] == nk[j]) ? (int)i : -1;
      ++j;
    }
  }
  const int grow = std::min(s->prm.max_bricks, std::max(nbn + nbn / 4 + 8, 16));
  if (s->spare_cap < nbn) {
    free_cells(s->spare);
    SRC(alloc_cells(s->spare, grow));
    s->spare_cap = grow;
  }
  if (s->keys_cap < nbn || !s->d_keys) {
    hipFree(s->d_keys);
    hipFree(s->nbr);
    hipFree(s->d_bflag);
    hipFree(s->d_from);
    hipFree(s->d_part);
    SCHK(hipMalloc(&s->d_keys, sizeof(int) * 3 * grow));
    SCHK(hipMalloc(&s->nbr, sizeof(int) * 27 * grow));
    SCHK(hipMalloc(&s->d_bflag, sizeof(int) * grow));
    SCHK(hipMalloc(&s->d_from, sizeof(int) * grow));
    SCHK(hipMalloc(&s->d_part, sizeof(double) * grow));
    s->keys_cap = grow;
  }
  std::vector<int> hk3(3 * std::max(nbn, 1), 0);
  for (int b = 0; b < nbn; ++b) {
    hk3[3 * b] = nk[b].x;
    hk3[3 * b + 1] = nk[b].y;
    hk3[3 * b + 2] = nk[b].z;
  }
  SCHK(hipMemcpyAsync(s->d_keys, hk3.data(), sizeof(int) * hk3.size(), hipMemcpyHostToDevice, s->st));
  SCHK(hipMemcpyAsync(s->d_from, from.data(), sizeof(int) * from.size(), hipMemcpyHostToDevice, s->st));
  mac_sim::Cells& to = s->spare;
  if (nbn > 0) {
    const int* d_from = s->d_from;
    k_relayout_u8<<<nbn, 256, 0, s->st>>>(nbn, d_from, s->flags, to.flags, (uint8_t)0);
    k_relayout_u8<<<nbn, 256, 0, s->st>>>(nbn, d_from, (const uint8_t*)s->layer, (uint8_t*)to.layer, (uint8_t)0xFF);
    k_relayout_i16<<<nbn, 256, 0, s->st>>>(nbn, d_from, s->depth, to.depth, (int16_t)-1);
    RelayoutF64 arr;
    int na = 0;
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < 3; ++i) {
        arr.src[na] = s->S[a][i];
        arr.dst[na++] = to.S[a][i];
      }
    for (int i = 0; i < 3; ++i) {
      arr.src[na] = s->A[i];
      arr.dst[na++] = to.A[i];
    }
    arr.src[na] = s->P;
    arr.dst[na++] = to.P;
    k_relayout_f64<<<dim3(nbn, na), 256, 0, s->st>>>(nbn, d_from, arr);
  }
  SCHK(hipGetLastError());
  {
    mac_sim::Cells old = cur_cells(s);
    set_cur_cells(s, to);
    s->spare = old;
    std::swap(s->cur_cap, s->spare_cap);
  }
  s->keys = nk;
  s->nb = nbn;
  // hash (load factor <= 1/2) and neighbour table
  s->hcap = ceil_pow2(std::max(2 * nbn, 64));
  if (s->hash_cap < s->hcap) {
    hipFree(s->hk);
    hipFree(s->hv);
    SCHK(hipMalloc(&s->hk, sizeof(uint64_t) * s->hcap));
    SCHK(hipMalloc(&s->hv, sizeof(int) * s->hcap));
    s->hash_cap = s->hcap;
  }
  k_hash_clear<<<(s->hcap + 255) / 256, 256, 0, s->st>>>(s->hcap, s->hk, s->hv);
  if (nbn > 0) {
    k_hash_insert<<<(nbn + 255) / 256, 256, 0, s->st>>>(nbn, s->d_keys, s->hk, s->hv, s->hcap);
    k_build_nbr<<<(nbn * 27 + 255) / 256, 256, 0, s->st>>>(nbn, s->d_keys, s->hk, s->hv, s->hcap, s->nbr);
  }
  SCHK(hipGetLastError());
  SCHK(hipStreamSynchronize(s->st));
  s->grid_ready = false;  // the projection hierarchy must be rebuilt
  return 0;
}

int sim_alloc_particles(mac_sim* s, int cap) {
  s->capp = std::max(cap, 1);
  for (double** p : {&s->px, &s->py, &s->pz, &s->qx, &s->qy, &s->qz, &s->P1[0], &s->P1[1], &s->P1[2], &s->P2[0], &s->P2[1],
                     &s->P2[2]})
    SCHK(hipMalloc(p, sizeof(double) * s->capp));
  SCHK(hipMalloc(&s->d_dest, sizeof(int) * s->capp));
  SCHK(hipMalloc(&s->d_holes, sizeof(int) * s->capp));
  return 0;
}

// Compacts slot 0 after a pass that marked its dropped particles with a NaN x
// and counted them in d_int[3]: the kept ones fill the dropped ones' places
// below the new count.
int sim_compact_dropped(mac_sim* s) {
  int dropped = 0;
  SCHK(hipMemcpyAsync(&dropped, s->d_int + 3, sizeof(int), hipMemcpyDeviceToHost, s->st));
  SCHK(hipStreamSynchronize(s->st));
  if (dropped < 0 || dropped > s->np) return mac_internal_set_error(MAC_E_HIP, "dropped-particle count out of range");
  if (dropped == 0) return 0;
  const int n_new = s->np - dropped;
  SCHK(hipMemsetAsync(s->d_int + 4, 0, 2 * sizeof(int), s->st));
  if (n_new > 0) {
    k_list_holes<<<(n_new + 255) / 256, 256, 0, s->st>>>(n_new, s->px, s->d_dest, s->d_int + 4);
    k_fill_holes<<<(dropped + 255) / 256, 256, 0, s->st>>>(n_new, s->np, s->px, s->py, s->pz, s->d_dest, s->d_int + 5);
    SCHK(hipGetLastError());
  }
  s->np = n_new;
  return 0;
}

// Depth search of Algorithms_Update_CellType_Layer_Depth.h:96-133: batches
// of four relaxation launches (a launch after one that changed nothing
// returns at once), each batch's change flags read back once, until a
// launch changes nothing; then the reached cells' types.
int sim_depth_search(mac_sim* s) {
  const mac_sim_params& P = s->prm;
  constexpr int batch = 4;
  for (int round = 0;; ++round) {
    if (round > 100000) return mac_internal_set_error(MAC_E_STATE, "depth search did not settle");
    SCHK(hipMemsetAsync(s->d_any, 0, (batch + 1) * sizeof(int), s->st));
    for (int k = 0; k < batch; ++k)
      k_depth_relax<<<s->nb, 256, 0, s->st>>>(s->nb, s->nbr, s->flags, s->layer, s->depth, k ? s->d_any + k - 1 : nullptr,
                                               s->d_any + k, s->d_any + batch);
    SCHK(hipGetLastError());
    int any[batch + 1];
    SCHK(hipMemcpyAsync(any, s->d_any, sizeof(any), hipMemcpyDeviceToHost, s->st));
    SCHK(hipStreamSynchronize(s->st));
    if (any[batch]) return mac_internal_set_error(MAC_E_ARG, "fluid deeper than 32766 cells");
    bool settled = false;
    for (int k = 0; k < batch; ++k) settled |= !any[k];
    if (settled) break;
  }
  k_depth_types<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->layer, s->depth, P.depth_interior);
  SCHK(hipGetLastError());
  return 0;
}

// Compacts the position slots in `sl` after k_mark_particles_keep marked the
// kept particles in d_dest and counted the others in d_int[3].
int sim_compact_marked(mac_sim* s, const PartSlots& sl) {
  int dropped = 0;
  SCHK(hipMemcpyAsync(&dropped, s->d_int + 3, sizeof(int), hipMemcpyDeviceToHost, s->st));
  SCHK(hipStreamSynchronize(s->st));
  if (dropped < 0 || dropped > s->np) return mac_internal_set_error(MAC_E_HIP, "erased-particle count out of range");
  if (dropped == 0) return 0;
  const int n_new = s->np - dropped;
  SCHK(hipMemsetAsync(s->d_int + 4, 0, 2 * sizeof(int), s->st));
  if (n_new > 0) {
    k_list_holes_marked<<<(n_new + 255) / 256, 256, 0, s->st>>>(n_new, s->d_dest, s->d_holes, s->d_int + 4);
    k_fill_holes_marked<<<(dropped + 255) / 256, 256, 0, s->st>>>(n_new, s->np, s->d_dest, s->d_holes, s->d_int + 5, sl);
    SCHK(hipGetLastError());
  }
  s->np = n_new;
  return 0;
}

// full: the step's cell-type chain (particles created, floor applied);
// otherwise the chain inside an RK4 stage (Algorithms_Create_Interior_Depth:
// no particles created, no floor), with the particle keys taken from
// position slot ks (0 or 2) and all three slots compacted together.
int sim_topology(mac_sim* s, bool full = true, int ks = 0) {
  const mac_sim_params& P = s->prm;
  SimIH ih{{1.0 / P.h[0], 1.0 / P.h[1], 1.0 / P.h[2]}};
  const bool dmode = P.depth_interior > 0;
  double** kp[3] = {ks == 2 ? &s->P2[0] : &s->px, ks == 2 ? &s->P2[1] : &s->py, ks == 2 ? &s->P2[2] : &s->pz};
  const int floor_axis = full ? P.floor_axis : -1;
  // Algorithms_Initialize_MacCell / Algorithms_Initialize_Layer_And_Depth
  if (s->nb > 0 && dmode)
    k_reset_layer_depth<<<1024, 256, 0, s->st>>>((size_t)s->nb * MAC_BC, s->flags, s->layer, s->depth);
  else if (s->nb > 0)
    k_reset_layer<<<1024, 256, 0, s->st>>>((size_t)s->nb * MAC_BC, s->layer);
  // bricks needed this step: missing particle bricks, and the 27-neighbourhood
  // of every brick that will hold a layer-0 cell (particles, inflow); the air
  // layers reach at most layer_max + 1 <= 8 cells, so they stay inside it
  SCHK(hipMemsetAsync(s->d_int, 0, 4 * sizeof(int), s->st));
  if (s->nb > 0) {
    SCHK(hipMemsetAsync(s->d_bflag, 0, sizeof(int) * s->nb, s->st));
    k_inflow_bricks<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->d_bflag, dmode ? 1 : 0);
  }
  if (s->np > 0)
    k_particle_bricks<<<(s->np + 255) / 256, 256, 0, s->st>>>(s->np, *kp[0], *kp[1], *kp[2], ih, s->hk, s->hv, s->hcap, s->d_bflag,
                                                             s->d_missing, s->d_int + 1, s->maxmissing, floor_axis,
                                                             P.floor_value);
  // the flagged bricks' absent neighbours, found and de-duplicated on the device
  if (s->nb > 0) {
    k_set_clear<<<256, 256, 0, s->st>>>(2 * s->maxmissing2, s->d_cset);
    k_missing_neighbours<<<(s->nb * 27 + 255) / 256, 256, 0, s->st>>>(s->nb, s->d_bflag, s->nbr, s->d_keys, s->d_cset,
                                                                      2 * s->maxmissing2, s->d_missing2, s->d_int + 3,
                                                                      s->maxmissing2);
  }
  SCHK(hipGetLastError());
  int cnt[4];
  SCHK(hipMemcpyAsync(cnt, s->d_int, sizeof(cnt), hipMemcpyDeviceToHost, s->st));
  SCHK(hipStreamSynchronize(s->st));
  if (cnt[1] > s->maxmissing) return mac_internal_set_error(MAC_E_ARG, "too many particles outside the grid in one step");
  if (cnt[3] > s->maxmissing2) return mac_internal_set_error(MAC_E_ARG, "too many bricks to create in one step");
  const int nmiss = cnt[1], nnb = cnt[3];
  std::vector<Key3> add(nnb);
  if (nnb > 0) {
    std::vector<int> nbm(3 * nnb);
    SCHK(hipMemcpyAsync(nbm.data(), s->d_missing2, sizeof(int) * 3 * nnb, hipMemcpyDeviceToHost, s->st));
    SCHK(hipStreamSynchronize(s->st));
    for (int i = 0; i < nnb; ++i) add[i] = {nbm[3 * i], nbm[3 * i + 1], nbm[3 * i + 2]};
  }
  if (nmiss > 0) {
    // bricks of particles outside the grid, with their whole neighbourhood
    std::vector<int> miss(3 * nmiss);
    SCHK(hipMemcpyAsync(miss.data(), s->d_missing, sizeof(int) * 3 * nmiss, hipMemcpyDeviceToHost, s->st));
    SCHK(hipStreamSynchronize(s->st));
    std::unordered_set<uint64_t> have;
    have.reserve(2 * (s->keys.size() + add.size() + 27 * (size_t)nmiss));
    for (const Key3& k : s->keys) have.insert(sim_pack(k.x, k.y, k.z));
    for (const Key3& k : add) have.insert(sim_pack(k.x, k.y, k.z));
    for (int i = 0; i < nmiss; ++i)
      for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
          for (int ox = -1; ox <= 1; ++ox) {
            const int x = miss[3 * i] + ox, y = miss[3 * i + 1] + oy, z = miss[3 * i + 2] + oz;
            if (have.insert(sim_pack(x, y, z)).second) add.push_back({x, y, z});
          }
  }
  if (!add.empty()) {
    std::sort(add.begin(), add.end());
    std::vector<Key3> nk;
    nk.reserve(s->keys.size() + add.size());
    std::merge(s->keys.begin(), s->keys.end(), add.begin(), add.end(), std::back_inserter(nk));
    SRC(sim_set_bricks(s, nk));
  }
  // Algorithms_Layer_Initial_With_Particle, or Algorithm_DelParticle +
  // Algorithms_Layer_Initial_With_Particle_Depth (particles erased: the kept
  // ones are appended to the second buffer, whose count starts the fill)
  const bool erase = dmode || floor_axis >= 0;
  if (s->np > 0 && erase) {
    SCHK(hipMemsetAsync(s->d_int + 3, 0, sizeof(int), s->st));
    k_mark_particles_keep<<<(s->np + 255) / 256, 256, 0, s->st>>>(s->np, *kp[0], *kp[1], *kp[2], ih, s->hk, s->hv, s->hcap, s->flags,
                                                                 s->layer, dmode ? s->depth : nullptr, P.depth_remove,
                                                                 floor_axis, P.floor_value, s->d_dest, s->d_int + 3);
    SCHK(hipGetLastError());
    // inside a stage the other two position slots follow the same erasures
    PartSlots sl;
    double* slots[3][3] = {{s->px, s->py, s->pz}, {s->P1[0], s->P1[1], s->P1[2]}, {s->P2[0], s->P2[1], s->P2[2]}};
    sl.n = full ? 1 : 3;
    for (int k = 0; k < 9; ++k) sl.p[k] = full ? (k < 3 ? *kp[k] : nullptr) : slots[k / 3][k % 3];
    SRC(sim_compact_marked(s, sl));
  } else if (s->np > 0) {
    k_mark_particles<<<(s->np + 255) / 256, 256, 0, s->st>>>(s->np, *kp[0], *kp[1], *kp[2], ih, s->hk, s->hv, s->hcap, s->flags,
                                                            s->layer);
  }
  // the particle count starts the fill's appends
  SCHK(hipMemcpyAsync(s->d_int, &s->np, sizeof(int), hipMemcpyHostToDevice, s->st));
  // Algorithms_Create_Fluid_Particle(_Depth), or inside a stage
  // Algorithms_Create_Interior_Depth (no particles)
  if (s->nb > 0 && (P.particles_per_inflow_cell > 0 || P.fill_cube_n > 0 || dmode))
    k_inflow_fill<<<s->nb, 256, 0, s->st>>>(s->nb, s->d_keys, s->flags, s->layer, P.h[0], P.h[1], P.h[2],
                                             P.particles_per_inflow_cell, P.seed, s->tick, s->px, s->py, s->pz, s->d_int,
                                             s->capp, s->d_int + 2, dmode ? s->depth : nullptr, P.depth_remove, full ? 1 : 0,
                                             P.fill_cube_n);
  // Algorithms_Update_CellType_Layer(_Depth): surface depth before the layers
  if (s->nb > 0 && dmode) k_depth_surface<<<s->nb, 256, 0, s->st>>>(s->nb, s->nbr, s->flags, s->layer, s->depth);
  if (s->nb > 0)
    for (int lay = 1; lay <= P.layer_max; ++lay) k_layer_round<<<s->nb, 256, 0, s->st>>>(s->nb, s->nbr, s->flags, s->layer, lay);
  if (s->nb > 0) k_layer_extra<<<s->nb, 256, 0, s->st>>>(s->nb, s->nbr, s->flags, s->layer, P.layer_max);
  if (s->nb > 0 && dmode) SRC(sim_depth_search(s));
  // Algorithms_Delete_MacCell
  if (s->nb > 0) k_delete_cells<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->layer, s->d_bflag);
  SCHK(hipGetLastError());
  SCHK(hipMemcpyAsync(cnt, s->d_int, sizeof(cnt), hipMemcpyDeviceToHost, s->st));
  std::vector<int> alive(s->nb);
  if (s->nb > 0) SCHK(hipMemcpyAsync(alive.data(), s->d_bflag, sizeof(int) * s->nb, hipMemcpyDeviceToHost, s->st));
  SCHK(hipStreamSynchronize(s->st));
  if (cnt[2]) return mac_internal_set_error(MAC_E_ARG, "particle pool exhausted (max_particles)");
  s->np = cnt[0];
  if (full) s->tick++;
  // free the bricks whose last cell was erased (ChunkHashTable.tpp:133-161)
  std::vector<Key3> keep;
  keep.reserve(s->nb);
  for (int b = 0; b < s->nb; ++b)
    if (alive[b] > 0) keep.push_back(s->keys[b]);
  if ((int)keep.size() != s->nb) SRC(sim_set_bricks(s, keep));
  s->grid_ready = false;
  return 0;
}

int sim_ensure_projection(mac_sim* s) {
  if (s->grid_ready) return 0;
  if (s->nb == 0) return mac_internal_set_error(MAC_E_STATE, "empty grid");
  SCHK(hipStreamSynchronize(s->st));
  std::vector<int32_t> k3(3 * s->nb);
  for (int b = 0; b < s->nb; ++b) {
    k3[3 * b] = s->keys[b].x;
    k3[3 * b + 1] = s->keys[b].y;
    k3[3 * b + 2] = s->keys[b].z;
  }
  SRC(mac_internal_build(s->proj, s->nb, k3.data(), s->flags));
  s->grid_ready = true;
  return 0;
}

// rk4: the RK4 "all" stages -- forces added to the acceleration held, in
// Jet3's order (gravity, convection, viscosity)
int sim_forces(mac_sim* s, int slot, bool rk4 = false) {
  if (s->nb == 0) return 0;
  const mac_sim_params& P = s->prm;
  SimForce F;
  for (int i = 0; i < 3; ++i) {
    F.ih[i] = 1.0 / P.h[i];
    F.g[i] = P.gravity[i];
  }
  F.nu = P.viscosity;
  F.accumulate = rk4 ? 1 : 0;
  F.conv_first = rk4 ? 1 : 0;
  k_forces<<<s->nb, 256, 0, s->st>>>(s->nb, s->nbr, s->flags, s->layer, s->S[slot][0], s->S[slot][1], s->S[slot][2], s->A[0],
                                      s->A[1], s->A[2], F);
  SCHK(hipGetLastError());
  return 0;
}

int sim_project(mac_sim* s, int slot, mac_stats* st) {
  // the projection context works on its own stream: the velocity written by
  // this stream's kernels must be complete before it is copied over
  SCHK(hipStreamSynchronize(s->st));
  SRC(sim_ensure_projection(s));
  SCHK(hipMemcpyAsync(s->Pprev, s->P, (size_t)s->nb * MAC_BC * sizeof(double), hipMemcpyDeviceToDevice, s->st));
  SCHK(hipStreamSynchronize(s->st));
  int guess = 0;
  if (s->guess_mode == 1) {
    guess = 1;
  } else if (s->guess_mode == 2) {
    SRC(mac_internal_load(s->proj, s->S[slot][0], s->S[slot][1], s->S[slot][2]));
    SRC(mac_internal_unet_guess(s->net, s->proj, s->Pprev, s->out_scale));
    guess = 2;
  }
  return mac_internal_project(s->proj, s->S[slot][0], s->S[slot][1], s->S[slot][2], s->P, guess, st);
}

SimGrid sim_grid(mac_sim* s, int slot) {
  SimGrid G;
  G.hk = s->hk;
  G.hv = s->hv;
  G.cap = s->hcap;
  G.nbr27 = s->nbr;
  for (int i = 0; i < 3; ++i) {
    G.U[i] = s->S[slot][i];
    G.ih[i] = 1.0 / s->prm.h[i];
  }
  return G;
}

}  // namespace

extern "C" {

void mac_sim_default_params(mac_sim_params* p) {
  memset(p, 0, sizeof(*p));
  p->h[0] = p->h[1] = p->h[2] = 0.01;
  p->gravity[2] = -9.81;
  p->viscosity = 0.0;
  p->cfl = 0.5;
  p->dt_min = 0.0;
  p->dt_max = 1e30;
  p->layer_max = 2;
  p->integrator = 0;
  p->particle_rk = 2;
  p->particles_per_inflow_cell = 8;
  p->seed = 0;
  p->max_particles = 1 << 22;
  p->max_bricks = 1 << 16;
  p->depth_interior = 0;
  p->depth_remove = 0;
  p->floor_axis = -1;
  p->floor_value = 0.0;
  mac_default_params(&p->proj);
}

int mac_sim_create(mac_sim** out, const mac_sim_params* p) {
  if (!out || !p) return mac_internal_set_error(MAC_E_ARG, "null argument");
  *out = nullptr;
  if (p->layer_max < 1 || p->layer_max > 7) return mac_internal_set_error(MAC_E_ARG, "layer_max must be in 1..7");
  if (p->integrator < 0 || p->integrator > 2) return mac_internal_set_error(MAC_E_ARG, "integrator must be 0, 1 or 2");
  if (p->fill_cube_n < 0 || p->fill_cube_n > 3) return mac_internal_set_error(MAC_E_ARG, "fill_cube_n must be in 0..3");
  if (!std::isfinite(p->inflow_amp) || (p->inflow_amp != 0.0 && p->inflow_period < 1))
    return mac_internal_set_error(MAC_E_ARG, "inflow_amp must be finite, with inflow_period >= 1 when it is not 0");
  if (p->surface_bc != 0 && p->surface_bc != 1) return mac_internal_set_error(MAC_E_ARG, "surface_bc must be 0 or 1");
  if (p->depth_interior < 0 || p->depth_interior > 32767) return mac_internal_set_error(MAC_E_ARG, "depth_interior must be in 0..32767");
  if (p->floor_axis < -1 || p->floor_axis > 2) return mac_internal_set_error(MAC_E_ARG, "floor_axis must be -1, 0, 1 or 2");
  for (int i = 0; i < 3; ++i)
    if (!(p->h[i] > 0)) return mac_internal_set_error(MAC_E_ARG, "h must be positive");
  mac_sim* s = new mac_sim();
  s->prm = *p;
  mac_params pp = p->proj;
  for (int i = 0; i < 3; ++i) pp.h[i] = p->h[i];
  int rc = mac_create(&s->proj, &pp);
  if (rc) {
    delete s;
    return rc;
  }
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
    mac_destroy(s->proj);
    delete s;
    return mac_internal_set_error(MAC_E_HIP, "hipStreamCreate failed");
  }
  s->maxmissing = 1 << 16;
  s->maxmissing2 = 1 << 18;
  if (hipMalloc(&s->d_int, 8 * sizeof(int)) != hipSuccess || hipMalloc(&s->d_any, 64 * sizeof(int)) != hipSuccess ||
      hipMalloc(&s->d_missing, sizeof(int) * 3 * s->maxmissing) != hipSuccess ||
      hipMalloc(&s->d_missing2, sizeof(int) * 3 * s->maxmissing2) != hipSuccess ||
      hipMalloc(&s->d_cset, sizeof(uint64_t) * 2 * s->maxmissing2) != hipSuccess ||
      sim_alloc_particles(s, p->max_particles) != 0) {
    mac_sim_destroy(s);
    return mac_internal_set_error(MAC_E_HIP, "allocation failed");
  }
  *out = s;
  return 0;
}

void mac_sim_destroy(mac_sim* s) {
  if (!s) return;
  if (s->st) hipStreamSynchronize(s->st);
  sim_free_bricks(s);
  for (double* p : {s->px, s->py, s->pz, s->qx, s->qy, s->qz, s->P1[0], s->P1[1], s->P1[2], s->P2[0], s->P2[1], s->P2[2]}) hipFree(p);
  hipFree(s->d_dest);
  hipFree(s->d_holes);
  hipFree(s->d_int);
  hipFree(s->d_any);
  hipFree(s->d_front);
  hipFree(s->bc_tmp);
  hipFree(s->bc_pend);
  hipFree(s->bc_brick);
  hipFree(s->bc_cls);
  hipFree(s->d_missing);
  hipFree(s->d_missing2);
  hipFree(s->d_cset);
  if (s->proj) mac_destroy(s->proj);
  if (s->st) hipStreamDestroy(s->st);
  delete s;
}

int mac_sim_upload(mac_sim* s, int nbricks, const int32_t* keys3, const uint8_t* flags, const int8_t* layer, const double* u,
                   const double* v, const double* w, int nparticles, const double* pxyz) {
  if (!s || nbricks < 0 || (nbricks > 0 && (!keys3 || !flags)) || nparticles < 0 || (nparticles > 0 && !pxyz))
    return mac_internal_set_error(MAC_E_ARG, "bad upload arguments");
  if (nparticles > s->capp) return mac_internal_set_error(MAC_E_ARG, "more particles than max_particles");
  // sort the input bricks; keys must be unique
  std::vector<int> order(nbricks);
  for (int b = 0; b < nbricks; ++b) order[b] = b;
  auto kof = [&](int b) { return Key3{keys3[3 * b], keys3[3 * b + 1], keys3[3 * b + 2]}; };
  std::sort(order.begin(), order.end(), [&](int a, int b) { return kof(a) < kof(b); });
  std::vector<Key3> nk(nbricks);
  for (int j = 0; j < nbricks; ++j) {
    nk[j] = kof(order[j]);
    if (j > 0 && nk[j] == nk[j - 1]) return mac_internal_set_error(MAC_E_ARG, "duplicate chunk key");
  }
  s->keys.clear();
  s->nb = 0;
  SRC(sim_set_bricks(s, nk));
  // scatter host data into sorted brick order
  const size_t cells = (size_t)nbricks * MAC_BC;
  std::vector<uint8_t> f8(cells);
  std::vector<int8_t> l8(cells, -1);
  std::vector<double> buf(cells);
  for (int j = 0; j < nbricks; ++j) {
    memcpy(&f8[(size_t)j * MAC_BC], &flags[(size_t)order[j] * MAC_BC], MAC_BC);
    if (layer) memcpy(&l8[(size_t)j * MAC_BC], &layer[(size_t)order[j] * MAC_BC], MAC_BC);
  }
  if (nbricks > 0) {
    SCHK(hipMemcpy(s->flags, f8.data(), cells, hipMemcpyHostToDevice));
    SCHK(hipMemcpy(s->layer, l8.data(), cells, hipMemcpyHostToDevice));
    const double* src[3] = {u, v, w};
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < nbricks; ++j) {
        if (src[i]) memcpy(&buf[(size_t)j * MAC_BC], &src[i][(size_t)order[j] * MAC_BC], MAC_BC * sizeof(double));
        else memset(&buf[(size_t)j * MAC_BC], 0, MAC_BC * sizeof(double));
      }
      SCHK(hipMemcpy(s->S[0][i], buf.data(), cells * sizeof(double), hipMemcpyHostToDevice));
      SCHK(hipMemcpy(s->S[1][i], buf.data(), cells * sizeof(double), hipMemcpyHostToDevice));
      SCHK(hipMemcpy(s->S[2][i], buf.data(), cells * sizeof(double), hipMemcpyHostToDevice));
      SCHK(hipMemset(s->A[i], 0, cells * sizeof(double)));
    }
    SCHK(hipMemset(s->P, 0, cells * sizeof(double)));
  }
  std::vector<double> c(std::max(nparticles, 1));
  double* dst[3] = {s->px, s->py, s->pz};
  for (int i = 0; i < 3; ++i) {
    for (int k = 0; k < nparticles; ++k) c[k] = pxyz[3 * k + i];
    if (nparticles > 0) SCHK(hipMemcpy(dst[i], c.data(), sizeof(double) * nparticles, hipMemcpyHostToDevice));
  }
  s->np = nparticles;
  s->grid_ready = false;
  return 0;
}

int mac_sim_get_depth(mac_sim* s, int16_t* depth) {
  if (!s || (!depth && s->nb > 0)) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  if (s->nb > 0) SCHK(hipMemcpyAsync(depth, s->depth, sizeof(int16_t) * (size_t)s->nb * MAC_BC, hipMemcpyDeviceToHost, s->st));
  SCHK(hipStreamSynchronize(s->st));
  return 0;
}
int mac_sim_set_depth(mac_sim* s, const int16_t* depth) {
  if (!s || (!depth && s->nb > 0)) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  if (s->nb > 0) SCHK(hipMemcpyAsync(s->depth, depth, sizeof(int16_t) * (size_t)s->nb * MAC_BC, hipMemcpyHostToDevice, s->st));
  SCHK(hipStreamSynchronize(s->st));
  return 0;
}
int mac_sim_counts(mac_sim* s, int* nbricks, int* nparticles) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  if (nbricks) *nbricks = s->nb;
  if (nparticles) *nparticles = s->np;
  return 0;
}

int mac_sim_get_step(mac_sim* s, unsigned long long* step) {
  if (!s || !step) return mac_internal_set_error(MAC_E_ARG, "null argument");
  *step = s->tick;
  return 0;
}

int mac_sim_set_step(mac_sim* s, unsigned long long step) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  s->tick = step;
  return 0;
}

int mac_sim_download(mac_sim* s, int32_t* keys3, uint8_t* flags, int8_t* layer, double* u, double* v, double* w, double* ax,
                     double* ay, double* az, double* p, double* pxyz) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  SCHK(hipStreamSynchronize(s->st));
  const size_t cells = (size_t)s->nb * MAC_BC;
  if (keys3)
    for (int b = 0; b < s->nb; ++b) {
      keys3[3 * b] = s->keys[b].x;
      keys3[3 * b + 1] = s->keys[b].y;
      keys3[3 * b + 2] = s->keys[b].z;
    }
  if (cells > 0) {
    if (flags) SCHK(hipMemcpy(flags, s->flags, cells, hipMemcpyDeviceToHost));
    if (layer) SCHK(hipMemcpy(layer, s->layer, cells, hipMemcpyDeviceToHost));
    double* dst[7] = {u, v, w, ax, ay, az, p};
    double* src[7] = {s->S[0][0], s->S[0][1], s->S[0][2], s->A[0], s->A[1], s->A[2], s->P};
    for (int i = 0; i < 7; ++i)
      if (dst[i]) SCHK(hipMemcpy(dst[i], src[i], cells * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (pxyz && s->np > 0) {
    std::vector<double> c(s->np);
    double* srcp[3] = {s->px, s->py, s->pz};
    for (int i = 0; i < 3; ++i) {
      SCHK(hipMemcpy(c.data(), srcp[i], sizeof(double) * s->np, hipMemcpyDeviceToHost));
      for (int k = 0; k < s->np; ++k) pxyz[3 * k + i] = c[k];
    }
  }
  return 0;
}

int mac_sim_topology(mac_sim* s) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  return sim_topology(s);
}

int mac_sim_compute_dt(mac_sim* s, double* dt) {
  if (!s || !dt) return mac_internal_set_error(MAC_E_ARG, "null argument");
  const mac_sim_params& P = s->prm;
  double mx = 0.0;
  if (s->nb > 0) {
    SimIH ih{{1.0 / P.h[0], 1.0 / P.h[1], 1.0 / P.h[2]}};
    k_dt_max<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->S[0][0], s->S[0][1], s->S[0][2], ih, s->d_part);
    SCHK(hipGetLastError());
    std::vector<double> part(s->nb);
    SCHK(hipMemcpyAsync(part.data(), s->d_part, sizeof(double) * s->nb, hipMemcpyDeviceToHost, s->st));
    SCHK(hipStreamSynchronize(s->st));
    for (double x : part) mx = std::max(mx, x);
  }
  // m_dt = 1/sqrt(max) * m_factor with FP exceptions held (division by 0 -> inf),
  // then Policy_CheckDT
  double d = (mx > 0.0) ? 1.0 / std::sqrt(mx) * P.cfl : HUGE_VAL;
  if (d < P.dt_min) d = P.dt_min;
  if (d > P.dt_max) d = P.dt_max;
  *dt = d;
  return 0;
}

int mac_sim_forces(mac_sim* s, int slot) {
  if (!s || slot < 0 || slot > 1) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  SRC(sim_forces(s, slot));
  SCHK(hipStreamSynchronize(s->st));
  return 0;
}

int mac_sim_advance(mac_sim* s, double dt, mac_sim_stats* stats) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  if (s->nb == 0) return 0;
  const bool tvd = s->prm.integrator == 1;
  const size_t cells = (size_t)s->nb * MAC_BC;
  auto t0 = std::chrono::steady_clock::now();
  SRC(sim_ensure_projection(s));
  mac_stats st;
  // stage 1: forces at u0, u1 = u0 + a dt (/2), acceleration reset, project u1
  SRC(sim_forces(s, 0));
  k_rk_combine<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->S[0][0], s->S[0][1], s->S[0][2], s->S[1][0], s->S[1][1], s->S[1][2],
                                         s->A[0], s->A[1], s->A[2], dt, tvd ? 1 : 0);
  k_zero_accel<<<1024, 256, 0, s->st>>>(cells, s->A[0], s->A[1], s->A[2]);
  SCHK(hipGetLastError());
  int rc = sim_project(s, 1, &st);
  if (rc) return rc;
  if (stats) {
    stats->pcg_iters[0] = st.iters;
    stats->true_rel_resid = st.true_rel_resid;
  }
  // stage 2: forces at u1, combine into u0, project u0
  SRC(sim_forces(s, 1));
  k_rk_combine<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->S[0][0], s->S[0][1], s->S[0][2], s->S[1][0], s->S[1][1], s->S[1][2],
                                         s->A[0], s->A[1], s->A[2], dt, tvd ? 3 : 2);
  SCHK(hipGetLastError());
  rc = sim_project(s, 0, &st);
  if (rc) return rc;
  if (stats) {
    stats->pcg_iters[1] = st.iters;
    stats->max_div = st.max_div;
    if (st.true_rel_resid > stats->true_rel_resid) stats->true_rel_resid = st.true_rel_resid;
    stats->nfluid = st.n_unknowns;
    stats->ms_project = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return 0;
}

int mac_sim_move_particles(mac_sim* s, double dt) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  if (s->np == 0) return 0;
  const mac_sim_params& P = s->prm;
  bool box = !(P.box_lo[0] == P.box_hi[0] && P.box_lo[1] == P.box_hi[1] && P.box_lo[2] == P.box_hi[2]);
  SCHK(hipMemsetAsync(s->d_int + 3, 0, sizeof(int), s->st));
  k_move_particles<<<(s->np + 255) / 256, 256, 0, s->st>>>(s->np, s->px, s->py, s->pz, s->qx, s->qy, s->qz, s->d_int + 3,
                                                            sim_grid(s, 0), dt, P.particle_rk == 4 ? 4 : 2, box ? 1 : 0, P.box_lo[0],
                                                            P.box_lo[1], P.box_lo[2], P.box_hi[0], P.box_hi[1], P.box_hi[2]);
  SCHK(hipGetLastError());
  std::swap(s->px, s->qx);
  std::swap(s->py, s->qy);
  std::swap(s->pz, s->qz);
  return sim_compact_dropped(s);
}

int mac_sim_interpolate(mac_sim* s, int n, const double* pos, double* vel) {
  if (!s || n < 0 || (n > 0 && (!pos || !vel))) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  if (n == 0) return 0;
  double *dp, *dv;
  SCHK(hipMalloc(&dp, sizeof(double) * 3 * n));
  SCHK(hipMalloc(&dv, sizeof(double) * 3 * n));
  SCHK(hipMemcpy(dp, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  k_interpolate<<<(n + 255) / 256, 256, 0, s->st>>>(n, dp, dv, sim_grid(s, 0));
  SCHK(hipGetLastError());
  SCHK(hipStreamSynchronize(s->st));
  SCHK(hipMemcpy(vel, dv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  hipFree(dp);
  hipFree(dv);
  return 0;
}

namespace {
// Algorithm_Extrapolate_Boundary_Condition::Do on the face field (u, v, w) of
// the current bricks (surface_bc.h): seven phases in a fixed order.
int sim_surface_bc(mac_sim* s, double* u, double* v, double* w, int dim) {
  if (s->nb == 0) return 0;
  const size_t nb = (size_t)s->nb;
  if (s->bc_cap < nb) {
    hipFree(s->bc_tmp);
    hipFree(s->bc_pend);
    hipFree(s->bc_brick);
    hipFree(s->bc_cls);
    s->bc_tmp = nullptr;
    s->bc_pend = nullptr;
    s->bc_brick = nullptr;
    s->bc_cls = nullptr;
    s->bc_cap = 0;
    const size_t cap = nb + nb / 4 + 16;
    SCHK(hipMalloc(&s->bc_tmp, 3 * cap * MAC_BC * sizeof(double)));
    SCHK(hipMalloc(&s->bc_pend, 3 * cap * MAC_BC));
    SCHK(hipMalloc(&s->bc_cls, cap * MAC_BC));
    SCHK(hipMalloc(&s->bc_brick, 2 * cap));
    SCHK(hipMemsetAsync(s->bc_pend, 0, 3 * cap * MAC_BC, s->st));
    s->bc_cap = cap;
  }
  const size_t cells = s->bc_cap * MAC_BC;
  BcArgs a;
  a.nb = s->nb;
  a.nbr = s->nbr;
  a.flags = s->flags;
  a.U[0] = u;
  a.U[1] = v;
  a.U[2] = w;
  for (int i = 0; i < 3; ++i) {
    a.tmp[i] = s->bc_tmp + i * cells;
    a.pend[i] = s->bc_pend + i * cells;
    a.h[i] = s->prm.h[i];
    a.ih[i] = 1.0 / s->prm.h[i];
  }
  uint8_t* summ = s->bc_brick;
  uint8_t* active = s->bc_brick + s->bc_cap;
  a.active = active;
  a.cls = s->bc_cls;
  a.dim = dim;
  const int g = s->nb;
  k_bc_summary<<<g, 256, 0, s->st>>>(s->nb, s->flags, summ);
  k_bc_active<<<(g + 255) / 256, 256, 0, s->st>>>(s->nb, s->nbr, summ, active);
  k_bc_classify<<<g, 256, 0, s->st>>>(a);
  k_bc_set_nan<<<g, 256, 0, s->st>>>(a);
  k_bc_div<<<g, 256, 0, s->st>>>(a);
  k_bc_plane<<<g, 256, 0, s->st>>>(a);
  k_bc_apply<<<g, 256, 0, s->st>>>(a);
  k_bc_corner45<<<g, 256, 0, s->st>>>(a);
  if (dim == 3) {
    k_bc_corner3<<<g, 256, 0, s->st>>>(a);
    k_bc_apply<<<g, 256, 0, s->st>>>(a);
  }
  k_bc_mark<<<g, 256, 0, s->st>>>(a);
  k_bc_nan_extrap<<<g, 256, 0, s->st>>>(a);
  k_bc_apply<<<g, 256, 0, s->st>>>(a);
  SCHK(hipGetLastError());
  return 0;
}

// Algorithms_Extrapolate::Do on the face field (u, v, w) of the current
// bricks, after the free-surface condition when surface_bc is set.
int sim_extrapolate_fields(mac_sim* s, double* u, double* v, double* w) {
  if (s->nb == 0) return 0;
  if (s->prm.surface_bc) SRC(sim_surface_bc(s, u, v, w, 3));
  const size_t cells = (size_t)s->nb * MAC_BC;
  if (s->front_cap < 2 * cells) {
    hipFree(s->d_front);
    s->d_front = nullptr;
    s->front_cap = 0;
    SCHK(hipMalloc(&s->d_front, 2 * cells + cells / 4));
    s->front_cap = 2 * cells + cells / 4;
  }
  uint8_t* front = s->d_front;
  uint8_t* next = s->d_front + cells;
  int* dany = s->d_any + 32;
  k_extrap_seed<<<(int)std::min<size_t>((cells + 255) / 256, 4096), 256, 0, s->st>>>(s->nb, s->flags, s->layer, front);
  // one frontier step per layer; layers are int8, so at most 127 steps
  for (int lay = 1; lay < 128; ++lay) {
    int any = 0;
    SCHK(hipMemsetAsync(next, 0, cells, s->st));
    SCHK(hipMemsetAsync(dany, 0, sizeof(int), s->st));
    k_extrapolate<<<s->nb, 256, 0, s->st>>>(s->nb, s->nbr, s->flags, s->layer, u, v, w, front, next, dany, lay);
    SCHK(hipGetLastError());
    SCHK(hipMemcpyAsync(&any, dany, sizeof(int), hipMemcpyDeviceToHost, s->st));
    SCHK(hipStreamSynchronize(s->st));
    if (!any) break;
    std::swap(front, next);
  }
  return 0;
}

// Algorithms_RungeKutta_all::Do (:35-156) for one step of size dt; the cell
// types are those of the step's own cell-type update.  Each stage runs the
// reference's Policy::Do of test/Jet3.cpp:585-650 -- the cell-type chain of
// m_alg_init2 from the stage's particle positions (slot 0 in stage 1, slot 2
// after) with the extrapolation of the stage's speed slot, gravity,
// convection and viscosity, the projection of the acceleration
// (Algorithms_Solve_Pressure_Force; the viscous air pressures read the stage's
// real speed) and the extrapolation of the acceleration
// (Algorithms_Accel_To_Speed) -- then the grid and particle combinations.
int sim_rk4all(mac_sim* s, double dt, mac_sim_stats* stats) {
  mac_stats st;
  int iters[2] = {0, 0};
  double maxdiv = 0.0, worst_resid = 0.0;
  long long nfluid = 0;
  // Acceleration_GetRef().SetZero(true) on the cells that exist (before the
  // first stage and after each of the first three)
  auto zero_accel = [&]() {
    if (s->nb > 0) k_zero_accel_existing<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->A[0], s->A[1], s->A[2]);
  };
  zero_accel();
  double ms_topo = 0.0, ms_proj = 0.0;  // host wall time of the stages' cell-type updates and projections
  auto ms_since = [](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
  };
  for (int stage = 1; stage <= 4; ++stage) {
    const int ks = stage == 1 ? 0 : 2;
    auto ta = std::chrono::steady_clock::now();
    SRC(sim_topology(s, false, ks));
    ms_topo += ms_since(ta);
    if (s->nb == 0) return 0;
    SRC(sim_extrapolate_fields(s, s->S[ks][0], s->S[ks][1], s->S[ks][2]));
    SRC(sim_forces(s, ks, true));
    SCHK(hipStreamSynchronize(s->st));
    auto tp = std::chrono::steady_clock::now();
    SRC(sim_ensure_projection(s));
    mac_internal_set_real_velocity(s->proj, s->S[ks][0], s->S[ks][1], s->S[ks][2]);
    int rc = mac_internal_project(s->proj, s->A[0], s->A[1], s->A[2], s->P, 0, &st);
    mac_internal_set_real_velocity(s->proj, nullptr, nullptr, nullptr);
    if (rc) return rc;
    ms_proj += ms_since(tp);
    iters[stage > 2] += st.iters;
    maxdiv = st.max_div;
    if (st.true_rel_resid > worst_resid) worst_resid = st.true_rel_resid;
    nfluid = st.n_unknowns;
    SRC(sim_extrapolate_fields(s, s->A[0], s->A[1], s->A[2]));
    k_rk4_combine<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, stage, dt, s->S[0][0], s->S[0][1], s->S[0][2], s->S[1][0], s->S[1][1],
                                            s->S[1][2], s->S[2][0], s->S[2][1], s->S[2][2], s->A[0], s->A[1], s->A[2]);
    if (s->np > 0)
      k_rk4_particles<<<(s->np + 255) / 256, 256, 0, s->st>>>(s->np, stage, dt, s->px, s->py, s->pz, s->P1[0], s->P1[1], s->P1[2],
                                                             s->P2[0], s->P2[1], s->P2[2], sim_grid(s, ks));
    if (stage < 4) zero_accel();
    SCHK(hipGetLastError());
  }
  if (stats) {
    stats->pcg_iters[0] = iters[0];
    stats->pcg_iters[1] = iters[1];
    stats->max_div = maxdiv;
    stats->true_rel_resid = worst_resid;
    stats->nfluid = nfluid;
    stats->ms_topology += ms_topo;
    stats->ms_project += ms_proj;
  }
  return 0;
}

// The inflow pulsation of step n = tick (mac_sim_params.inflow_amp).
int sim_pulse_inflow(mac_sim* s) {
  const mac_sim_params& P = s->prm;
  if (P.inflow_amp == 0.0 || s->nb == 0) return 0;
  const double phi = 2.0 * M_PI * sim_uniform(P.seed, 0x1f10, 0, 0);
  const double f = 1.0 + P.inflow_amp * std::sin(2.0 * M_PI * (double)s->tick / (double)P.inflow_period + phi);
  k_inflow_speed<<<s->nb, 256, 0, s->st>>>(s->nb, s->flags, s->S[0][0], s->S[0][1], s->S[0][2], P.inflow_speed[0] * f,
                                            P.inflow_speed[1] * f, P.inflow_speed[2] * f);
  SCHK(hipGetLastError());
  return 0;
}

// Particles leaving [box_lo, box_hi) are deleted (slot 0).
int sim_box_delete(mac_sim* s) {
  const mac_sim_params& P = s->prm;
  if (s->np == 0 || (P.box_lo[0] == P.box_hi[0] && P.box_lo[1] == P.box_hi[1] && P.box_lo[2] == P.box_hi[2])) return 0;
  SCHK(hipMemsetAsync(s->d_int + 3, 0, sizeof(int), s->st));
  k_box_delete<<<(s->np + 255) / 256, 256, 0, s->st>>>(s->np, s->px, s->py, s->pz, s->d_int + 3, P.box_lo[0], P.box_lo[1],
                                                        P.box_lo[2], P.box_hi[0], P.box_hi[1], P.box_hi[2]);
  SCHK(hipGetLastError());
  return sim_compact_dropped(s);
}
}  // namespace

int mac_sim_extrapolate(mac_sim* s, int slot) {
  if (!s || slot < 0 || slot > 2) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  return sim_extrapolate_fields(s, s->S[slot][0], s->S[slot][1], s->S[slot][2]);
}

int mac_sim_surface_bc(mac_sim* s, int slot, int dim) {
  if (!s || slot < 0 || slot > 3 || (dim != 2 && dim != 3)) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  double** f = slot == 3 ? s->A : s->S[slot];
  SRC(sim_surface_bc(s, f[0], f[1], f[2], dim));
  SCHK(hipStreamSynchronize(s->st));
  return 0;
}

int mac_sim_set_guess(mac_sim* s, int mode, mac_unet* net, float out_scale) {
  if (!s || mode < 0 || mode > 2) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  if (mode == 2 && (!net || !(out_scale > 0.f))) return mac_internal_set_error(MAC_E_ARG, "learned guess needs a network and out_scale > 0");
  s->guess_mode = mode;
  s->net = mode == 2 ? net : nullptr;
  s->out_scale = out_scale;
  return 0;
}

int mac_sim_capture(mac_sim* s, mac_unet* net, float out_scale, float* in5, float* target, float* mask) {
  if (!s || !net) return mac_internal_set_error(MAC_E_ARG, "null argument");
  if (!s->grid_ready) return mac_internal_set_error(MAC_E_STATE, "no projection since the grid last changed");
  return mac_internal_unet_sample(net, s->proj, s->Pprev, out_scale, in5, target, mask);
}

int mac_sim_step(mac_sim* s, mac_sim_stats* stats) {
  if (!s) return mac_internal_set_error(MAC_E_ARG, "null sim");
  mac_sim_stats local;
  memset(&local, 0, sizeof(local));
  auto t0 = std::chrono::steady_clock::now();
  SRC(sim_topology(s));
  auto t1 = std::chrono::steady_clock::now();
  double dt;
  std::chrono::steady_clock::time_point t2, t3;
  if (s->prm.integrator == 2) {
    // the init chain of test/Jet3.cpp:545-547 ends with the time step and the
    // extrapolation of speed slot 0; the particles move inside the stages
    SRC(mac_sim_compute_dt(s, &dt));
    local.dt = dt;
    SRC(sim_pulse_inflow(s));
    SRC(sim_extrapolate_fields(s, s->S[0][0], s->S[0][1], s->S[0][2]));
    SRC(sim_rk4all(s, dt, &local));
    t2 = std::chrono::steady_clock::now();
    SRC(sim_box_delete(s));
    t3 = std::chrono::steady_clock::now();
  } else {
    SRC(mac_sim_compute_dt(s, &dt));
    local.dt = dt;
    SRC(sim_pulse_inflow(s));
    SRC(mac_sim_advance(s, dt, &local));
    t2 = std::chrono::steady_clock::now();
    SRC(mac_sim_move_particles(s, dt));
    t3 = std::chrono::steady_clock::now();
  }
  local.nbricks = s->nb;
  local.nparticles = s->np;
  // the RK4 stages' own cell-type updates and projections are already in
  // ms_topology / ms_project; the rest of the stage work counts as forces
  const double stage_topo = local.ms_topology;
  local.ms_topology += std::chrono::duration<double, std::milli>(t1 - t0).count();
  local.ms_forces = std::chrono::duration<double, std::milli>(t2 - t1).count() - local.ms_project - stage_topo;
  local.ms_particles = std::chrono::duration<double, std::milli>(t3 - t2).count();
  local.ms_total = std::chrono::duration<double, std::milli>(t3 - t0).count();
  if (stats) *stats = local;
  return 0;
}

}  // extern "C"
