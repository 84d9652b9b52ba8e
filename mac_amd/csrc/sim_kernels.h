// sim_kernels.h -- HIP kernels of the per-timestep MAC-grid loop on the
// sparse brick grid (topology, time step, forces, RK2 combination,
// staggered interpolation, particle advection).
//
// Cell addressing: brick b, cell t = x + 8y + 64z (x fastest); global key of
// the cell = 8 * key(b) + (x, y, z).  Missing bricks read as zero; a cell
// whose existence bit is off keeps the data it had (as a chunk keeps a
// deleted cell's data until the chunk itself is freed,
// ChunkHashTable.tpp:133-161).
#pragma once
#include <stdint.h>

#include "brick.h"

#define SIM_HEMPTY 0xFFFFFFFFFFFFFFFFull

__host__ __device__ inline uint64_t sim_pack(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
         (uint64_t)(uint32_t)(z + (1 << 20));
}

__host__ __device__ inline uint64_t sim_mix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Counter-based uniform in (0,1) of (seed, stream, a, b) -- SplitMix64 finaliser.
__host__ __device__ inline double sim_uniform(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  uint64_t z = seed * 0xD1B54A32D192ED03ull + stream * 0x8CB92BA72F3D8DD7ull + a * 0x9E3779B97F4A7C15ull +
               (b + 1) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Chunk key of a cell key (ChunkHashTable.tpp:181-196).
__host__ __device__ inline int sim_chunk(int k) { return k >= 0 ? k / 8 : (k + 1) / 8 - 1; }

// Materials (MaterialToInt numbering) and predicates of
// Data_CellType_Fluid_Air_Solid_Inflow_Interior_0_Pres.tpp:25-233.
__host__ __device__ inline bool mat_is_fluid(unsigned m) { return m == 1u || m == 3u || m == 4u || m == 5u; }
__host__ __device__ inline unsigned mat_set_fluid(unsigned m) { return (m == 0u || m == 4u) ? 1u : m; }
__host__ __device__ inline unsigned mat_set_air(unsigned m) { return (m == 1u || m == 4u) ? 0u : m; }

#define F_EXISTS 0x80u
#define F_NODEL 0x40u

// Slot of an appended element: one atomic per wave on *counter, lanes ranked
// by lane id among the wave's keepers.  Every active lane of the wave must
// call it; returns -1 where keep is false.
__device__ inline int wave_append(bool keep, int* counter) {
  const unsigned long long m = __ballot(keep);
  if (m == 0ull) return -1;
  const int lane = (int)__lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, (int)__popcll(m));
  base = __shfl(base, leader);
  return keep ? base + (int)__popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Adds the number of the wave's lanes with `flag` to *counter: one atomic in
// each wave that has any, none elsewhere.
__device__ inline void wave_count(bool flag, int* counter) {
  const unsigned long long m = __ballot(flag);
  if (m != 0ull && (int)__lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(counter, (int)__popcll(m));
}

// In-place compaction of particle positions whose dropped entries hold a NaN
// x (a kept particle lies inside the box, so its x is never NaN).  With d
// dropped among np, the kept ones end in [0, np - d): k_list_holes lists the
// dropped indices below np - d, k_fill_holes moves the kept ones at or above
// it into them; the two counts are equal.  Particle order is not kept.
__global__ void k_list_holes(int n_new, const double* __restrict__ x, int* __restrict__ holes, int* __restrict__ nholes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_new) return;
  const int o = wave_append(__builtin_isnan(x[i]), nholes);
  if (o >= 0) holes[o] = i;
}

__global__ void k_fill_holes(int n_new, int np, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                             const int* __restrict__ holes, int* __restrict__ nfilled) {
  const int i = n_new + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double xv = x[i];
  const int o = wave_append(!__builtin_isnan(xv), nfilled);
  if (o < 0) return;
  const int j = holes[o];
  x[j] = xv;
  y[j] = y[i];
  z[j] = z[i];
}

__device__ inline int dh_lookup(const uint64_t* __restrict__ hk, const int* __restrict__ hv, int cap, uint64_t k) {
  const uint32_t m = (uint32_t)cap - 1u;
  uint32_t i = (uint32_t)sim_mix(k) & m;
  for (int n = 0; n < cap; ++n) {
    uint64_t kk = hk[i];
    if (kk == k) return hv[i];
    if (kk == SIM_HEMPTY) return -1;
    i = (i + 1u) & m;
  }
  return -1;
}

__global__ void k_hash_clear(int cap, uint64_t* __restrict__ hk, int* __restrict__ hv) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
    hk[i] = SIM_HEMPTY;
    hv[i] = -1;
  }
}

// Insert every brick key (open addressing, linear probing; keys are unique).
__global__ void k_hash_insert(int nb, const int* __restrict__ keys3, uint64_t* __restrict__ hk, int* __restrict__ hv, int cap) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const uint64_t k = sim_pack(keys3[3 * b], keys3[3 * b + 1], keys3[3 * b + 2]);
  const uint32_t m = (uint32_t)cap - 1u;
  uint32_t i = (uint32_t)sim_mix(k) & m;
  for (int n = 0; n < cap; ++n) {
    unsigned long long prev = atomicCAS((unsigned long long*)&hk[i], (unsigned long long)SIM_HEMPTY, (unsigned long long)k);
    if (prev == SIM_HEMPTY || prev == k) {
      hv[i] = b;
      return;
    }
    i = (i + 1u) & m;
  }
}

// 27-neighbour brick table from the hash (Hook_Neighbour_List_key.h:97-120).
__global__ void k_build_nbr(int nb, const int* __restrict__ keys3, const uint64_t* __restrict__ hk, const int* __restrict__ hv,
                            int cap, int* __restrict__ nbr27) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * 27) return;
  int b = idx / 27, s = idx % 27;
  int ox = s % 3 - 1, oy = (s / 3) % 3 - 1, oz = s / 9 - 1;
  nbr27[idx] = dh_lookup(hk, hv, cap, sim_pack(keys3[3 * b] + ox, keys3[3 * b + 1] + oy, keys3[3 * b + 2] + oz));
}

// Cell of a global cell key through the hash: returns brick*512+t or -1.
__device__ inline long long cell_of_key(const uint64_t* hk, const int* hv, int cap, int kx, int ky, int kz) {
  int b = dh_lookup(hk, hv, cap, sim_pack(sim_chunk(kx), sim_chunk(ky), sim_chunk(kz)));
  if (b < 0) return -1;
  return (long long)b * MAC_BC + (kx & 7) + 8 * (ky & 7) + 64 * (kz & 7);
}

// Particle -> cell key: key_i = lround(x_i * (1/h_i)) (Policy_Particle_To_Key.h:24-35).
__device__ inline void part_key(double x, double y, double z, const double* ih, int& kx, int& ky, int& kz) {
#pragma clang fp contract(off)
  kx = (int)lround(x * ih[0]);
  ky = (int)lround(y * ih[1]);
  kz = (int)lround(z * ih[2]);
}

struct SimIH {
  double ih[3];
};

// ---------------------------------------------------------------- topology --

// Algorithms_Initialize_MacCell::Do: every layer := empty.
__global__ void k_reset_layer(size_t ncell, int8_t* __restrict__ layer) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ncell; i += (size_t)gridDim.x * blockDim.x) layer[i] = -1;
}

// Bricks the particles land in: missing ones are appended to `missing`
// (duplicates allowed), existing ones flagged.
__global__ void k_particle_bricks(int np, const double* __restrict__ px, const double* __restrict__ py,
                                  const double* __restrict__ pz, SimIH ih, const uint64_t* __restrict__ hk,
                                  const int* __restrict__ hv, int cap, int* __restrict__ bflag, int* __restrict__ missing,
                                  int* __restrict__ nmissing, int maxmissing, int floor_axis, double floor_value) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  if (floor_axis >= 0 && (floor_axis == 0 ? px[i] : floor_axis == 1 ? py[i] : pz[i]) < floor_value) return;
  int kx, ky, kz;
  part_key(px[i], py[i], pz[i], ih.ih, kx, ky, kz);
  int cx = sim_chunk(kx), cy = sim_chunk(ky), cz = sim_chunk(kz);
  int b = dh_lookup(hk, hv, cap, sim_pack(cx, cy, cz));
  if (b >= 0) {
    bflag[b] = 1;
  } else {
    int slot = atomicAdd(nmissing, 1);
    if (slot < maxmissing) {
      missing[3 * slot] = cx;
      missing[3 * slot + 1] = cy;
      missing[3 * slot + 2] = cz;
    }
  }
}

__global__ void k_set_clear(int cap, uint64_t* __restrict__ set) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) set[i] = SIM_HEMPTY;
}

// The absent bricks of the flagged bricks' 26-neighbourhoods, each listed
// once: a candidate key is inserted into the open-addressing set `set` (cap a
// power of two, above twice maxn) and only its first inserter appends it.
// *n counts the distinct keys; out holds at most maxn of them.
__global__ void k_missing_neighbours(int nb, const int* __restrict__ bflag, const int* __restrict__ nbr27,
                                     const int* __restrict__ keys3, uint64_t* __restrict__ set, int cap,
                                     int* __restrict__ out, int* __restrict__ n, int maxn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool fresh = false;
  int kx = 0, ky = 0, kz = 0;
  if (i < nb * 27) {
    const int b = i / 27, k = i % 27;
    if (k != 13 && bflag[b] && nbr27[i] < 0) {
      kx = keys3[3 * b] + k % 3 - 1;
      ky = keys3[3 * b + 1] + (k / 3) % 3 - 1;
      kz = keys3[3 * b + 2] + k / 9 - 1;
      const uint64_t key = sim_pack(kx, ky, kz);
      const uint32_t m = (uint32_t)cap - 1u;
      uint32_t h = (uint32_t)sim_mix(key) & m;
      for (int t = 0; t < cap; ++t) {
        const unsigned long long prev =
            atomicCAS((unsigned long long*)&set[h], (unsigned long long)SIM_HEMPTY, (unsigned long long)key);
        if (prev == SIM_HEMPTY) {
          fresh = true;
          break;
        }
        if (prev == key) break;
        h = (h + 1u) & m;
      }
    }
  }
  const int slot = wave_append(fresh, n);
  if (slot < 0 || slot >= maxn) return;
  out[3 * slot] = kx;
  out[3 * slot + 1] = ky;
  out[3 * slot + 2] = kz;
}

// The pulsed inflow: the const faces of every existing Inflow cell get v[i].
__global__ void k_inflow_speed(int nb, const uint8_t* __restrict__ flags, double* __restrict__ u, double* __restrict__ v,
                               double* __restrict__ w, double v0, double v1, double v2) {
  const int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t c = (size_t)b * MAC_BC + t;
    const unsigned f = flags[c];
    if (!(f & F_EXISTS) || (f & 7u) != 3u) continue;
    if (f & 0x08u) u[c] = v0;
    if (f & 0x10u) v[c] = v1;
    if (f & 0x20u) w[c] = v2;
  }
}

// Bricks holding a cell that the fill step can give layer 0: Inflow cells,
// and Interior cells when `interior` is set.
__global__ void k_inflow_bricks(int nb, const uint8_t* __restrict__ flags, int* __restrict__ bflag, int interior) {
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    unsigned f = flags[(size_t)b * MAC_BC + t];
    if ((f & F_EXISTS) && ((f & 7u) == 3u || (interior && (f & 7u) == 4u))) bflag[b] = 1;
  }
}

// Algorithms_Layer_Initial_With_Particle::Do (:20-35 of the non-depth
// variant): the cell of each particle gets layer 0 and SetFluid.  Every
// writer of a cell writes the same values, so concurrent updates agree.
__global__ void k_mark_particles(int np, const double* __restrict__ px, const double* __restrict__ py,
                                 const double* __restrict__ pz, SimIH ih, const uint64_t* __restrict__ hk,
                                 const int* __restrict__ hv, int cap, uint8_t* __restrict__ flags, int8_t* __restrict__ layer) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int kx, ky, kz;
  part_key(px[i], py[i], pz[i], ih.ih, kx, ky, kz);
  long long c = cell_of_key(hk, hv, cap, kx, ky, kz);
  if (c < 0) return;
  unsigned f = flags[c];
  unsigned m = mat_set_fluid(f & 7u);
  flags[c] = (uint8_t)((f & ~7u) | m | F_EXISTS);
  layer[c] = 0;
}

// ------------------------------------------------------- depth chain ----
// Algorithms_Initialize_Layer_And_Depth::Do (:16-33): layer := empty for every
// cell, depth := empty for the fluid cells that are not Interior.
__global__ void k_reset_layer_depth(size_t ncell, const uint8_t* __restrict__ flags, int8_t* __restrict__ layer,
                                    int16_t* __restrict__ depth) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ncell; i += (size_t)gridDim.x * blockDim.x) {
    layer[i] = -1;
    const unsigned f = flags[i];
    if ((f & F_EXISTS) && mat_is_fluid(f & 7u) && (f & 7u) != 4u) depth[i] = -1;
  }
}

// Algorithm_DelParticle::Do (test/Jet3.cpp:259-289) then
// Algorithms_Layer_Initial_With_Particle_Depth::Do (:20-51).  A particle with
// position[floor_axis] < floor_value is erased first (floor_axis < 0: none).
// A particle in an Interior cell gives the cell layer 0 and is erased when the
// cell's depth exceeds depth_remove; any other particle's cell gets layer 0
// and SetFluid.  depth == nullptr: no cell is treated as Interior (the
// non-depth Algorithms_Layer_Initial_With_Particle with the floor).  keep[i]
// records whether particle i stays; *ndropped counts the others, which
// k_list_holes_marked / k_fill_holes_marked then compact away.
__global__ void k_mark_particles_keep(int np, const double* __restrict__ px, const double* __restrict__ py,
                                      const double* __restrict__ pz, SimIH ih, const uint64_t* __restrict__ hk,
                                      const int* __restrict__ hv, int cap, uint8_t* __restrict__ flags, int8_t* __restrict__ layer,
                                      const int16_t* __restrict__ depth, int depth_remove, int floor_axis, double floor_value,
                                      int* __restrict__ keepmark, int* __restrict__ ndropped) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double x = px[i], y = py[i], z = pz[i];
  bool keep = true;
  if (floor_axis >= 0) keep = !((floor_axis == 0 ? x : floor_axis == 1 ? y : z) < floor_value);
  if (keep) {
    int kx, ky, kz;
    part_key(x, y, z, ih.ih, kx, ky, kz);
    long long c = cell_of_key(hk, hv, cap, kx, ky, kz);
    if (c >= 0) {
      const unsigned f = flags[c];
      if (depth && (f & 7u) == 4u) {
        // grid[key] sets the cell's bit and keeps its data (ChunkHashTable.tpp:63-87)
        flags[c] = (uint8_t)(f | F_EXISTS);
        layer[c] = 0;
        keep = depth[c] <= depth_remove;
      } else {
        flags[c] = (uint8_t)((f & ~7u) | mat_set_fluid(f & 7u) | F_EXISTS);
        layer[c] = 0;
      }
    }
  }
  keepmark[i] = keep ? 1 : 0;
  wave_count(!keep, ndropped);
}

// In-place compaction by a keep mark, over up to three position slots moved
// together: with d dropped among np, k_list_holes_marked lists the dropped
// indices below np - d, k_fill_holes_marked moves the kept particles at or
// above it into them (the two counts are equal).
__global__ void k_list_holes_marked(int n_new, const int* __restrict__ keepmark, int* __restrict__ holes,
                                    int* __restrict__ nholes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_new) return;
  const int o = wave_append(keepmark[i] == 0, nholes);
  if (o >= 0) holes[o] = i;
}

struct PartSlots {
  double* p[9];  // x, y, z of each slot
  int n;         // slots in use (1..3)
};

__global__ void k_fill_holes_marked(int n_new, int np, const int* __restrict__ keepmark, const int* __restrict__ holes,
                                    int* __restrict__ nfilled, PartSlots sl) {
  const int i = n_new + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int o = wave_append(keepmark[i] != 0, nfilled);
  if (o < 0) return;
  const int j = holes[o];
  for (int k = 0; k < 3 * sl.n; ++k) sl.p[k][j] = sl.p[k][i];
}

// Start of Algorithms_Update_CellType_Layer_Depth::Do (:31-56): depth := empty;
// a layer-0 cell with a face neighbour that is missing or layer-empty is on
// the surface: depth 0.  Runs before the air layers are laid.
__global__ void k_depth_surface(int nb, const int* __restrict__ nbr27, const uint8_t* __restrict__ flags,
                                const int8_t* __restrict__ layer, int16_t* __restrict__ depth) {
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    if (!(flags[c] & F_EXISTS)) continue;
    int16_t d = -1;
    if (layer[c] == 0) {
      const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
      const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
      for (int k = 0; k < 6; ++k) {
        int cell;
        const int nbk = locate(nbr27, b, x + dx[k], y + dy[k], z + dz[k], cell);
        if (nbk < 0) {
          d = 0;
          break;
        }
        const size_t nc = (size_t)nbk * MAC_BC + cell;
        if (!(flags[nc] & F_EXISTS) || layer[nc] == -1) {
          d = 0;
          break;
        }
      }
    }
    depth[c] = d;
  }
}

// The depth search of Algorithms_Update_CellType_Layer_Depth.h:96-133 as a
// relaxation.  The reference's level-by-level search gives every layer-0 cell
// of empty depth its distance, in face steps through layer-0 cells, from the
// surface cells (depth 0); that distance is the unique fixed point of
// depth(c) = 1 + min over face neighbours of depth >= 0, reached from any
// upper bounds.  One block per brick holds the brick's depths and the face
// planes of its six neighbours in LDS and relaxes in place (values only fall
// and are always lengths of real paths) until nothing changes in the brick;
// the neighbour planes are those of the launch's start.  *any is set when a
// launch changed a depth; launches repeat until one changes nothing.
// Depths are int16: a distance above 32766 sets *overflow.
// prev: the change flag of the launch before in the same batch (nullptr for
// the first); when it is 0 the search has settled and the launch returns.
__global__ void __launch_bounds__(256) k_depth_relax(int nb, const int* __restrict__ nbr27, const uint8_t* __restrict__ flags,
                                                     const int8_t* __restrict__ layer, int16_t* __restrict__ depth,
                                                     const int* __restrict__ prev, int* __restrict__ any, int* __restrict__ overflow) {
  __shared__ int16_t D[1000];  // [z+1][y+1][x+1], x, y, z in -1..8
  if (prev && *prev == 0) return;
  const int b = blockIdx.x;
  for (int i = threadIdx.x; i < 1000; i += blockDim.x) {
    const int hx = i % 10 - 1, hy = (i / 10) % 10 - 1, hz = i / 100 - 1;
    const int out = (hx < 0 || hx > 7) + (hy < 0 || hy > 7) + (hz < 0 || hz > 7);
    int16_t v = -1;
    if (out <= 1) {  // the brick and its face planes; edges and corners are never read
      int cell;
      const int nbk = locate(nbr27, b, hx, hy, hz, cell);
      if (nbk >= 0) v = depth[(size_t)nbk * MAC_BC + cell];
    }
    D[i] = v;
  }
  __syncthreads();
  int li[2];
  bool upd[2];
  int16_t start[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = threadIdx.x + 256 * k;
    const size_t c = (size_t)b * MAC_BC + t;
    li[k] = (t & 7) + 1 + 10 * (((t >> 3) & 7) + 1) + 100 * ((t >> 6) + 1);
    start[k] = D[li[k]];
    // surface cells (0) are fixed; every other existing layer-0 cell relaxes
    upd[k] = (flags[c] & F_EXISTS) && layer[c] == 0 && start[k] != 0;
  }
  bool changed_any = false, big = false;
  for (int it = 0; it < 4096; ++it) {
    bool ch = false;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!upd[k]) continue;
      const int l = li[k];
      int m = 1 << 30;
      const int nbv[6] = {D[l - 1], D[l + 1], D[l - 10], D[l + 10], D[l - 100], D[l + 100]};
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (nbv[q] >= 0) m = min(m, nbv[q] + 1);
      if (m == (1 << 30)) continue;
      const int cur = D[l];
      if (cur < 0 || m < cur) {
        if (m > 32766) {
          big = true;
          continue;
        }
        D[l] = (int16_t)m;
        ch = true;
      }
    }
    if (!__syncthreads_or(ch)) break;
    changed_any = true;
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = threadIdx.x + 256 * k;
    if (upd[k] && D[li[k]] != start[k]) depth[(size_t)b * MAC_BC + t] = D[li[k]];
  }
  if (big) *overflow = 1;
  if (changed_any && threadIdx.x == 0) *any = 1;
}

// After the search: a cell it reached (layer 0, depth >= 1) becomes Fluid
// (SetFluid) when its depth is below depth_interior, Interior otherwise.
__global__ void k_depth_types(int nb, uint8_t* __restrict__ flags, const int8_t* __restrict__ layer,
                              const int16_t* __restrict__ depth, int depth_interior) {
  const int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t c = (size_t)b * MAC_BC + t;
    const unsigned f = flags[c];
    const int d = depth[c];
    if (!(f & F_EXISTS) || layer[c] != 0 || d < 1) continue;
    flags[c] = (uint8_t)((f & ~7u) | (d < depth_interior ? mat_set_fluid(f & 7u) : 4u));
  }
}

// Algorithms_Create_Fluid_Particle::Do with Policy_Is_Inbound_Filling_Layer_CellType
// (:16-19): an existing Inflow cell with an empty layer gets particles, layer
// 0 and SetFluid.  The particles are `ppc` at seeded jitter inside the cell
// ((key + U - 1/2) h per axis), or with cube_n > 0 the (2n+1)^3 lattice of
// Policy_Add_Particle_Cube.
// With depth (Algorithms_Create_Fluid_Particle_Depth::Do, :20-47): an Interior
// cell with an empty layer and depth <= depth_remove also gets particles and
// layer 0 (no SetFluid); any other Interior cell gets layer 0.
// create == 0 is Algorithms_Create_Interior_Depth (the same cell updates,
// no particles), run inside the RK4 stages.
__global__ void k_inflow_fill(int nb, const int* __restrict__ keys3, uint8_t* __restrict__ flags, int8_t* __restrict__ layer,
                              double h0, double h1, double h2, int ppc, uint64_t seed, uint64_t tick, double* __restrict__ px,
                              double* __restrict__ py, double* __restrict__ pz, int* __restrict__ np, int maxp,
                              int* __restrict__ overflow, const int16_t* __restrict__ depth, int depth_remove, int create,
                              int cube_n) {
#pragma clang fp contract(off)
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if (!(f & F_EXISTS)) continue;
    const bool inflow = (ppc > 0 || cube_n > 0) && (f & 7u) == 3u && layer[c] == -1;
    if (!inflow) {
      if (!depth || (f & 7u) != 4u) continue;
      const bool refill = layer[c] == -1 && depth[c] <= depth_remove;
      layer[c] = 0;
      if (!refill) continue;
    }
    int kx = keys3[3 * b] * 8 + (t & 7), ky = keys3[3 * b + 1] * 8 + ((t >> 3) & 7), kz = keys3[3 * b + 2] * 8 + (t >> 6);
    uint64_t ck = sim_pack(kx, ky, kz);
    if (cube_n > 0 && create) {
      // Policy_Add_Particle_Cube::AddParticle (:29-59): (2n+1)^3 particles at
      // (k + j dh) h, j = -n..n, dh = 1/(2n+1), x fastest
      const int m = 2 * cube_n + 1, cnt = m * m * m;
      const double dh = 1. / (2 * cube_n + 1);
      const int base = atomicAdd(np, cnt);
      for (int k = 0; k < cnt; ++k) {
        if (base + k >= maxp) {
          *overflow = 1;
          break;
        }
        const int jx = k % m - cube_n, jy = (k / m) % m - cube_n, jz = k / (m * m) - cube_n;
        px[base + k] = ((double)kx + jx * dh) * h0;
        py[base + k] = ((double)ky + jy * dh) * h1;
        pz[base + k] = ((double)kz + jz * dh) * h2;
      }
    } else if (create) {
      int base = atomicAdd(np, ppc);
      for (int k = 0; k < ppc; ++k) {
        if (base + k >= maxp) {
          *overflow = 1;
          break;
        }
        double ux = sim_uniform(seed, tick, ck, 3 * k), uy = sim_uniform(seed, tick, ck, 3 * k + 1),
               uz = sim_uniform(seed, tick, ck, 3 * k + 2);
        px[base + k] = ((double)kx + (ux - 0.5)) * h0;
        py[base + k] = ((double)ky + (uy - 0.5)) * h1;
        pz[base + k] = ((double)kz + (uz - 0.5)) * h2;
      }
    }
    layer[c] = 0;
    if (inflow) flags[c] = (uint8_t)((f & ~7u) | mat_set_fluid(f & 7u));
  }
}

// One level of Algorithms_Update_CellType_Layer::Do (:22-76): every missing
// or layer-empty face neighbour of a layer-(lay-1) cell gets layer `lay` and
// SetAir.  Level-synchronous: a cell set in this pass has layer lay != lay-1,
// so the outcome does not depend on thread order (the reference's stack
// order does not matter either, each level being a set).
__global__ void k_layer_round(int nb, const int* __restrict__ nbr27, uint8_t* __restrict__ flags, int8_t* __restrict__ layer,
                              int lay) {
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if ((f & F_EXISTS) && layer[c] != -1) continue;
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      int cell;
      int nbk = locate(nbr27, b, x + dx[k], y + dy[k], z + dz[k], cell);
      if (nbk < 0) continue;
      size_t nc = (size_t)nbk * MAC_BC + cell;
      if ((flags[nc] & F_EXISTS) && layer[nc] == lay - 1) hit = true;
    }
    if (hit) {
      layer[c] = (int8_t)lay;
      flags[c] = (uint8_t)((f & ~7u) | mat_set_air(f & 7u) | F_EXISTS);
    }
  }
}

// The closing pass (:77-95): from each cell of the last layer, only the +1
// neighbours that are missing or empty get layer_max + 1 and SetAir.
__global__ void k_layer_extra(int nb, const int* __restrict__ nbr27, uint8_t* __restrict__ flags, int8_t* __restrict__ layer,
                              int lmax) {
  int b = blockIdx.x;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if ((f & F_EXISTS) && layer[c] != -1) continue;
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    bool hit = false;
    for (int k = 0; k < 3; ++k) {
      int cell;
      int nbk = locate(nbr27, b, x - (k == 0), y - (k == 1), z - (k == 2), cell);
      if (nbk < 0) continue;
      size_t nc = (size_t)nbk * MAC_BC + cell;
      if ((flags[nc] & F_EXISTS) && layer[nc] == lmax) hit = true;
    }
    if (hit) {
      layer[c] = (int8_t)(lmax + 1);
      flags[c] = (uint8_t)((f & ~7u) | mat_set_air(f & 7u) | F_EXISTS);
    }
  }
}

// Algorithms_Delete_MacCell::Do (:22-44): erase layer-empty cells unless
// NoDelete; count the cells left per brick.
__global__ void k_delete_cells(int nb, uint8_t* __restrict__ flags, const int8_t* __restrict__ layer, int* __restrict__ alive) {
  int b = blockIdx.x;
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int my = 0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if ((f & F_EXISTS) && layer[c] == -1 && !(f & F_NODEL)) {
      f &= ~F_EXISTS;
      flags[c] = (uint8_t)f;
    }
    my += (f & F_EXISTS) ? 1 : 0;
  }
  atomicAdd(&cnt, my);
  __syncthreads();
  if (threadIdx.x == 0) alive[b] = cnt;
}

// Re-layout after the brick set changed: new brick j takes old brick
// from[j], or the prototype cell (Air, no const, not existing, layer empty,
// zero fields) when from[j] < 0 -- a chunk allocated from the prototype.
__global__ void k_relayout_u8(int nb, const int* __restrict__ from, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                              uint8_t proto) {
  int j = blockIdx.x;
  int o = from[j];
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x)
    dst[(size_t)j * MAC_BC + t] = o >= 0 ? src[(size_t)o * MAC_BC + t] : proto;
}
__global__ void k_relayout_i16(int nb, const int* __restrict__ from, const int16_t* __restrict__ src, int16_t* __restrict__ dst,
                               int16_t proto) {
  int j = blockIdx.x;
  int o = from[j];
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x)
    dst[(size_t)j * MAC_BC + t] = o >= 0 ? src[(size_t)o * MAC_BC + t] : proto;
}
// Every FP64 per-cell array relaid in one launch: blockIdx.y picks the array.
struct RelayoutF64 {
  static constexpr int CAP = 16;
  const double* src[CAP];
  double* dst[CAP];
};
__global__ void k_relayout_f64(int nb, const int* __restrict__ from, RelayoutF64 arr) {
  const int j = blockIdx.x;
  const int o = from[j];
  const double* __restrict__ src = arr.src[blockIdx.y];
  double* __restrict__ dst = arr.dst[blockIdx.y];
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x)
    dst[(size_t)j * MAC_BC + t] = o >= 0 ? src[(size_t)o * MAC_BC + t] : 0.0;
}

// ----------------------------------------------------------------- dt -------

// max over existing fluid cells of sum_i (u_i inv_h_i)^2 (Algorithms_Calculate_Time_Step.h:41-54).
__global__ void k_dt_max(int nb, const uint8_t* __restrict__ flags, const double* __restrict__ u, const double* __restrict__ v,
                         const double* __restrict__ w, SimIH ih, double* __restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double sh[16];
  int b = blockIdx.x;
  double m = 0.0;
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if (!(f & F_EXISTS) || !mat_is_fluid(f & 7u)) continue;
    double a = u[c] * ih.ih[0], bb = v[c] * ih.ih[1], cc = w[c] * ih.ih[2];
    double temp = 0.0;
    temp += a * a;
    temp += bb * bb;
    temp += cc * cc;
    if (temp > m) m = temp;
  }
  double t = block_max(m, sh);
  if (threadIdx.x == 0) partials[b] = t;
}

// --------------------------------------------------------------- forces -----

struct SimForce {
  double ih[3];
  double g[3];
  double nu;
  int accumulate;  // add to the acceleration held (the RK4 "all" stages) instead of to 0
  int conv_first;  // convection before viscosity (test/Jet3.cpp:596-599) instead of after
};

__device__ inline bool cell_fluid(const uint8_t* flags, const int* nbr27, int b, int x, int y, int z, bool& valid) {
  int cell;
  int nb = locate(nbr27, b, x, y, z, cell);
  if (nb < 0) {
    valid = false;
    return false;
  }
  unsigned f = flags[(size_t)nb * MAC_BC + cell];
  valid = f & F_EXISTS;
  return valid && mat_is_fluid(f & 7u);
}

// acceleration := 0 (forced), then in the reference's order
// Algorithms<Gravity, Viscosity, Convection> (TestJetDEau_Basic.cpp:363):
//   gravity    Policy_Gravity.h:14-45 (fluid cell, or non-empty cell whose -e_i neighbour is fluid)
//   viscosity  Algorithms_Viscosity.h:19-35 + Policy_Laplacian_Speed.h:22-37 + Policy_Viscosity_Apply_If.h:16-39
//   convection Algorithms_Convection.h:18-37 + Policy_Upwind_1_Order.h:18-66 + Policy_Convection_Apply_If.h:41-61
// every accumulation a non-forced Set (const components untouched,
// Data_Grid_Speed.tpp:20-26).  Same operation order as the reference.
__global__ void k_forces(int nb, const int* __restrict__ nbr27, const uint8_t* __restrict__ flags,
                         const int8_t* __restrict__ layer, const double* __restrict__ u, const double* __restrict__ v,
                         const double* __restrict__ w, double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
                         SimForce P) {
#pragma clang fp contract(off)
  int b = blockIdx.x;
  const double* U[3] = {u, v, w};
  double* A[3] = {ax, ay, az};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if (!(f & F_EXISTS)) continue;
    int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    double a[3] = {0.0, 0.0, 0.0};
    if (P.accumulate)
      for (int i = 0; i < 3; ++i) a[i] = A[i][c];
    const bool fl = mat_is_fluid(f & 7u);
    const bool lay = layer[c] != -1;
    bool minus_fluid[3];
    for (int i = 0; i < 3; ++i) {
      bool valid;
      minus_fluid[i] = cell_fluid(flags, nbr27, b, x - (i == 0), y - (i == 1), z - (i == 2), valid);
    }
    // gravity
    for (int i = 0; i < 3; ++i) {
      if (P.g[i] == 0.0) continue;
      bool apply = fl || (lay && minus_fluid[i]);
      if (apply && !(f & (0x08u << i))) a[i] = a[i] + P.g[i];
    }
    auto viscosity = [&]() {
      if (P.nu != 0.0) {
        bool apply = fl || (lay && (minus_fluid[0] || minus_fluid[1] || minus_fluid[2]));
        if (apply) {
          for (int i = 0; i < 3; ++i) {
            double temp = 0.0;
            double uc = U[i][c];
            for (int j = 0; j < 3; ++j) {
              double ih2 = P.ih[j] * P.ih[j];
              temp -= 2 * uc * ih2;
              temp += ld_rel(U[i], nbr27, b, x + (j == 0), y + (j == 1), z + (j == 2)) * ih2;
              temp += ld_rel(U[i], nbr27, b, x - (j == 0), y - (j == 1), z - (j == 2)) * ih2;
            }
            if (!(f & (0x08u << i))) a[i] = a[i] + temp * P.nu;
          }
        }
      }
    };
    auto convection = [&]() {
      for (int i = 0; i < 3; ++i) {
        bool apply = fl || (lay && minus_fluid[i]);
        if (!apply) continue;
        double temp = 0.0;
        for (int j = 0; j < 3; ++j) {
          const int ex = (j == 0), ey = (j == 1), ez = (j == 2);
          const int mx = (i == 0), my = (i == 1), mz = (i == 2);
          double Ub;
          if (i == j) {
            Ub = U[j][c];
          } else {
            Ub = 0.25 * (U[j][c] + ld_rel(U[j], nbr27, b, x + ex, y + ey, z + ez) + ld_rel(U[j], nbr27, b, x - mx, y - my, z - mz) +
                         ld_rel(U[j], nbr27, b, x - mx + ex, y - my + ey, z - mz + ez));
          }
          double U1 = U[i][c], U2;
          if (Ub > 0) {
            U2 = ld_rel(U[i], nbr27, b, x - ex, y - ey, z - ez);
            Ub *= (U1 - U2) * P.ih[j];
          } else {
            U2 = ld_rel(U[i], nbr27, b, x + ex, y + ey, z + ez);
            Ub *= (U2 - U1) * P.ih[j];
          }
          temp -= Ub;
        }
        if (!(f & (0x08u << i))) a[i] = a[i] + temp;
      }
    };
    if (P.conv_first) {
      convection();
      viscosity();
    } else {
      viscosity();
      convection();
    }
    for (int i = 0; i < 3; ++i) A[i][c] = a[i];
  }
}

// ------------------------------------------------------------ RK2 stages ----
// mode 0: RK2 midpoint stage 1  u1 = u0 + a dt 0.5           (forced)   RK2.h:41
// mode 1: RK2-TVD stage 1       u1 = u0 + a dt               (forced)   RK2_TVD.h:40-43
// mode 2: RK2 midpoint stage 2  u0 = u0 + a dt               (non-forced) RK2.h:56
// mode 3: RK2-TVD stage 2       u0 = 0.5 u0 + 0.5 u1 + 0.5 a dt (non-forced) RK2_TVD.h:55-58
__global__ void k_rk_combine(int nb, const uint8_t* __restrict__ flags, double* __restrict__ u0, double* __restrict__ v0,
                             double* __restrict__ w0, double* __restrict__ u1, double* __restrict__ v1, double* __restrict__ w1,
                             const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
                             double dt, int mode) {
#pragma clang fp contract(off)
  int b = blockIdx.x;
  double* S0[3] = {u0, v0, w0};
  double* S1[3] = {u1, v1, w1};
  const double* A[3] = {ax, ay, az};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    size_t c = (size_t)b * MAC_BC + t;
    unsigned f = flags[c];
    if (!(f & F_EXISTS)) continue;
    for (int i = 0; i < 3; ++i) {
      const double a = A[i][c];
      const bool cst = f & (0x08u << i);
      switch (mode) {
        case 0: S1[i][c] = S0[i][c] + a * dt * 0.5; break;
        case 1: S1[i][c] = S0[i][c] + a * dt; break;
        case 2:
          if (!cst) S0[i][c] = S0[i][c] + a * dt;
          break;
        default:
          if (!cst) S0[i][c] = 0.5 * S0[i][c] + 0.5 * S1[i][c] + 0.5 * a * dt;
          break;
      }
    }
  }
}

// Algorithms_RungeKutta_all::Do (:35-156), grid part after stage `stage`'s
// forces and projection (a = the projected acceleration):
//   1: S2 = S0 + a dt 0.5, S1 = S0 + a dt 1./6.   (forced: const faces too)
//   2: S2 = S0 + a dt 0.5, S1 = S1 + a dt 1./3.   (forced)
//   3: S2 = S0 + a dt,     S1 = S1 + a dt 1./3.   (forced)
//   4: S0 = S1 + a dt 1./6.                       (const faces keep S0)
// on every existing cell, with the reference's left-to-right products.
__global__ void k_rk4_combine(int nb, const uint8_t* __restrict__ flags, int stage, double dt, double* __restrict__ s0x,
                              double* __restrict__ s0y, double* __restrict__ s0z, double* __restrict__ s1x,
                              double* __restrict__ s1y, double* __restrict__ s1z, double* __restrict__ s2x,
                              double* __restrict__ s2y, double* __restrict__ s2z, const double* __restrict__ ax,
                              const double* __restrict__ ay, const double* __restrict__ az) {
#pragma clang fp contract(off)
  const int b = blockIdx.x;
  double* S0[3] = {s0x, s0y, s0z};
  double* S1[3] = {s1x, s1y, s1z};
  double* S2[3] = {s2x, s2y, s2z};
  const double* A[3] = {ax, ay, az};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t c = (size_t)b * MAC_BC + t;
    const unsigned f = flags[c];
    if (!(f & F_EXISTS)) continue;
    for (int i = 0; i < 3; ++i) {
      const double a = A[i][c];
      switch (stage) {
        case 1:
          S2[i][c] = S0[i][c] + a * dt * 0.5;
          S1[i][c] = S0[i][c] + a * dt * 1. / 6.;
          break;
        case 2:
          S2[i][c] = S0[i][c] + a * dt * 0.5;
          S1[i][c] = S1[i][c] + a * dt * 1. / 3.;
          break;
        case 3:
          S2[i][c] = S0[i][c] + a * dt;
          S1[i][c] = S1[i][c] + a * dt * 1. / 3.;
          break;
        default:
          if (!(f & (0x08u << i))) S0[i][c] = S1[i][c] + a * dt * 1. / 6.;
          break;
      }
    }
  }
}

__global__ void k_zero_accel_existing(int nb, const uint8_t* __restrict__ flags, double* __restrict__ ax, double* __restrict__ ay,
                                      double* __restrict__ az) {
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t c = (size_t)blockIdx.x * MAC_BC + t;
    if (!(flags[c] & F_EXISTS)) continue;
    ax[c] = 0.0;
    ay[c] = 0.0;
    az[c] = 0.0;
  }
}
__global__ void k_zero_accel(size_t n, double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) ax[i] = ay[i] = az[i] = 0.0;
}

// ------------------------------------------------------- interpolation ------

struct SimGrid {
  const uint64_t* hk;
  const int* hv;
  int cap;
  const int* nbr27;  // 27-neighbour brick table (-1: absent)
  const double* U[3];
  double ih[3];
};

// The brick holding a particle's nearest cell, kept between the
// interpolations of one particle: the next position's brick is then a
// neighbour-table read instead of a hash lookup.
struct BrickHint {
  int b = -1;  // brick index, -1 absent
  int c[3] = {0, 0, 0};
  bool valid = false;
};

__device__ inline int hinted_brick(const SimGrid& G, BrickHint& h, const int cb[3]) {
  const int dx = cb[0] - h.c[0], dy = cb[1] - h.c[1], dz = cb[2] - h.c[2];
  int b;
  if (h.valid && h.b >= 0 && dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1)
    b = (dx | dy | dz) == 0 ? h.b : G.nbr27[h.b * 27 + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
  else
    b = dh_lookup(G.hk, G.hv, G.cap, sim_pack(cb[0], cb[1], cb[2]));
  h.b = b;
  h.c[0] = cb[0], h.c[1] = cb[1], h.c[2] = cb[2];
  h.valid = true;
  return b;
}

// Face speed of component k at cell (kx, ky, kz), 0 where no brick holds the
// cell.  The cell lies within one cell of brick cb, whose index is b0: its
// brick comes from b0's neighbour table, or from the hash when b0 is absent.
__device__ inline double speed_near(const SimGrid& G, int k, int b0, const int cb[3], int kx, int ky, int kz) {
  int b;
  if (b0 >= 0) {
    const int sl = (sim_chunk(kx) - cb[0] + 1) + 3 * (sim_chunk(ky) - cb[1] + 1) + 9 * (sim_chunk(kz) - cb[2] + 1);
    b = sl == 13 ? b0 : G.nbr27[b0 * 27 + sl];
  } else {
    b = dh_lookup(G.hk, G.hv, G.cap, sim_pack(sim_chunk(kx), sim_chunk(ky), sim_chunk(kz)));
  }
  return b >= 0 ? G.U[k][(size_t)b * MAC_BC + (kx & 7) + 8 * (ky & 7) + 64 * (kz & 7)] : 0.0;
}

// Policy_Speed_Interpolation_Linear_Symmetric::Get_Speed (:46-119) with the
// recursive Get_Speed_Impl (:13-44) unrolled in the same nesting (x outer,
// z inner) and operation order.
__device__ inline void interp_speed(const SimGrid& G, const double pos[3], double out[3], BrickHint& hint) {
#pragma clang fp contract(off)
  double d[3];
  int key0[3], cb[3];
  for (int i = 0; i < 3; ++i) {
    d[i] = pos[i] * G.ih[i];
    key0[i] = (int)lround(d[i]);
    d[i] -= key0[i];
    cb[i] = sim_chunk(key0[i]);
  }
  // every sample lies within one cell of key0, so inside cb's 27-neighbourhood
  const int b0 = hinted_brick(G, hint, cb);
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
    double w0[3], w1[3];
    for (int i = 0; i < 3; ++i) {
      if (i != k) {
        w0[i] = 1 - d2[i];
        w1[i] = d2[i];
      } else {
        w0[i] = 0.5 - d2[i];
        w1[i] = 0.5 + d2[i];
      }
    }
    double vx[2];
    for (int a = 0; a < 2; ++a) {
      double vy[2];
      for (int bb = 0; bb < 2; ++bb) {
        double s0 = speed_near(G, k, b0, cb, key[0] + a, key[1] + bb, key[2]);
        double s1 = speed_near(G, k, b0, cb, key[0] + a, key[1] + bb, key[2] + 1);
        vy[bb] = w0[2] * s0 + w1[2] * s1;
      }
      vx[a] = w0[1] * vy[0] + w1[1] * vy[1];
    }
    out[k] = w0[0] * vx[0] + w1[0] * vx[1];
  }
}

__device__ inline void interp_speed(const SimGrid& G, const double pos[3], double out[3]) {
  BrickHint h;
  interp_speed(G, pos, out, h);
}

__global__ void k_interpolate(int n, const double* __restrict__ pos, double* __restrict__ vel, SimGrid G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, o[3];
  interp_speed(G, p, o);
  vel[3 * i] = o[0];
  vel[3 * i + 1] = o[1];
  vel[3 * i + 2] = o[2];
}

// Algorithms_Move_Particles::Do (:22-36): rk == 4 is
// Policy_Advance_Ode_RungeKutta::AdvanceSolveODE (:15-27), rk == 2 the
// midpoint rule x += dt v(x + dt v(x) / 2).  Each particle's new position is
// written at its own index; one leaving [lo, hi) gets a NaN x there and is
// counted in *ndropped, for k_list_holes / k_fill_holes to compact.
__global__ void k_move_particles(int np, const double* __restrict__ px, const double* __restrict__ py,
                                 const double* __restrict__ pz, double* __restrict__ qx, double* __restrict__ qy,
                                 double* __restrict__ qz, int* __restrict__ ndropped, SimGrid G, double dt, int rk, int use_box,
                                 double lo0, double lo1, double lo2, double hi0, double hi1, double hi2) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double pos[3] = {px[i], py[i], pz[i]};
  double res[3] = {pos[0], pos[1], pos[2]};
  double v[3], k[3], q[3];
  BrickHint hint;
  if (rk == 4) {
    interp_speed(G, pos, v, hint);
    for (int j = 0; j < 3; ++j) {
      k[j] = dt * v[j];
      res[j] += (1. / 6.) * k[j];
      q[j] = pos[j] + k[j] * 0.5;
    }
    interp_speed(G, q, v, hint);
    for (int j = 0; j < 3; ++j) {
      k[j] = dt * v[j];
      res[j] += (1. / 3.) * k[j];
      q[j] = pos[j] + k[j] * 0.5;
    }
    interp_speed(G, q, v, hint);
    for (int j = 0; j < 3; ++j) {
      k[j] = dt * v[j];
      res[j] += (1. / 3.) * k[j];
      q[j] = pos[j] + k[j];
    }
    interp_speed(G, q, v, hint);
    for (int j = 0; j < 3; ++j) {
      k[j] = dt * v[j];
      res[j] += (1. / 6.) * k[j];
    }
  } else {
    interp_speed(G, pos, v, hint);
    for (int j = 0; j < 3; ++j) {
      k[j] = dt * v[j];
      q[j] = pos[j] + k[j] * 0.5;
    }
    interp_speed(G, q, v, hint);
    for (int j = 0; j < 3; ++j) res[j] = pos[j] + dt * v[j];
  }
  const bool keep =
      !use_box || (res[0] >= lo0 && res[0] < hi0 && res[1] >= lo1 && res[1] < hi1 && res[2] >= lo2 && res[2] < hi2);
  wave_count(!keep, ndropped);
  qx[i] = keep ? res[0] : __builtin_nan("");
  qy[i] = res[1];
  qz[i] = res[2];
}

// Particle part of the same stages (Policy_Move_Particle: the velocity is the
// interpolated grid speed of the stage's slot at the given position; in
// stages 2 and 3 the second update reads it at the position just written):
//   1: v = V(p0);  p2 = p0 + dt 0.5 v;  p1 = p0 + dt 1.0/6.0 v
//   2: p2 = p0 + dt 0.5 V(p2);  p1 = p1 + dt 1.0/3.0 V(p2)
//   3: p2 = p0 + dt V(p2);      p1 = p1 + dt 1.0/3.0 V(p2)
//   4: p0 = p1 + dt 1.0/6.0 V(p2)
__global__ void k_rk4_particles(int np, int stage, double dt, double* __restrict__ p0x, double* __restrict__ p0y,
                                double* __restrict__ p0z, double* __restrict__ p1x, double* __restrict__ p1y,
                                double* __restrict__ p1z, double* __restrict__ p2x, double* __restrict__ p2y,
                                double* __restrict__ p2z, SimGrid G) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double* P0[3] = {p0x, p0y, p0z};
  double* P1[3] = {p1x, p1y, p1z};
  double* P2[3] = {p2x, p2y, p2z};
  double q[3], v[3];
  BrickHint hint;
  if (stage == 1) {
    for (int k = 0; k < 3; ++k) q[k] = P0[k][i];
    interp_speed(G, q, v, hint);
    for (int k = 0; k < 3; ++k) {
      P2[k][i] = q[k] + dt * 0.5 * v[k];
      P1[k][i] = q[k] + dt * 1.0 / 6.0 * v[k];
    }
  } else if (stage == 4) {
    for (int k = 0; k < 3; ++k) q[k] = P2[k][i];
    interp_speed(G, q, v, hint);
    for (int k = 0; k < 3; ++k) P0[k][i] = P1[k][i] + dt * 1.0 / 6.0 * v[k];
  } else {
    for (int k = 0; k < 3; ++k) q[k] = P2[k][i];
    interp_speed(G, q, v, hint);
    for (int k = 0; k < 3; ++k) q[k] = P0[k][i] + (stage == 2 ? dt * 0.5 : dt) * v[k];
    for (int k = 0; k < 3; ++k) P2[k][i] = q[k];
    interp_speed(G, q, v, hint);
    for (int k = 0; k < 3; ++k) P1[k][i] = P1[k][i] + dt * 1.0 / 3.0 * v[k];
  }
}

// Particles outside [lo, hi) are deleted (slot 0 only; the other slots are
// rewritten from slot 0 at the next step's first stage): their x becomes NaN
// in place and they are counted in *ndropped, for k_list_holes /
// k_fill_holes to compact.
__global__ void k_box_delete(int np, double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
                             int* __restrict__ ndropped, double lo0, double lo1, double lo2, double hi0, double hi1, double hi2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double x = px[i], y = py[i], z = pz[i];
  const bool keep = x >= lo0 && x < hi0 && y >= lo1 && y < hi1 && z >= lo2 && z < hi2;
  wave_count(!keep, ndropped);
  if (!keep) px[i] = __builtin_nan("");
}


// ------------------------------------------------------ extrapolation ----
// Algorithms_Extrapolate::Do (Algorithms_Extrapolate.h:25-235), one frontier
// step per launch.  A cell is "ok" when it exists and its layer is not empty.
// For every frontier cell A (layer pass `lay`) and direction i:
//   +1: B = A+e_i ok with layer >= lay: face A|B (stored at B) := mean of the
//       faces parallel to it that touch a cell of layer < lay; B joins the
//       next frontier when its layer > lay;
//   -1: C = A-e_i ok with layer >= lay: face C|A (stored at A), likewise;
//       C missing or empty: A's -e_i face from its +e_i and transverse
//       neighbours.
// Non-forced Set (const faces keep their value); a mean over no face is 0/0.
// Within one pass the faces written join two cells of layer >= lay and the
// faces read touch a cell of layer < lay, so threads never read what others
// write; a face reached from both of its cells gets the same sum in the same
// order from each, so the duplicate writes agree bit for bit.
struct ExCell {
  long long c;  // slot, -1 when the brick is missing
  int lay;      // layer; -2 when not ok
};

__device__ inline ExCell ex_cell(const uint8_t* flags, const int8_t* layer, const int* nbr27, int b, int x, int y, int z) {
  int cell;
  const int nb = locate(nbr27, b, x, y, z, cell);
  ExCell e{-1, -2};
  if (nb < 0) return e;
  e.c = (long long)nb * MAC_BC + cell;
  if ((flags[e.c] & F_EXISTS) && layer[e.c] != -1) e.lay = layer[e.c];
  return e;
}

__global__ void k_extrap_seed(int nb, const uint8_t* __restrict__ flags, const int8_t* __restrict__ layer, uint8_t* __restrict__ front) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < (long long)nb * MAC_BC; c += (long long)gridDim.x * blockDim.x)
    front[c] = ((flags[c] & F_EXISTS) && layer[c] == 1) ? 1 : 0;
}

__global__ void k_extrapolate(int nb, const int* __restrict__ nbr27, const uint8_t* __restrict__ flags, const int8_t* __restrict__ layer,
                              double* u, double* v, double* w, const uint8_t* __restrict__ front, uint8_t* __restrict__ next,
                              int* __restrict__ any_next, int lay) {
#pragma clang fp contract(off)
  const int b = blockIdx.x;
  double* U[3] = {u, v, w};
  for (int t = threadIdx.x; t < MAC_BC; t += blockDim.x) {
    const size_t ca = (size_t)b * MAC_BC + t;
    if (!front[ca]) continue;
    const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    auto at = [&](int dx, int dy, int dz) { return ex_cell(flags, layer, nbr27, b, x + dx, y + dy, z + dz); };
    auto set = [&](size_t c, int i, double val) {
      if (!(flags[c] & (0x08u << i))) U[i][c] = val;
    };
    for (int i = 0; i < 3; ++i) {
      const int ex = (i == 0), ey = (i == 1), ez = (i == 2);
      // +1 direction
      const ExCell B = at(ex, ey, ez);
      if (B.lay >= lay) {
        if (B.lay > lay) {
          next[B.c] = 1;
          *any_next = 1;
        }
        int n = 0;
        double val = 0.0;
        for (int j = 0; j < 3; ++j)
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            const int jx = s2 * (j == 0), jy = s2 * (j == 1), jz = s2 * (j == 2);
            if (s2 == -1 && j == i) {
              const ExCell C = at(-ex, -ey, -ez);
              if (C.lay != -2 && C.lay < lay) {
                ++n;
                val += U[i][ca];
              }
            } else if (s2 == 1 && j == i) {
              const ExCell D = at(2 * ex, 2 * ey, 2 * ez);
              if (D.lay != -2 && D.lay < lay) {
                ++n;
                val += U[i][D.c];
              }
            } else {
              const ExCell N3 = at(ex + jx, ey + jy, ez + jz), N4 = at(jx, jy, jz);
              if (N3.lay != -2 && N4.lay != -2 && (N3.lay < lay || N4.lay < lay)) {
                ++n;
                val += U[i][N3.c];
              }
            }
          }
        set(B.c, i, val / n);
      }
      // -1 direction
      const ExCell C = at(-ex, -ey, -ez);
      if (C.lay != -2) {
        if (C.lay >= lay) {
          if (C.lay > lay) {
            next[C.c] = 1;
            *any_next = 1;
          }
          int n = 0;
          double val = 0.0;
          for (int j = 0; j < 3; ++j)
            for (int s2 = -1; s2 <= 1; s2 += 2) {
              const int jx = s2 * (j == 0), jy = s2 * (j == 1), jz = s2 * (j == 2);
              if (s2 == 1 && j == i) {
                const ExCell N3 = at(ex, ey, ez);
                if (N3.lay != -2 && N3.lay < lay) {
                  ++n;
                  val += U[j][N3.c];
                }
              } else if (s2 == -1 && j == i) {
                const ExCell N3 = at(-2 * ex, -2 * ey, -2 * ez);
                if (N3.lay != -2 && N3.lay < lay) {
                  ++n;
                  val += U[j][C.c];
                }
              } else {
                const ExCell N3 = at(-ex + jx, -ey + jy, -ez + jz), N4 = at(jx, jy, jz);
                if (N3.lay != -2 && N4.lay != -2 && (N3.lay < lay || N4.lay < lay)) {
                  ++n;
                  val += U[i][N4.c];
                }
              }
            }
          set(ca, i, val / n);
        }
      } else {
        int n = 0;
        double val = 0.0;
        for (int j = 0; j < 3; ++j)
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            const int jx = s2 * (j == 0), jy = s2 * (j == 1), jz = s2 * (j == 2);
            if (s2 == 1 && j == i) {
              const ExCell N3 = at(ex, ey, ez);
              if (N3.lay != -2 && N3.lay < lay) {
                ++n;
                val += U[j][N3.c];
              }
            } else if (s2 == -1 && j == i) {
            } else {
              const ExCell N4 = at(jx, jy, jz);
              if (N4.lay != -2) {
                const ExCell N3 = at(jx - ex, jy - ey, jz - ez);
                if (N3.lay != -2 && (N3.lay < lay || N4.lay < lay)) {
                  ++n;
                  val += U[i][N4.c];
                }
              }
            }
          }
        set(ca, i, val / n);
      }
    }
  }
}
