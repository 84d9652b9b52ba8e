/*
 * mac_amd.h -- C-ABI of the MI355X-native MAC-grid projection library
 * (libmac_amd.so).  Plain pointers and sizes only; every host array is owned
 * by the caller and copied in or out, every device buffer is owned by the
 * context.  One context per host thread; all work of a context is issued on
 * its own HIP stream and serialises there.  Every entry point returns 0 on
 * success and a negative MAC_E_* code on failure (the reason is available
 * from mac_last_error()).
 *
 * Grid format (shared by every entry point that takes cells):
 *   The grid is a set of 8x8x8 bricks ("chunks" in the reference).  Brick b
 *   has integer chunk key keys3[3b+0..2]; its cells are stored x-fastest,
 *   cell (x,y,z) at offset x + 8y + 64z (reference Offset.h:14-24), and a
 *   cell with global integer key k lives in chunk floor(k/8)
 *   (ChunkHashTable.tpp:181-196).  Per cell one flag byte:
 *     bits 0..2  material, MaterialToInt numbering
 *                (Data_CellType_Fluid_Air_Solid_Inflow_Interior_0_Pres.tpp:3-21):
 *                Air=0 Fluid=1 Solid=2 Inflow=3 Interior=4 _0_Pres=5
 *     bits 3..5  const bit of the x/y/z face on the cell's low side
 *                (Data_Grid_Speed m_const, Data_Grid_Speed.tpp:1-26)
 *     bit  7     cell exists (the chunk's existence bitset,
 *                Data_Chunk_Bool_Array.h:8-28); a cell without it is an
 *                invalid neighbour (Offset_Base_Neighbour.h:14-33)
 *   Face velocities u,v,w are per cell, the face on the cell's low side in
 *   that direction (staggered-left MAC layout, Policy_Divergence.h:20-29).
 */
#ifndef MAC_AMD_H
#define MAC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAC_OK 0
#define MAC_E_HIP -1      /* a HIP runtime call failed                        */
#define MAC_E_ARG -2      /* bad argument / inconsistent sizes                */
#define MAC_E_NOCONV -3   /* the solver hit max_iter before rtol              */
#define MAC_E_STATE -4    /* call order violated (e.g. project before grid)   */
#define MAC_E_NODEV -5    /* no GPU visible                                   */

#define MAC_FLAG_MAT_MASK 0x07u
#define MAC_FLAG_CONST_X 0x08u
#define MAC_FLAG_CONST_Y 0x10u
#define MAC_FLAG_CONST_Z 0x20u
#define MAC_FLAG_EXISTS 0x80u

enum mac_material { MAC_AIR = 0, MAC_FLUID = 1, MAC_SOLID = 2, MAC_INFLOW = 3, MAC_INTERIOR = 4, MAC_0_PRES = 5 };

enum mac_precision {
  MAC_FP64 = 0,  /* FP64 V-cycle inside FP64 CG                              */
  MAC_MIXED = 1  /* FP32 V-cycle (smooth/restrict/prolong) inside FP64 CG    */
};

enum mac_smoother {
  MAC_SMOOTH_JACOBI = 0,    /* nu damped-Jacobi sweeps with weight omega                  */
  MAC_SMOOTH_CHEBYSHEV = 1  /* nu Jacobi sweeps with the Chebyshev weights of [a, 2]:
                               omega_k = 1 / (1 + a/2 + (1 - a/2) cos((2k-1) pi / (2 nu))),
                               k = 1..nu before the coarse correction, reversed after      */
};

typedef struct mac_ctx mac_ctx;

typedef struct {
  double h[3];     /* cell spacing per direction; 1/h is formed as 1.0/h like
                      Inversible_Value (Inversible_Value_Vect.h)                  */
  double rtol;     /* stop when ||f - L p||_2 <= rtol * ||f||_2                    */
  int max_iter;    /* PCG iteration cap                                            */
  int precision;   /* enum mac_precision                                           */
  int nu;          /* damped-Jacobi sweeps before and after each coarse correction */
  double omega;    /* Jacobi damping                                               */
  int check_every; /* iterations before the host's first poll of the device
                      convergence flag; later polls come after as many
                      iterations as the residual decay so far predicts are
                      still needed.  From the third solve on the same grid
                      (kernel timing off) each iteration is one captured HIP
                      graph launch                                                */
  int stall_iters; /* guided solve (guess = 1): restart from zero if not converged
                      after this many iterations (0: never; the host polls at
                      that iteration); a guess whose first residual is no
                      smaller than ||f|| is dropped at once                       */
  int smoother;    /* enum mac_smoother                                            */
  double cheb_lower; /* MAC_SMOOTH_CHEBYSHEV: lower end a of the eigenvalue range
                      [a, 2] of D^-1 A that the nu sweeps damp                    */
  int nu_coarse;   /* smoothing launches before and after the correction on the
                      coarse levels (0: nu); they are short, latency-bound
                      launches                                                    */
  double cheb_lower_coarse; /* a of the coarse levels' Chebyshev weights          */
  int coarse_tail; /* 1 (default): the last two levels of each V-cycle (the top brick
                      and its at most 8 children) run as one single-workgroup launch
                      with the level in LDS; 0: one launch per step                */
  double viscosity; /* > 0: the viscous free-surface pressure of Policy_Dirichlet_Boundary
                      (Algorithms_Solve_Pressure_Dirichlet): every air or solid cell
                      next to a non-air cell holds p = -sum 2 nu (u_i(c+e_i) - u_i(c))
                      / h_i * s over its non-air neighbours, which enters the right
                      side, the written pressure and the velocity update; 0: p = 0 */
  int coarse_inner; /* > 1 (default 4): each smoothing launch on the coarse levels
                      other than the last two is a block-Jacobi step, the brick's unknowns taking this many
                      Chebyshev sweeps ([cheb_lower_coarse, 2]) among themselves
                      with the neighbour bricks frozen (nu_coarse counts launches);
                      1: point Jacobi                                             */
  double coarse_damp; /* the block step's correction is scaled by this (0, 1]     */
  int fuse_coarse;  /* 1 (default): each restriction into a block-smoothed level also
                      takes that level's first smoothing step (from zero, brick-local),
                      and each prolongation takes the first post-smoothing step of the
                      level it lands on (the prolongated face planes of the neighbour
                      bricks formed in registers; on level 0 in the FP32 V-cycle
                      only); 0: separate launches                                 */
} mac_params;

typedef struct {
  int iters;             /* PCG iterations in total (a dropped guided start included) */
  int restarts;          /* true-residual restarts taken                         */
  int fallback_used;     /* 1 when the guided solve stalled and was redone       */
  int levels;            /* multigrid levels                                     */
  long long n_unknowns;  /* in-domain cells                                      */
  long long n_bricks;    /* bricks at level 0 / active bricks at level 0         */
  long long n_active_bricks;
  double rel_resid;      /* recursive residual ||r||/||f|| at exit               */
  double true_rel_resid; /* recomputed ||f - L p|| / ||f||                       */
  double max_div;        /* max |div u| over in-domain cells after the update    */
  double ms_total;       /* wall time of the projection (device-resident)        */
  double ms_solve;       /* PCG part only                                        */
} mac_stats;

/* Defaults: h=1, rtol=1e-8, max_iter=500, FP64, nu=2, omega=0.9, check_every=4, stall_iters=0,
 * Chebyshev smoothing with a = 0.25 on level 0; below it nu_coarse=2 block-Jacobi
 * launches of coarse_inner=4 Chebyshev sweeps on [0.05, 2], damped by 0.8 (from
 * level 3 on one undamped launch of nu_coarse * coarse_inner <= 8 sweeps); the
 * last two levels fused (coarse_tail=1); viscosity 0. */
void mac_default_params(mac_params* prm);

/* Create / destroy a context on the current HIP device. */
int mac_create(mac_ctx** out, const mac_params* prm);
void mac_destroy(mac_ctx* ctx);
int mac_set_params(mac_ctx* ctx, const mac_params* prm);
const char* mac_last_error(void);

/* Upload a brick grid (keys + flags) and build the device brick hash, the
 * 27-neighbour brick table, the per-cell stencil codes and the multigrid
 * hierarchy.  Replaces the reference's per-projection matrix assembly
 * (Algorithms_Solve_Pressure.h:44-157: the operator is applied matrix-free
 * from the codes instead of being stored).  u,v,w may be NULL (zeroed). */
int mac_upload_bricks(mac_ctx* ctx, int nbricks, const int32_t* keys3, const uint8_t* flags,
                      const double* u, const double* v, const double* w);

/* Host <-> device copies of the cell fields, brick layout, nbricks*512 each. */
int mac_set_velocity(mac_ctx* ctx, const double* u, const double* v, const double* w);
int mac_get_velocity(mac_ctx* ctx, double* u, double* v, double* w);
/* The real velocity the viscous air pressures are formed from
 * (Policy_Dirichlet_Boundary.h:84-109 reads Speed_GetRef(-1,false), the
 * speed slot, while Algorithms_Solve_Pressure_Force.h:16-24 has redirected
 * Speed_GetRef() to the acceleration that is being projected).  Brick
 * layout, nbricks*512 each, copied to the device; it holds for the current
 * grid until the next mac_upload_bricks.  All three NULL: the air pressures
 * are formed from the projected field again (the default).  Only read when
 * mac_params.viscosity > 0. */
int mac_set_real_velocity(mac_ctx* ctx, const double* u, const double* v, const double* w);
int mac_get_pressure(mac_ctx* ctx, double* p);
int mac_set_pressure(mac_ctx* ctx, const double* p);
/* Device-side snapshot / restore of u,v,w (to re-run a projection on the same input). */
int mac_save_velocity(mac_ctx* ctx);
int mac_restore_velocity(mac_ctx* ctx);

/* One free-surface projection on the device-resident fields, the
 * Divergence_Projection of Algorithms_Solve_Pressure.h:168-221:
 *   f = div(u) per in-domain cell            (Policy_Divergence.h:20-29)
 *   solve A p = f, A the 7-point operator of Init_Iteration (air / missing
 *     neighbours Dirichlet p=0, const faces Neumann), by MG-preconditioned CG
 *   p := solution in the domain, 0 elsewhere     (:184-197)
 *   u_i -= p(c)/h_i - p(c-e_i)/h_i on correctable, non-const faces (:198-207,
 *     Policy_Gradiant.h:32-37, Policy_Pressure_If_Correction.h:33-49)
 *   max |div u| over the domain                  (:209-219, reported)
 * guess: 0 = zero start, 1 = start from the pressure currently held on the
 * device (previous step or mac_set_pressure).  stats may be NULL. */
int mac_project(mac_ctx* ctx, int guess, mac_stats* stats);

/* Pieces of the projection, for parity tests (all on device-resident data):
 * div_out[c] = Policy_Divergence::Get_Divergence(c) for in-domain cells, else 0. */
int mac_divergence(mac_ctx* ctx, double* div_out);
/* y = A x for the assembled operator (reference sign: diagonal negative),
 * x,y host arrays in brick layout; entries outside the domain read as 0. */
int mac_apply_operator(mac_ctx* ctx, const double* x, double* y);
/* u -= grad p using the device pressure (the write-back + update half of
 * Divergence_Projection). */
int mac_gradient_update(mac_ctx* ctx);
/* max |div u| over in-domain cells of the device velocity. */
int mac_max_divergence(mac_ctx* ctx, double* out);

/* Slot A of the reference (Policy_Solve_Linear_Umfpack.h:7-71): solve
 * A x = b for a matrix in compressed columns, UMFPACK's layout (offset[n+1]
 * column pointers, indice[nnz] row indices, value[nnz]).  n == 0 is a no-op;
 * n <= 3 is solved on the host by Gaussian elimination with partial
 * pivoting (the non-symmetric 3x3 corner systems of the extrapolation
 * boundary condition, Algorithm_Extrapolate_Boundary_Condition.h:557-576);
 * larger systems must be symmetric (the pressure system of
 * Algorithms_Solve_Pressure::Init_Iteration, where columns and rows
 * coincide) and are solved by Jacobi-preconditioned CG in FP64 on the device
 * to rtol. */
int mac_solve_csr(mac_ctx* ctx, int n, const int* offset, const int* indice, const double* value,
                  const double* b, double* x, double rtol);

/* The slot A solve without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and answer as mac_solve_csr, but every array is a caller-owned
 * DEVICE array (d_offset[n+1], d_indice/d_value[nnz], d_b and d_x[n]; the CSR
 * is checked on the device: off[0] = 0, off[n] = nnz, non-decreasing offsets,
 * column indices in [0, n)), and the solve's work arrays live in a fixed
 * static device array of the library (about 21.1 MB per GPU, held until the
 * process ends).  n <= MAC_CSR_NOALLOC_MAX_N, larger n is rejected with
 * MAC_E_ARG.  Calls that overlap on one GPU share that array and race: issue
 * them from one thread, or serialise them.  Runs on the context's stream and
 * returns when x is written; *iters (may be NULL) gets the CG iterations. */
#define MAC_CSR_NOALLOC_MAX_N 262144
int mac_solve_csr_noalloc(mac_ctx* ctx, int n, int nnz, const int* d_offset, const int* d_indice, const double* d_value,
                          const double* d_b, double* d_x, double rtol, int* iters);

/* Kernel timing for roofline figures: when enabled, mac_project records HIP
 * events around every launch of the named kernel classes on the context's
 * stream and accumulates their durations. */
int mac_enable_kernel_timing(mac_ctx* ctx, int on);
/* Returns the number of kernel classes; fills up to cap entries. */
int mac_kernel_times(mac_ctx* ctx, int cap, const char** names, double* total_ms, long long* launches,
                     double* bytes);
int mac_reset_kernel_times(mac_ctx* ctx);
/* Mean duration, in microseconds, of an event pair around an empty launch on
 * the context's stream (n launches): the part of each event-timed launch
 * above that is not the kernel's own run time.  No counterpart in the
 * reference; a measurement aid for the figures above. */
int mac_timing_overhead(mac_ctx* ctx, int n, double* us);

/* Synchronise the context's stream. */
int mac_sync(mac_ctx* ctx);

/* ===================================================================== */
/* Per-timestep MAC-grid loop on a sparse device brick grid.              */
/* ===================================================================== */
/* The grid is a device hash of chunk keys -> 8^3 bricks (ChunkHashTable,
 * ChunkHashTable.tpp:63-161) with a 27-neighbour brick table; bricks are
 * created when particles or air layers reach them and freed when their last
 * cell is deleted.  Flag byte as above plus bit 6 = NoDelete
 * (Data_CellType_Fluid_Air_Solid_Inflow_Interior_0_Pres.h:354-361).  Per
 * cell: layer (int8, -1 = empty, Data_Grid_Layer_Empty), two speed slots and
 * one acceleration slot (Data_Chunk_Speed.h:5-96), pressure.  Particles are
 * positions (x,y,z), FP64. */
#define MAC_FLAG_NODELETE 0x40u

typedef struct mac_sim mac_sim;

typedef struct {
  double h[3];
  double gravity[3];    /* Policy_Gravity(g, i) for every axis i with g != 0          */
  double viscosity;     /* nu of Algorithms_Viscosity; 0 disables the term             */
  double cfl;           /* m_factor of Algorithms_Calculate_Time_Step.h:35-66          */
  double dt_min, dt_max;/* Policy_CheckDT.h:16-29                                      */
  int layer_max;        /* Policy_Layer_Max: air layers built around the fluid         */
  int integrator;       /* 0 = RK2 midpoint (Algorithms_RungeKutta_RK2.h:20-64),
                           1 = RK2-TVD (Algorithms_RungeKutta_RK2_TVD.h:20-66),
                           2 = RK4 "all" (Algorithms_RungeKutta_all.h:35-156): four
                           stages, each with the cell-type update from the stage's
                           particle positions, velocity extrapolation, forces, the
                           projection of the acceleration and its extrapolation;
                           particles advance inside the stages                        */
  int particle_rk;      /* 2 = RK2 midpoint, 4 = classical RK4
                           (Policy_Advance_ODE_RungeKutta.h:15-27)                    */
  int particles_per_inflow_cell; /* seeded-jitter fill of empty inflow cells           */
  unsigned long long seed;
  double box_lo[3], box_hi[3];   /* particles leaving [lo, hi) are deleted (lo == hi: none) */
  int max_particles;
  int max_bricks;
  int depth_interior;   /* 0: the cell-type update without depth (TestJetDEau_Basic.cpp:324-326);
                           >= 1: the depth chain of test/Jet3.cpp:523-547 with this
                           DepthInterior (Policy_Depth.h, Jet3: 3)                    */
  int depth_remove;     /* DepthRemoveParticle of Policy_Depth (Jet3: 5)               */
  int floor_axis;       /* Algorithm_DelParticle (test/Jet3.cpp:259-289): particles with
                           position[floor_axis] < floor_value are erased at the start of
                           the cell-type update; -1 = none (Jet3: axis y, value 0)    */
  double floor_value;
  int fill_cube_n;      /* > 0: new particles on the (2n+1)^3 lattice of
                           Policy_Add_Particle_Cube (Jet3: n = 1) instead of the
                           particles_per_inflow_cell jitter                           */
  /* Inflow pulsation, the seeded counterpart of Algorithm_Init_Perturbation
     (test/Jet.cpp:188-256): after the time step of every step, the const
     faces of every existing Inflow cell get inflow_speed[i] * (1 + inflow_amp
     * sin(2 pi n / inflow_period + phi)), n the step number (1 in the first
     step), phi = 2 pi U(seed).  inflow_amp = 0: off.                           */
  double inflow_speed[3];
  double inflow_amp;
  int inflow_period;
  int surface_bc;       /* 1: every velocity extrapolation (mac_sim_extrapolate and,
                           with integrator 2, those of the speeds and of the
                           acceleration in every stage) first applies the
                           free-surface velocity condition
                           (Algorithm_Extrapolate_Boundary_Condition, chained before
                           Algorithms_Extrapolate as test/Jet3.cpp:541-550,636-639
                           does); 0: off.  Integrators 0 and 1 extrapolate nothing
                           (TestJetDEau_Basic.cpp's chain).                          */
  mac_params proj;      /* projection solver settings                                  */
} mac_sim_params;

typedef struct {
  double dt;            /* time step used                                              */
  int nbricks;          /* bricks after the topology update                            */
  int nparticles;       /* particles after the move                                    */
  long long nfluid;     /* in-domain cells                                             */
  int pcg_iters[2];     /* PCG iterations of the two projections (RK4 "all": of stages
                           1-2 and of stages 3-4)                                       */
  double max_div;       /* max |div u| after the last projection                       */
  /* host wall time per part (ms): cell-type updates (RK4 "all": the step's and
     the four stages'), forces and velocity updates, projections, particle
     moves (RK4 "all": the box deletion; the stages move the particles)       */
  double ms_topology, ms_forces, ms_project, ms_particles, ms_total;
  double true_rel_resid; /* largest recomputed ||f - L p|| / ||f|| over the step's
                            projections                                            */
} mac_sim_stats;

void mac_sim_default_params(mac_sim_params* p);
int mac_sim_create(mac_sim** out, const mac_sim_params* p);
void mac_sim_destroy(mac_sim* sim);
/* Upload a state: bricks (keys, flags, optional layers, speed slot 0) and particles
 * (pxyz interleaved x,y,z per particle). */
int mac_sim_upload(mac_sim* sim, int nbricks, const int32_t* keys3, const uint8_t* flags, const int8_t* layer,
                   const double* u, const double* v, const double* w, int nparticles, const double* pxyz);
/* Bricks and particles currently held. */
int mac_sim_counts(mac_sim* sim, int* nbricks, int* nparticles);
/* Steps taken (the counter the seeded inflow jitter and the inflow pulsation
 * read: the fill of step n draws from (seed, n), the pulsation of step n is
 * sin(2 pi n / period + phi)).  A run resumed from a checkpoint sets it to the
 * checkpoint's step so that it continues the same sequence; mac_sim_upload
 * leaves it unchanged.  No counterpart in the reference, whose jitter is
 * std::rand's global state. */
int mac_sim_get_step(mac_sim* sim, unsigned long long* step);
int mac_sim_set_step(mac_sim* sim, unsigned long long step);
/* Download the state; bricks come in ascending (z, y, x) chunk-key order.  Any
 * pointer may be NULL.  a* = acceleration slot, p = pressure. */
int mac_sim_download(mac_sim* sim, int32_t* keys3, uint8_t* flags, int8_t* layer, double* u, double* v, double* w,
                     double* ax, double* ay, double* az, double* p, double* pxyz);
/* Per-cell depth (int16, -1 = empty, Data_Grid_Depth_Empty.h), bricks in the
 * order mac_sim_download gives them.  Only the depth chain reads and writes it;
 * set it after mac_sim_upload to start from Interior cells. */
int mac_sim_get_depth(mac_sim* sim, int16_t* depth);
int mac_sim_set_depth(mac_sim* sim, const int16_t* depth);
/* The cell-type/layer update of one step, in the order of the reference's
 * init chain (TestJetDEau_Basic.cpp:324-326): Algorithms_Initialize_MacCell,
 * Algorithms_Layer_Initial_With_Particle (particle -> key = lround(x/h),
 * Policy_Particle_To_Key.h:24-35), Algorithms_Create_Fluid_Particle (empty
 * inflow cells get particles), Algorithms_Update_CellType_Layer (air layers
 * 1..layer_max by BFS, then the +1-direction extra layer),
 * Algorithms_Delete_MacCell (layer-empty cells erased, empty bricks freed).
 * With depth_interior >= 1 it is the depth chain of test/Jet3.cpp:545-547:
 * Algorithms_Initialize_Layer_And_Depth, Algorithm_DelParticle (floor_axis),
 * Algorithms_Layer_Initial_With_Particle_Depth (particles in Interior cells
 * deeper than depth_remove are erased), Algorithms_Create_Fluid_Particle_Depth
 * (empty inflow cells, and empty Interior cells no deeper than depth_remove,
 * get particles), Algorithms_Update_CellType_Layer_Depth (surface depth 0, air
 * layers, then the depth search inside the fluid: depth < depth_interior gives
 * Fluid, deeper cells become Interior), Algorithms_Delete_MacCell. */
int mac_sim_topology(mac_sim* sim);
/* Algorithms_Calculate_Time_Step + Policy_CheckDT. */
int mac_sim_compute_dt(mac_sim* sim, double* dt);
/* acceleration := 0, then gravity, viscosity and first-order upwind
 * convection of speed slot `slot` (Algorithms_Gravity/Viscosity/Convection). */
int mac_sim_forces(mac_sim* sim, int slot);
/* One RK2 / RK2-TVD velocity step with both projections. */
int mac_sim_advance(mac_sim* sim, double dt, mac_sim_stats* stats);
/* Algorithms_Move_Particles with the interpolated velocity of slot 0, then
 * the box deletion. */
int mac_sim_move_particles(mac_sim* sim, double dt);
/* Policy_Speed_Interpolation_Linear_Symmetric::Get_Speed of slot 0 at n
 * positions (xyz interleaved in and out). */
int mac_sim_interpolate(mac_sim* sim, int n, const double* pos, double* vel);
/* One whole time step: topology, dt, advance, move particles. */
int mac_sim_step(mac_sim* sim, mac_sim_stats* stats);
/* Algorithms_Extrapolate::Do (Algorithms_Extrapolate.h:25-235) on speed slot
 * `slot`: starting from the cells of layer 1 and moving out layer by layer,
 * each face between two cells of layer >= L gets the mean of the parallel
 * faces next to it that touch a cell of layer < L (non-forced: const faces
 * keep their value; a mean over no face is 0/0 = NaN, as in the reference). */
int mac_sim_extrapolate(mac_sim* sim, int slot);
/* Algorithm_Extrapolate_Boundary_Condition::Do
 * (Algorithm_Extrapolate_Boundary_Condition.h:44-986) on speed slot `slot`
 * (0..2) or, slot 3, the acceleration (Algorithms_Accel_To_Speed); dim 3, or
 * 2 for Algorithm_Extrapolate_Boundary_Condition_Dim<2> on the x-y plane (z
 * faces untouched).  The reference's rules run in a defined order of seven
 * phases, each reading what the phases before it wrote (set NaN between air
 * cells; the zero-divergence face of one-air cells; the zero-shear
 * tangential faces of face-cell pairs; 45-degree corners; the 3x3 corner
 * solves; NaN on air faces outside the domain; apply_nan_extrap), so the
 * result does not depend on the order of the bricks.  Non-forced: const
 * faces keep their value.  Faces between two non-air cells are read, never
 * written; they must be finite. */
int mac_sim_surface_bc(mac_sim* sim, int slot, int dim);

/* ===================================================================== */
/* Learned initial guess: 3-level 3-D U-Net (no counterpart in the        */
/* reference; it only seeds the PCG of mac_project, which then solves to   */
/* the same tolerance).  16/32/64 channels, 3x3x3 convolutions on the BF16 */
/* matrix cores with FP32 accumulation, ReLU, 2x2x2 average pooling,       */
/* 2x2x2 transposed-convolution upsampling, skip concatenation, 1x1x1 head.*/
/* Volumes are D x H x W voxels, z slowest, x fastest; D and H multiples   */
/* of 4, W a multiple of 64.  Input: 5 FP32 channels per voxel             */
/* (divergence, fluid, air, solid masks, previous pressure), [V][5].       */
/* ===================================================================== */
typedef struct mac_unet mac_unet;
int mac_unet_create(mac_unet** out, int D, int H, int W, unsigned long long seed);
void mac_unet_destroy(mac_unet* net);
long long mac_unet_num_params(mac_unet* net);
/* FP32 master parameters, layer by layer (weights [Cout][27 Cin] with K
 * ordered (tap, cin), tap = 9(dz+1)+3(dy+1)+(dx+1), then biases). */
int mac_unet_get_params(mac_unet* net, float* out);
int mac_unet_set_params(mac_unet* net, const float* in);
int mac_unet_forward(mac_unet* net, const float* in5, float* out);
/* Parameter gradients of sum(out * gout) for the last forward. */
int mac_unet_backward(mac_unet* net, const float* gout, float* grads);
/* One Adam step on the masked mean-squared error to target. */
int mac_unet_train_step(mac_unet* net, const float* in5, const float* target, const float* mask, float lr, float* loss);
/* Event-timed forward convolutions: returns the time and FLOPs accumulated
 * since the previous call, then switches timing on or off. */
int mac_unet_conv_timing(mac_unet* net, int on, double* ms, double* flop);
/* Weight gradients of the 3x3x3 convolutions: mode 0 (default) on the BF16
 * matrix cores (implicit GEMM over the voxels, FP32 accumulate; network
 * volume at most 128 wide), mode 1 on FP32 FMA (the same sums in another
 * order, kept to check the first). */
int mac_unet_wgrad_mode(mac_unet* net, int mode);

/* The network on a projection context.  The grid's bounding box (chunk keys
 * shifted to a minimum of 0 per axis) is mapped to the network volume, cell
 * (8kx+x, 8ky+y, 8kz+z); it must fit.  Cells outside the grid are solid.
 * Scaling: s = rms of the right-hand side over the unknowns, hm = mean of
 * 1/h^2 over the axes; inputs f/s and prev*hm/s, output p*hm/(s*out_scale).
 *
 * mac_unet_guess: the learned starting iterate for mac_project(ctx, 1, ...)
 * from the context's current face velocities, written into its pressure.
 * use_prev = 1 takes the context's pressure as it stands as the previous
 * pressure input, 0 a zero one. */
int mac_unet_guess(mac_unet* net, mac_ctx* ctx, int use_prev, float out_scale);
/* Training pair of the context's last mac_project: in5 [V][5] for its
 * right-hand side with prev (host, brick layout; NULL = zero) as the previous
 * pressure, target [V] its pressure in output units, mask [V] 1 on unknowns. */
int mac_unet_sample(mac_unet* net, mac_ctx* ctx, const double* prev, float out_scale, float* in5, float* target, float* mask);
/* Starting iterate of the simulation's projections: mode 0 zero, 1 the last
 * projection's pressure, 2 the learned guess of net with that pressure as its
 * previous-pressure input (the network stays owned by the caller). */
int mac_sim_set_guess(mac_sim* sim, int mode, mac_unet* net, float out_scale);
/* Training pair of the simulation's last projection (previous pressure = the
 * pressure that projection started from); outputs as for mac_unet_sample. */
int mac_sim_capture(mac_sim* sim, mac_unet* net, float out_scale, float* in5, float* target, float* mask);

#ifdef __cplusplus
}
#endif
#endif
