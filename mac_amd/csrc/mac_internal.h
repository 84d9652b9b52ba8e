// mac_internal.h -- entry points shared between the translation units of
// libmac_amd.so (not exported).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mac_amd.h"

#define MAC_HIDDEN __attribute__((visibility("hidden")))

extern "C" {

// (Re)build a projection context's brick hierarchy from host chunk keys and a
// flag array that may live on the device or the host.
MAC_HIDDEN int mac_internal_build(mac_ctx* c, int nbricks, const int32_t* keys3_host, const uint8_t* flags_any);
// Project device-resident face velocities (brick layout of the last build) in
// place and write the pressure to p (device).  guess = 1 starts from p,
// guess = 2 from the context's pressure as it stands (a learned guess).
MAC_HIDDEN int mac_internal_project(mac_ctx* c, double* u, double* v, double* w, double* p, int guess, mac_stats* st);
MAC_HIDDEN int mac_internal_set_error(int code, const char* msg);
// The context's stream and device ordinal.
MAC_HIDDEN hipStream_t mac_internal_stream(mac_ctx* c);
MAC_HIDDEN int mac_internal_device(mac_ctx* c);
// Device workspace of at least `bytes` for the CSR solve (slot A), owned by
// the context: grown when a call needs more, freed by mac_destroy.  *pinned
// receives a 256-byte pinned host block of the context for the solve's state.
MAC_HIDDEN int mac_internal_csr_workspace(mac_ctx* c, size_t bytes, void** ws, void** pinned);
// Device velocity the viscous air pressures are formed from (the real speed
// of Speed_GetRef(-1, false) when an acceleration is projected); NULL: the
// projected field itself.
MAC_HIDDEN void mac_internal_set_real_velocity(mac_ctx* c, const double* u, const double* v, const double* w);

// Device view of a projection context's level-0 grid (valid until the next
// build).
struct mac_grid_view {
  hipStream_t st;       // the context's stream
  int nb;               // bricks
  int ext[3];           // bounding box in bricks (x, y, z)
  const int* dkeys;     // device: 3 chunk keys per brick, shifted to a minimum of 0 per axis
  const uint8_t* flags; // device: cell flags
  const uint8_t* code;  // device: stencil codes (bit 6: unknown)
  double* x;            // device: pressure (the PCG's starting iterate with guess = 1)
  const double* f;      // device: right-hand side of the last divergence
  const double* ff;     // device: its squared norm over the unknowns
  long long ndom;       // unknowns
  double ih2m;          // mean over the axes of 1/h^2
};
MAC_HIDDEN int mac_internal_view(mac_ctx* c, mac_grid_view* v);
// Right-hand side (and its norm) of the context's current face velocities,
// on the context's stream.
MAC_HIDDEN int mac_internal_rhs(mac_ctx* c);
// Copy device face velocities into the context.
MAC_HIDDEN int mac_internal_load(mac_ctx* c, const double* u, const double* v, const double* w);
// Learned guess from the context's current velocities and `prev` (device,
// brick layout; may alias the context's pressure or be null): writes the
// context's pressure.  Returns after the guess is on the context's stream.
MAC_HIDDEN int mac_internal_unet_guess(mac_unet* net, mac_ctx* c, const double* prev, float out_scale);
// Training pair of the context's last projection (its right-hand side and
// answer) with `prev` (device, may be null) as the previous pressure; host outputs.
MAC_HIDDEN int mac_internal_unet_sample(mac_unet* net, mac_ctx* c, const double* prev, float out_scale, float* in5, float* target,
                                        float* mask);
}
