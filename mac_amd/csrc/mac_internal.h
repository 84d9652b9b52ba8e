// mac_internal.h -- entry points shared between the translation units of
// libmac_amd.so (not exported).
#pragma once
#include <stdint.h>

#include "../../include/mac_amd.h"

#define MAC_HIDDEN __attribute__((visibility("hidden")))

extern "C" {

// (Re)build a projection context's brick hierarchy from host chunk keys and a
// flag array that may live on the device or the host.
MAC_HIDDEN int mac_internal_build(mac_ctx* c, int nbricks, const int32_t* keys3_host, const uint8_t* flags_any);
// Project device-resident face velocities (brick layout of the last build) in
// place and write the pressure to p (device).  guess = 1 starts from p.
MAC_HIDDEN int mac_internal_project(mac_ctx* c, double* u, double* v, double* w, double* p, int guess, mac_stats* st);
MAC_HIDDEN int mac_internal_set_error(int code, const char* msg);
}
