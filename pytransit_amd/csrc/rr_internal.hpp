// Types and helpers shared by the library's translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

namespace rr {

// Per-parameter-vector state written by the setup kernel and read by the flux
// kernels: Taylor coefficients (x then y, divided by n!), the contact window
// (first/fourth contact, before the per-light-curve margin), period, 1/period,
// the radius ratio the window and the k-table lerp used, two validity flags
// (orbit parameters valid; limb-darkening profile not NaN), and the per-pv values
// of light curve 0 (its zero epoch, radius ratio and stellar disk integral), so a
// one-light-curve flux tile needs no loads beyond this record.
struct alignas(16) PvRec {
    double c[10];
    double bt1, bt4;
    double p, inv_p;
    double k0;
    int32_t good;
    int32_t ldok;
    double t0_lc0, k_lc0, ist_lc0;
    double pad_[5];
};
static_assert(sizeof(PvRec) == 192, "PvRec must be 192 bytes");

int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);
size_t align_up(size_t x, size_t a);

}  // namespace rr
