// mchip_runtime.hpp -- runtime services of mchip_api.hip shared with the
// other host-side translation units of the library (hidden symbols).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "mchip_internal.hpp"

#define MCHIP_HIDDEN __attribute__((visibility("hidden")))

namespace mchip {

// Record an error message for mchip_last_error(); returns -1.
MCHIP_HIDDEN int rt_fail(const char *fmt, ...);
// Slot of the per-quality table for q (built and uploaded on first use).
MCHIP_HIDDEN int rt_qtab_slot(float q, uint32_t *slot);
// Device array of per-quality tables (valid after the first rt_qtab_slot).
MCHIP_HIDDEN const QTab *rt_qtabs();
// Entropy decoder variant selected by mchip_set_tuning("decoder", ...).
MCHIP_HIDDEN int rt_decoder();

}  // namespace mchip

#define MCHIP_RT_CHECK(x)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return mchip::rt_fail("%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)
