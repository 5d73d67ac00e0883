// nqm_limits.h -- limits shared by device and host-only code (no HIP here).
#pragma once

#include <cstdint>

namespace nqm {

// Largest n this build counts (u32 bitsets), as deffinitions.cuh:116-118.
constexpr uint32_t kMaxN = 31;

}  // namespace nqm
