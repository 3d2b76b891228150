// Structures shared by the batching-form kernels (batch.hip) and their host
// planner (batch.cpp).
#pragma once

#include <string>

#include "eq_core.h"

namespace mega {
namespace detail {

// One cell of a batch as the kernels see it (device memory).
struct BatchCell {
  EqParams eq;         // csi = W (the ZF output), slots = 1, Q = the cell's order
  const float2* a;     // channel estimates A = H^H, [G][B][U]
  float2* w;           // W, [G][B][U]
  int32_t* info;       // optional per-group status [G]
};

struct ZfClassLaunch {
  uint32_t UT;
  bool a_lds;
  size_t lds;             // dynamic LDS bytes (max over the class's cells)
  uint32_t blocks;        // total groups of the class
  const BatchCell* cells;  // device, ncells entries
  const uint32_t* prefix;  // device, ncells + 1 group prefix sums
  uint32_t ncells;
};

enum class EqKind : uint32_t { Stream = 0, Reg = 1 };

struct EqClassLaunch {
  EqKind kind;
  uint32_t MT, KS, nch;    // tile plan (nch: stream chunk count, 0 = runtime)
  uint32_t blocks;         // grid (persistent for Stream)
  const BatchCell* cells;
  uint32_t ncells;
  const uint32_t* prefix;   // Reg: device, ncells + 1 unit prefix sums
  const uint32_t* xprefix;  // Stream: [8][ncells + 1] per-XCD queue prefix sums
  const uint32_t* xbase;    // Stream: [8][ncells] first unit of each cell's part
  uint32_t* counters;       // Stream: the 8 queue counters
};

// Sets the class kernel's dynamic-LDS limit; call before launching (and
// outside stream capture).
hipError_t prepare_zf_class(const ZfClassLaunch& L);
hipError_t launch_zf_class(const ZfClassLaunch& L, hipStream_t st);
hipError_t launch_eq_class(const EqClassLaunch& L, hipStream_t st);
hipError_t launch_reset_counters(uint32_t* counters, uint32_t n, hipStream_t st);
size_t zf_class_lds(uint32_t UT, uint32_t U, uint32_t B, bool a_lds);
size_t zf_class_static_lds(uint32_t UT);

}  // namespace detail
}  // namespace mega
