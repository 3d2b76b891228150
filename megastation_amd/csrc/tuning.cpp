// Tuning overrides (MEGA_* environment variables) read once, not per launch.
//
// The launchers consult these on every call, so the environment is parsed the
// first time any of them is needed and again only on mega_tuning_reload().
// Every value is sanitized here: anything outside a key's accepted set reads
// as that key's default (0 = the launcher's own choice), so no launcher sees
// an out-of-range index.
#include <atomic>
#include <cstdlib>
#include <mutex>

#include "internal.h"

namespace mega {
namespace detail {

namespace {

struct TuneSpec {
  const char* env;
  int def;
  int lo, hi;  // accepted range; outside it the default applies
};

// Indexed by TuneKey.
constexpr TuneSpec kSpecs[kTuneCount] = {
    {"MEGA_EQ_KS", 0, 8, 16},     // FP32 equalize, register plan: antenna steps per wave
    {"MEGA_EQ_MODE", 0, 0, 3},    // FP32 equalize: 1 streamed W, 2 W in registers, 3 wide
    {"MEGA_EQ_SCH", 0, 0, 2},     // FP32 streamed W: chunk shape
    {"MEGA_EQ_SMT", 0, 1, 4},     // FP32 streamed W: users per wave / 8
    {"MEGA_EQ_TPW", 0, 1, 1 << 20},  // FP32 register plan: column tiles per workgroup
    {"MEGA_EQ_XMT", 0, 1, 2},     // mixed equalize: 16-user tiles per wave
    {"MEGA_EQ_XREG", 0, 0, 1},    // mixed equalize: 1 forces streamed W
    {"MEGA_EQ_XT", 0, 1, 2},      // mixed equalize: column tiles per wave
    {"MEGA_EQ_XKS", 0, 1, 4},     // mixed equalize: waves per unit (antenna split)
    {"MEGA_PREC_ACC", 0, 8, 8},   // precode: 8 = 64 antennas per wave
    {"MEGA_DET_PF", 1, 0, 1},     // fused detect: next-tile y prefetch
    {"MEGA_EQ_WTC", 0, 1, 1 << 20},  // FP32 wide form: column tiles per workgroup
    {"MEGA_ZF_NTH", 0, 256, 512},   // FP32 ZF threads per group: 512 (17-32 users), 256 (33-64)
    {"MEGA_ZF_WIDE", 0, 1, 2},      // FP32 ZF: 1 one workgroup per group, 2 three kernels
    {"MEGA_UPLINK_STREAMS", 0, 1, 2},  // uplink data path: 1 one stream, 2 two streams
    {"MEGA_EQ_WNW", 0, 4, 8},       // FP32 wide form: waves per workgroup (4 or 8)
    {"MEGA_EQ_WP", 0, 1, 8},        // FP32 wide form: persistent, k * 256 workgroups
};

std::atomic<int> g_vals[kTuneCount];
std::once_flag g_once;
std::mutex g_mu;

void read_env() {
  for (int k = 0; k < kTuneCount; ++k) {
    const TuneSpec& s = kSpecs[k];
    const char* e = std::getenv(s.env);
    int v = s.def;
    if (e && *e) {
      char* end = nullptr;
      const long x = std::strtol(e, &end, 10);
      if (end != e && x >= s.lo && x <= s.hi) v = (int)x;
    }
    g_vals[k].store(v, std::memory_order_relaxed);
  }
}

}  // namespace

int tune(TuneKey k) {
  std::call_once(g_once, read_env);
  return g_vals[k].load(std::memory_order_relaxed);
}

void tuning_reload() {
  std::call_once(g_once, read_env);
  std::lock_guard<std::mutex> lk(g_mu);
  read_env();
}

}  // namespace detail
}  // namespace mega

extern "C" int mega_tuning_reload(void) {
  mega::detail::tuning_reload();
  return MEGA_OK;
}
