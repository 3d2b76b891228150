// Host side of the batching form: plans a batch of uneven cells into kernel
// classes, picks each equalize class's tile plan and the launch schedule by
// timing the candidates on the batch's own buffers (the autotuner), and
// captures the chosen launches into one hipGraph that mega_batch_run replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/megadetect.h"
#include "batch.h"

namespace mega {
namespace detail {
namespace {

struct ZfClass {
  uint32_t UT = 0;
  bool a_lds = false;
  size_t lds = 0;
  std::vector<uint32_t> cells;  // indices into the batch
};

struct EqPlanChoice {
  EqKind kind = EqKind::Stream;
  uint32_t MT = 4, KS = 0, waves_per_cu = 12;
  std::string str() const {
    std::ostringstream o;
    if (kind == EqKind::Stream)
      o << "stream MT=" << MT << " waves/CU=" << waves_per_cu;
    else
      o << "reg MT=" << MT << " KS=" << KS;
    return o.str();
  }
  static bool parse(const std::string& s, EqPlanChoice& c) {
    if (std::sscanf(s.c_str(), "stream MT=%u waves/CU=%u", &c.MT, &c.waves_per_cu) == 2) {
      c.kind = EqKind::Stream;
      return true;
    }
    if (std::sscanf(s.c_str(), "reg MT=%u KS=%u", &c.MT, &c.KS) == 2) {
      c.kind = EqKind::Reg;
      return true;
    }
    return false;
  }
};

struct EqClass {
  uint32_t B = 0;
  std::vector<uint32_t> cells;  // longest first
  EqPlanChoice plan;
  std::string key;
};

constexpr uint32_t kCUs = 256;

// The tile plan's per-cell fields and unit count.
uint32_t eq_units(EqParams& p, const EqPlanChoice& c) {
  if (c.kind == EqKind::Stream) {
    p.nmg = (p.U + 8 * c.MT - 1) / (8 * c.MT);
    p.nwk = 1;
    p.chunks = 1;
    p.tiles_per_wg = p.col_tiles;
    return p.G * p.col_tiles * p.nmg;
  }
  p.nmg = (p.U + 8 * c.MT - 1) / (8 * c.MT);
  p.nwk = 1;
  // column tiles of a group split into chunks of about 4 tiles
  p.tiles_per_wg = std::min<uint32_t>(4, p.col_tiles);
  p.chunks = (p.col_tiles + p.tiles_per_wg - 1) / p.tiles_per_wg;
  return p.G * p.chunks * p.nmg;
}

std::vector<EqPlanChoice> eq_candidates(uint32_t B, uint32_t maxU) {
  std::vector<EqPlanChoice> v;
  if (B % 32 == 0) {
    for (uint32_t MT : {1u, 2u, 4u}) {
      if (MT > 1 && 8 * (MT / 2) >= maxU) continue;  // a smaller tile already covers all users
      for (uint32_t w : {8u, 12u, 16u}) {
        EqPlanChoice c;
        c.kind = EqKind::Stream;
        c.MT = MT;
        c.waves_per_cu = w;
        v.push_back(c);
      }
    }
  }
  for (uint32_t MT : {1u, 2u}) {
    if (MT > 1 && 8 >= maxU) continue;
    for (uint32_t KS : {8u, 16u}) {
      if (B > 4 * KS) continue;  // W must fit one wave's registers
      EqPlanChoice c;
      c.kind = EqKind::Reg;
      c.MT = MT;
      c.KS = KS;
      v.push_back(c);
    }
  }
  return v;
}

}  // namespace
}  // namespace detail
}  // namespace mega

using namespace mega::detail;

struct mega_batch {
  int device = 0;
  std::vector<mega_cell_desc> desc;
  std::vector<EqParams> eqp;  // per cell, layout fields filled
  std::vector<ZfClass> zf;
  std::vector<EqClass> eq;
  std::vector<ZfClassLaunch> zf_launch;
  std::vector<EqClassLaunch> eq_launch;
  void* dmem = nullptr;         // tables and counters
  uint32_t* counters = nullptr;  // 8 per equalize class
  uint32_t ncounters = 0;
  hipStream_t stream = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool parallel_zf = false;
  std::string plan_text;
  double tuned_ms = 0;
};

namespace {

int bfail(int code, const std::string& m) { return set_error(code, m); }

// Device tables of one planning pass: BatchCell arrays per class and the
// prefix tables, in one allocation.
struct TableBuilder {
  std::vector<uint8_t> bytes;
  size_t add(const void* p, size_t n) {
    const size_t off = (bytes.size() + 255) / 256 * 256;
    bytes.resize(off + n);
    std::memcpy(bytes.data() + off, p, n);
    return off;
  }
};

hipError_t run_launches(mega_batch* b, hipStream_t st, bool parallel) {
  hipError_t e = launch_reset_counters(b->counters, b->ncounters, st);
  if (e != hipSuccess) return e;
  // ZF classes: the widest on the side stream when parallel
  if (parallel && b->zf_launch.size() > 1) {
    if ((e = hipEventRecord(b->fork, st)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(b->side, b->fork, 0)) != hipSuccess) return e;
    for (size_t i = 0; i < b->zf_launch.size(); ++i) {
      e = launch_zf_class(b->zf_launch[i], i == 0 ? b->side : st);
      if (e != hipSuccess) return e;
    }
    if ((e = hipEventRecord(b->join, b->side)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(st, b->join, 0)) != hipSuccess) return e;
  } else {
    for (auto& L : b->zf_launch)
      if ((e = launch_zf_class(L, st)) != hipSuccess) return e;
  }
  for (auto& L : b->eq_launch)
    if ((e = launch_eq_class(L, st)) != hipSuccess) return e;
  return hipSuccess;
}

float time_ms(hipStream_t st, int reps, const std::function<hipError_t()>& f, hipError_t& err) {
  hipEvent_t a, c;
  err = hipEventCreate(&a);
  if (err != hipSuccess) return 0;
  err = hipEventCreate(&c);
  if (err != hipSuccess) return 0;
  float best = 1e30f;
  for (int r = 0; r < reps && err == hipSuccess; ++r) {
    err = hipEventRecord(a, st);
    if (err == hipSuccess) err = f();
    if (err == hipSuccess) err = hipEventRecord(c, st);
    if (err == hipSuccess) err = hipEventSynchronize(c);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, c);
    if (r > 0) best = std::min(best, ms);  // the first run warms up
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(c);
  return best;
}

std::map<std::string, std::string> read_cache(const char* path) {
  std::map<std::string, std::string> m;
  if (!path) return m;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t t = line.find('\t');
    if (t != std::string::npos) m[line.substr(0, t)] = line.substr(t + 1);
  }
  return m;
}

void append_cache(const char* path, const std::string& key, const std::string& val) {
  if (!path) return;
  std::ofstream f(path, std::ios::app);
  f << key << '\t' << val << '\n';
}

}  // namespace

extern "C" {

int mega_batch_destroy(mega_batch_t b) {
  if (!b) return MEGA_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->dmem) (void)hipFree(b->dmem);
  delete b;
  return MEGA_OK;
}

int mega_batch_create(const mega_cell_desc* cells, uint32_t num_cells, const char* tune_path,
                      mega_batch_t* out) {
  if (!out) return bfail(MEGA_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!cells || num_cells == 0) return bfail(MEGA_ERR_INVALID_ARG, "no cells");
  auto* b = new mega_batch;
  auto bail = [&](int code, const std::string& m) {
    mega_batch_destroy(b);
    return bfail(code, m);
  };
  auto hip_bail = [&](hipError_t e, const char* what) {
    return bail(MEGA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e = hipGetDevice(&b->device);
  if (e != hipSuccess) return hip_bail(e, "hipGetDevice");

  // ---- validate, per-cell parameters --------------------------------------
  b->desc.assign(cells, cells + num_cells);
  for (uint32_t i = 0; i < num_cells; ++i) {
    const mega_cell_desc& c = cells[i];
    const std::string who = "cell " + std::to_string(i) + ": ";
    if (c.mod_order != 2 && c.mod_order != 4 && c.mod_order != 6 && c.mod_order != 8)
      return bail(MEGA_ERR_UNSUPPORTED, who + "unsupported modulation order");
    if (c.num_ue == 0 || c.num_bs == 0 || c.num_sc == 0 || c.sc_group == 0 || c.num_syms == 0 ||
        c.cb_size == 0)
      return bail(MEGA_ERR_INVALID_ARG, who + "zero size");
    if (c.num_ue > kMaxUsers) return bail(MEGA_ERR_UNSUPPORTED, who + "num_ue > 64");
    if (!c.csi || !c.y || !c.llr || !c.w) return bail(MEGA_ERR_INVALID_ARG, who + "NULL buffer");
    EqualizeArgs a{};
    a.csi = static_cast<const float2*>(c.w);
    a.y = static_cast<const float2*>(c.y);
    a.llr = static_cast<_Float16*>(c.llr);
    a.xhat = static_cast<float2*>(c.xhat);
    a.U = c.num_ue;
    a.B = c.num_bs;
    a.N = c.num_sc;
    a.sg = c.sc_group;
    a.G = (c.num_sc + c.sc_group - 1) / c.sc_group;
    a.S = c.num_syms;
    a.slots = 1;
    a.Q = c.mod_order;
    a.ue_stride = c.ue_stride;
    a.cb_size = c.cb_size;
    a.cb_stride = c.cb_stride;
    a.zeros = c.zeros;
    a.valid_bits = c.valid_bits;
    a.y_sym = c.y_sym_stride;
    a.llr_sym = c.llr_sym_stride;
    EqParams p;
    const char* why = nullptr;
    if (fill_eq_params(a, p, &why) != hipSuccess) return bail(MEGA_ERR_UNSUPPORTED, who + why);
    b->eqp.push_back(p);
  }

  // ---- classes --------------------------------------------------------------
  for (uint32_t UT : {1u, 2u, 4u}) {
    ZfClass z;
    z.UT = UT;
    for (uint32_t i = 0; i < num_cells; ++i) {
      const uint32_t u = b->eqp[i].U;
      if ((UT == 1 && u <= 16) || (UT == 2 && u > 16 && u <= 32) || (UT == 4 && u > 32))
        z.cells.push_back(i);
    }
    if (z.cells.empty()) continue;
    z.a_lds = UT == 1;
    for (uint32_t i : z.cells)
      if (zf_class_static_lds(UT) + zf_class_lds(UT, b->eqp[i].U, b->eqp[i].B, true) >
          120 * 1024)
        z.a_lds = false;
    for (uint32_t i : z.cells)
      z.lds = std::max(z.lds, zf_class_lds(UT, b->eqp[i].U, b->eqp[i].B, z.a_lds));
    b->zf.push_back(z);
  }
  {
    std::map<uint32_t, EqClass> byB;
    for (uint32_t i = 0; i < num_cells; ++i) {
      EqClass& c = byB[b->eqp[i].B];
      c.B = b->eqp[i].B;
      c.cells.push_back(i);
    }
    for (auto& kv : byB) {
      EqClass c = kv.second;
      std::sort(c.cells.begin(), c.cells.end(), [&](uint32_t x, uint32_t y) {
        const EqParams &px = b->eqp[x], &py = b->eqp[y];
        return (uint64_t)px.G * px.col_tiles * px.U > (uint64_t)py.G * py.col_tiles * py.U;
      });
      std::ostringstream k;
      k << "B=" << c.B;
      for (uint32_t i : c.cells)
        k << ";" << b->eqp[i].U << "x" << b->eqp[i].N << "/" << b->eqp[i].sg << "x"
          << b->eqp[i].S;
      c.key = k.str();
      b->eq.push_back(c);
    }
  }

  // ---- streams, events, counters ---------------------------------------------
  if ((e = hipStreamCreate(&b->stream)) != hipSuccess) return hip_bail(e, "hipStreamCreate");
  if ((e = hipStreamCreate(&b->side)) != hipSuccess) return hip_bail(e, "hipStreamCreate");
  if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming)) != hipSuccess)
    return hip_bail(e, "hipEventCreate");

  // Builds the device tables and launch records for the current plans.
  auto build = [&]() -> hipError_t {
    TableBuilder tb;
    std::vector<size_t> zf_cells_off, zf_prefix_off, eq_cells_off, eq_prefix_off, eq_x_off,
        eq_xb_off;
    for (auto& z : b->zf) {
      std::vector<BatchCell> bc;
      std::vector<uint32_t> prefix{0};
      for (uint32_t i : z.cells) {
        BatchCell c{};
        c.eq = b->eqp[i];
        c.a = static_cast<const float2*>(b->desc[i].csi);
        c.w = static_cast<float2*>(b->desc[i].w);
        c.info = b->desc[i].info_dev;
        bc.push_back(c);
        prefix.push_back(prefix.back() + b->eqp[i].G);
      }
      zf_cells_off.push_back(tb.add(bc.data(), bc.size() * sizeof(BatchCell)));
      zf_prefix_off.push_back(tb.add(prefix.data(), prefix.size() * sizeof(uint32_t)));
    }
    for (auto& q : b->eq) {
      std::vector<BatchCell> bc;
      std::vector<uint32_t> prefix{0}, units;
      for (uint32_t i : q.cells) {
        BatchCell c{};
        c.eq = b->eqp[i];
        units.push_back(eq_units(c.eq, q.plan));
        c.a = static_cast<const float2*>(b->desc[i].csi);
        c.w = static_cast<float2*>(b->desc[i].w);
        c.info = b->desc[i].info_dev;
        bc.push_back(c);
        prefix.push_back(prefix.back() + units.back());
      }
      const uint32_t n = (uint32_t)q.cells.size();
      std::vector<uint32_t> xp(8 * (n + 1)), xb(8 * n);
      for (uint32_t x = 0; x < 8; ++x) {
        xp[x * (n + 1)] = 0;
        for (uint32_t c = 0; c < n; ++c) {
          const uint32_t lo = (uint32_t)((uint64_t)units[c] * x / 8);
          const uint32_t hi = (uint32_t)((uint64_t)units[c] * (x + 1) / 8);
          xb[x * n + c] = lo;
          xp[x * (n + 1) + c + 1] = xp[x * (n + 1) + c] + (hi - lo);
        }
      }
      eq_cells_off.push_back(tb.add(bc.data(), bc.size() * sizeof(BatchCell)));
      eq_prefix_off.push_back(tb.add(prefix.data(), prefix.size() * sizeof(uint32_t)));
      eq_x_off.push_back(tb.add(xp.data(), xp.size() * sizeof(uint32_t)));
      eq_xb_off.push_back(tb.add(xb.data(), xb.size() * sizeof(uint32_t)));
    }
    b->ncounters = 8 * (uint32_t)b->eq.size();
    std::vector<uint32_t> zeros(b->ncounters, 0);
    const size_t cnt_off = tb.add(zeros.data(), zeros.size() * sizeof(uint32_t));
    hipError_t err = hipSuccess;
    if (b->dmem) {
      err = hipStreamSynchronize(b->stream);
      if (err == hipSuccess) err = hipFree(b->dmem);
      b->dmem = nullptr;
      if (err != hipSuccess) return err;
    }
    err = hipMalloc(&b->dmem, tb.bytes.size());
    if (err != hipSuccess) return err;
    err = hipMemcpy(b->dmem, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    auto at = [&](size_t off) { return static_cast<uint8_t*>(b->dmem) + off; };
    b->counters = reinterpret_cast<uint32_t*>(at(cnt_off));
    b->zf_launch.clear();
    for (size_t k = 0; k < b->zf.size(); ++k) {
      ZfClassLaunch L{};
      L.UT = b->zf[k].UT;
      L.a_lds = b->zf[k].a_lds;
      L.lds = b->zf[k].lds;
      L.cells = reinterpret_cast<const BatchCell*>(at(zf_cells_off[k]));
      L.prefix = reinterpret_cast<const uint32_t*>(at(zf_prefix_off[k]));
      L.ncells = (uint32_t)b->zf[k].cells.size();
      L.blocks = 0;
      for (uint32_t i : b->zf[k].cells) L.blocks += b->eqp[i].G;
      err = prepare_zf_class(L);  // LDS attribute, outside any capture
      if (err != hipSuccess) return err;
      b->zf_launch.push_back(L);
    }
    // widest class first (the parallel schedule puts it on the side stream)
    std::sort(b->zf_launch.begin(), b->zf_launch.end(),
              [](const ZfClassLaunch& x, const ZfClassLaunch& y) { return x.UT > y.UT; });
    b->eq_launch.clear();
    for (size_t k = 0; k < b->eq.size(); ++k) {
      const EqClass& q = b->eq[k];
      EqClassLaunch L{};
      L.kind = q.plan.kind;
      L.MT = q.plan.MT;
      L.KS = q.plan.KS;
      L.nch = (q.B == 64 || q.B == 128 || q.B == 256) ? q.B / 8 : 0;
      L.cells = reinterpret_cast<const BatchCell*>(at(eq_cells_off[k]));
      L.ncells = (uint32_t)q.cells.size();
      L.prefix = reinterpret_cast<const uint32_t*>(at(eq_prefix_off[k]));
      L.xprefix = reinterpret_cast<const uint32_t*>(at(eq_x_off[k]));
      L.xbase = reinterpret_cast<const uint32_t*>(at(eq_xb_off[k]));
      L.counters = b->counters + 8 * k;
      uint32_t total = 0;
      for (uint32_t i : q.cells) {
        EqParams p = b->eqp[i];
        total += eq_units(p, q.plan);
      }
      const uint32_t wgs = (total + 3) / 4;
      L.blocks = q.plan.kind == EqKind::Stream
                     ? std::min<uint32_t>(wgs, kCUs * q.plan.waves_per_cu / 4)
                     : wgs;
      b->eq_launch.push_back(L);
    }
    return hipSuccess;
  };

  // ---- autotune the equalize classes ---------------------------------------
  // Every class starts on its first candidate (or the cached choice); one run
  // with those computes W, then each class without a cached choice times its
  // candidates alone on the batch's own buffers and keeps the fastest.
  auto cache = read_cache(tune_path);
  std::vector<bool> tune(b->eq.size(), true);
  for (size_t k = 0; k < b->eq.size(); ++k) {
    EqClass& q = b->eq[k];
    uint32_t maxU = 0;
    for (uint32_t i : q.cells) maxU = std::max(maxU, b->eqp[i].U);
    const auto cands = eq_candidates(q.B, maxU);
    if (cands.empty())
      return bail(MEGA_ERR_UNSUPPORTED, "no equalize plan for B=" + std::to_string(q.B));
    q.plan = cands[0];
    auto it = cache.find(q.key);
    if (it != cache.end() && EqPlanChoice::parse(it->second, q.plan)) tune[k] = false;
  }
  if ((e = build()) != hipSuccess) return hip_bail(e, "batch tables");
  if ((e = run_launches(b, b->stream, false)) != hipSuccess) return hip_bail(e, "batch warm-up");
  if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return hip_bail(e, "batch warm-up");
  std::ostringstream tune_log;
  for (size_t k = 0; k < b->eq.size(); ++k) {
    EqClass& q = b->eq[k];
    if (!tune[k]) {
      tune_log << "  " << q.key << ": " << q.plan.str() << " (cached)\n";
      continue;
    }
    uint32_t maxU = 0;
    for (uint32_t i : q.cells) maxU = std::max(maxU, b->eqp[i].U);
    float best = 1e30f;
    EqPlanChoice best_c = q.plan;
    for (const auto& c : eq_candidates(q.B, maxU)) {
      q.plan = c;
      if ((e = build()) != hipSuccess) return hip_bail(e, "batch tables");
      const EqClassLaunch L = b->eq_launch[k];
      const float ms = time_ms(b->stream, 4, [&]() {
        hipError_t r = launch_reset_counters(L.counters, 8, b->stream);
        return r == hipSuccess ? launch_eq_class(L, b->stream) : r;
      }, e);
      if (e != hipSuccess) return hip_bail(e, "batch tuning");
      if (ms < best) {
        best = ms;
        best_c = c;
      }
    }
    q.plan = best_c;
    tune_log << "  " << q.key << ": " << q.plan.str() << " (" << best * 1e3 << " us)\n";
    append_cache(tune_path, q.key, q.plan.str());
  }
  if ((e = build()) != hipSuccess) return hip_bail(e, "batch tables");

  // ---- schedule: ZF classes serial or two-way parallel; capture -----------
  auto capture = [&](bool parallel) -> hipError_t {
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    if (b->graph) {
      (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
    }
    hipError_t r = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (r != hipSuccess) return r;
    hipError_t r2 = run_launches(b, b->stream, parallel);
    r = hipStreamEndCapture(b->stream, &b->graph);
    if (r2 != hipSuccess) return r2;
    if (r != hipSuccess) return r;
    return hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  };
  float best_ms = 1e30f;
  bool best_par = false;
  for (bool par : {false, true}) {
    if (par && b->zf_launch.size() < 2) continue;
    if ((e = capture(par)) != hipSuccess) return hip_bail(e, "graph capture");
    const float ms = time_ms(b->stream, 4, [&]() { return hipGraphLaunch(b->exec, b->stream); }, e);
    if (e != hipSuccess) return hip_bail(e, "graph timing");
    if (ms < best_ms) {
      best_ms = ms;
      best_par = par;
    }
  }
  if ((e = capture(best_par)) != hipSuccess) return hip_bail(e, "graph capture");
  b->parallel_zf = best_par;
  b->tuned_ms = best_ms;

  std::ostringstream plan;
  plan << num_cells << " cells; ZF classes:";
  for (auto& z : b->zf)
    plan << " UT=" << z.UT << "(" << z.cells.size() << " cells" << (z.a_lds ? ", A in LDS" : "")
         << ")";
  plan << (best_par ? " [two streams]" : " [serial]") << "\nequalize classes:\n" << tune_log.str();
  plan << "graph: " << best_ms * 1e3 << " us per launch\n";
  b->plan_text = plan.str();
  *out = b;
  return MEGA_OK;
}

int mega_batch_run(mega_batch_t b, mega_stream_t stream) {
  if (!b) return bfail(MEGA_ERR_INVALID_ARG, "batch handle is NULL");
  hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return bfail(MEGA_ERR_HIP, std::string("mega_batch_run: ") + hipGetErrorString(e));
  return MEGA_OK;
}

const char* mega_batch_plan(mega_batch_t b) { return b ? b->plan_text.c_str() : ""; }

}  // extern "C"
