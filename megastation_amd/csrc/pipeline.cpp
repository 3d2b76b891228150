// Slot pipeline of one cell (mega_pipeline_*): the ZF of slot i+1 runs on the
// pipeline's own stream beside the equalize + demod of slot i on the caller's
// stream. The reference runs the beamformer of a frame, then the equalize
// calls of its symbols, frame after frame (src/baseband/process.cc:172-185,
// src/scheduler/gpu_engine.cc:289-343); here consecutive slots overlap.
//
// push(slot i) issues ZF(i) first, then equalize(i-1): issued the other way
// round, the ZF would wait for the equalize waves to drain before any of its
// workgroups found room. Weights live in `depth` buffers owned by the
// pipeline; ZF(i) writes buffer i % depth after equalize(i - depth) has read
// it. csi is read on the ZF stream: its producer runs there (or that stream
// is made to wait for it), so the ZF stream runs ahead of the equalize stream
// without a cross-stream wait on the critical path. Events are created once,
// without timing.
//
// mega_pipeline_graph_*: n consecutive slots with their buffers bound,
// scheduled exactly as n pushes and a flush, captured once into a graph and
// replayed with one launch; the host then issues nothing per slot, and the
// two streams' kernels follow each other without host gaps.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/megadetect.h"
#include "internal.h"
#include "prec_core.h"

struct mega_pipeline {
  mega::detail::EqualizeArgs a;  // one slot; buffers filled per push
  int form;
  uint32_t depth;
  hipStream_t st;  // caller's stream (equalize)
  hipStream_t zs;  // ZF stream
  std::vector<void*> w;
  void* zf_work = nullptr;  // the three-kernel ZF's workspace (FP32 form, few wide groups)
  uint64_t zf_work_bytes = 0;
  std::vector<hipEvent_t> zf_ev, eq_ev;
  std::vector<bool> eq_rec;
  uint64_t pushed = 0;
  bool pending = false;
  const void* py = nullptr;
  void* pllr = nullptr;
  void* pxhat = nullptr;
};

namespace {

int hip_err(hipError_t e, const char* who, const char* why = nullptr) {
  std::string m = std::string(who) + ": " + hipGetErrorString(e);
  if (why) m += std::string(" (") + why + ")";
  return mega::detail::set_error(MEGA_ERR_HIP, m);
}

void release(mega_pipeline* p) {
  for (void* b : p->w)
    if (b) (void)hipFree(b);
  if (p->zf_work) (void)hipFree(p->zf_work);
  for (hipEvent_t e : p->zf_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->eq_ev)
    if (e) (void)hipEventDestroy(e);
  if (p->zs) (void)hipStreamDestroy(p->zs);
  delete p;
}

int issue_equalize(mega_pipeline* p, uint32_t k) {
  hipError_t e = hipStreamWaitEvent(p->st, p->zf_ev[k], 0);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_push");
  mega::detail::EqualizeArgs a = p->a;
  a.csi = static_cast<const float2*>(p->w[k]);
  a.y = static_cast<const float2*>(p->py);
  a.llr = static_cast<_Float16*>(p->pllr);
  a.xhat = static_cast<float2*>(p->pxhat);
  const char* why = nullptr;
  e = p->form == MEGA_FORM_MIXED ? mega::detail::launch_equalize_mixed(a, p->w[k], p->st, &why)
                                 : mega::detail::launch_equalize(a, p->st, &why);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_push", why);
  e = hipEventRecord(p->eq_ev[k], p->st);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_push");
  p->eq_rec[k] = true;
  p->pending = false;
  return MEGA_OK;
}

}  // namespace

extern "C" {

int mega_pipeline_create(const mega_equalize_desc* shape, uint32_t depth, mega_stream_t stream,
                         mega_pipeline_t* out) {
  if (!out) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, "mega_pipeline_create: out is NULL");
  *out = nullptr;
  if (depth < 2)
    return mega::detail::set_error(MEGA_ERR_INVALID_ARG,
                                   "mega_pipeline_create: depth must be at least 2");
  mega::detail::EqualizeArgs a;
  bool empty = false;
  int r = mega::detail::equalize_args_checked(shape, a, empty, "mega_pipeline_create", true,
                                              false);
  if (r != MEGA_OK) return r;
  if (empty)
    return mega::detail::set_error(MEGA_ERR_INVALID_ARG, "mega_pipeline_create: empty cell");
  // a.slots = shape->num_slots: slots per push (their ZF is one launch over
  // slots * G groups, their equalize one launch)
  const uint64_t Bp = mega::detail::split_row(a.B);
  a.csi_slot = shape->form == MEGA_FORM_MIXED ? (uint64_t)a.G * a.U * Bp
                                              : (uint64_t)a.G * a.B * a.U;
  if (a.y_slot == 0) a.y_slot = (uint64_t)a.S * a.y_sym;
  if (a.llr_slot == 0) a.llr_slot = (uint64_t)a.S * a.llr_sym;
  if ((uint64_t)a.slots * a.G > 0xffffffffull)
    return mega::detail::set_error(MEGA_ERR_UNSUPPORTED, "mega_pipeline_create: too many groups");
  mega_pipeline* p = new mega_pipeline();
  p->a = a;
  p->form = shape->form;
  p->depth = depth;
  p->st = static_cast<hipStream_t>(stream);
  p->w.assign(depth, nullptr);
  p->zf_ev.assign(depth, nullptr);
  p->eq_ev.assign(depth, nullptr);
  p->eq_rec.assign(depth, false);
  hipError_t e = hipStreamCreateWithFlags(&p->zs, hipStreamNonBlocking);
  const size_t wbytes = a.csi_slot * 8 * a.slots;  // complex64 or a split word pair
  for (uint32_t k = 0; k < depth && e == hipSuccess; ++k) {
    e = hipMalloc(&p->w[k], wbytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->zf_ev[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->eq_ev[k], hipEventDisableTiming);
  }
  // one ZF workspace: the ZF launches follow each other on the ZF stream
  const uint64_t zb = mega::detail::zf_wide_work_bytes(a.U, a.B, a.slots * a.G);
  if (e == hipSuccess && p->form == MEGA_FORM_FP32 && zb > 0 &&
      mega::detail::zf_wide_pays(a.U, a.B, a.slots * a.G)) {
    e = hipMalloc(&p->zf_work, zb);
    if (e == hipSuccess) p->zf_work_bytes = zb;
  }
  if (e != hipSuccess) {
    release(p);
    return hip_err(e, "mega_pipeline_create");
  }
  *out = p;
  return MEGA_OK;
}

int mega_pipeline_push(mega_pipeline_t p, const void* csi, const void* y, void* llr,
                       void* xhat) {
  if (!p) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, "mega_pipeline_push: NULL handle");
  if (!csi || !y || !llr)
    return mega::detail::set_error(MEGA_ERR_INVALID_ARG, "mega_pipeline_push: NULL buffer");
  const uint32_t k = (uint32_t)(p->pushed % p->depth);
  // buffer k is written after the equalize that last read it
  hipError_t e = hipSuccess;
  if (p->eq_rec[k]) e = hipStreamWaitEvent(p->zs, p->eq_ev[k], 0);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_push");
  const char* why = nullptr;
  const float2* in = static_cast<const float2*>(csi);
  e = p->form == MEGA_FORM_MIXED
          ? mega::detail::launch_zf_mixed(p->a.U, p->a.B, p->a.slots * p->a.G, in, nullptr,
                                          p->w[k], nullptr, nullptr, p->zs, &why)
          : mega::detail::launch_zf(p->a.U, p->a.B, p->a.slots * p->a.G, in,
                                    static_cast<float2*>(p->w[k]), nullptr, p->zs, &why,
                                    p->zf_work, p->zf_work_bytes, false);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_push", why);
  e = hipEventRecord(p->zf_ev[k], p->zs);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_push");
  if (p->pending) {
    const int r = issue_equalize(p, (uint32_t)((p->pushed - 1) % p->depth));
    if (r != MEGA_OK) return r;
  }
  p->py = y;
  p->pllr = llr;
  p->pxhat = xhat;
  p->pending = true;
  ++p->pushed;
  return MEGA_OK;
}

int mega_pipeline_flush(mega_pipeline_t p) {
  if (!p) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, "mega_pipeline_flush: NULL handle");
  if (!p->pending) return MEGA_OK;
  return issue_equalize(p, (uint32_t)((p->pushed - 1) % p->depth));
}

mega_stream_t mega_pipeline_zf_stream(mega_pipeline_t p) { return p ? p->zs : nullptr; }

int mega_pipeline_destroy(mega_pipeline_t p) {
  if (!p) return MEGA_OK;
  hipError_t e = hipStreamSynchronize(p->zs);
  for (uint32_t k = 0; k < p->depth && e == hipSuccess; ++k)
    if (p->eq_rec[k]) e = hipEventSynchronize(p->eq_ev[k]);
  release(p);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_destroy");
  return MEGA_OK;
}

struct mega_pipeline_graph {
  mega_pipeline* p = nullptr;  // weight buffers and events the nodes use
  hipStream_t cs = nullptr;    // capture stream
  hipEvent_t fork = nullptr, join = nullptr, done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool launched = false;
};

static void release_graph(mega_pipeline_graph* g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->p) {
    // its event records were captured, not run: nothing to wait for
    g->p->eq_rec.assign(g->p->depth, false);
    (void)mega_pipeline_destroy(g->p);
  }
  for (hipEvent_t e : {g->fork, g->join, g->done})
    if (e) (void)hipEventDestroy(e);
  if (g->cs) (void)hipStreamDestroy(g->cs);
  delete g;
}

int mega_pipeline_graph_create(const mega_equalize_desc* shape, uint32_t n,
                               const void* const* csi, const void* const* y,
                               void* const* llr, void* const* xhat,
                               mega_pipeline_graph_t* out) {
  const char* who = "mega_pipeline_graph_create";
  if (!out) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": out is NULL");
  *out = nullptr;
  if (n == 0) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": no slots");
  if (!csi || !y || !llr)
    return mega::detail::set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL buffer array");
  for (uint32_t i = 0; i < n; ++i)
    if (!csi[i] || !y[i] || !llr[i])
      return mega::detail::set_error(MEGA_ERR_INVALID_ARG,
                                     std::string(who) + ": NULL buffer of slot " + std::to_string(i));
  mega_pipeline_graph* g = new mega_pipeline_graph();
  hipError_t e = hipStreamCreateWithFlags(&g->cs, hipStreamNonBlocking);
  for (hipEvent_t* ev : {&g->fork, &g->join, &g->done})
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    release_graph(g);
    return hip_err(e, who);
  }
  int r = mega_pipeline_create(shape, 3, g->cs, &g->p);
  if (r != MEGA_OK) {
    release_graph(g);
    return r;
  }
  e = hipStreamBeginCapture(g->cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    release_graph(g);
    return hip_err(e, who);
  }
  e = hipEventRecord(g->fork, g->cs);
  if (e == hipSuccess) e = hipStreamWaitEvent(g->p->zs, g->fork, 0);
  r = e == hipSuccess ? MEGA_OK : hip_err(e, who);
  for (uint32_t i = 0; i < n && r == MEGA_OK; ++i)
    r = mega_pipeline_push(g->p, csi[i], y[i], llr[i], xhat ? xhat[i] : nullptr);
  if (r == MEGA_OK) r = mega_pipeline_flush(g->p);
  if (r == MEGA_OK) {
    e = hipEventRecord(g->join, g->p->zs);
    if (e == hipSuccess) e = hipStreamWaitEvent(g->cs, g->join, 0);
    if (e != hipSuccess) r = hip_err(e, who);
  }
  // the capture ends on every path (a stream left capturing is unusable)
  const hipError_t ee = hipStreamEndCapture(g->cs, &g->graph);
  if (r == MEGA_OK && ee != hipSuccess) r = hip_err(ee, who);
  if (r == MEGA_OK) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) r = hip_err(e, who);
  }
  if (r != MEGA_OK) {
    release_graph(g);
    return r;
  }
  *out = g;
  return MEGA_OK;
}

int mega_pipeline_graph_launch(mega_pipeline_graph_t g, mega_stream_t stream) {
  if (!g) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, "mega_pipeline_graph_launch: NULL handle");
  hipError_t e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
  if (e == hipSuccess) e = hipEventRecord(g->done, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_graph_launch");
  g->launched = true;
  return MEGA_OK;
}

int mega_pipeline_graph_destroy(mega_pipeline_graph_t g) {
  if (!g) return MEGA_OK;
  hipError_t e = g->launched ? hipEventSynchronize(g->done) : hipSuccess;
  release_graph(g);
  if (e != hipSuccess) return hip_err(e, "mega_pipeline_graph_destroy");
  return MEGA_OK;
}

}  // extern "C"
