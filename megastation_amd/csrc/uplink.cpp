// Uplink data path of one slot from received samples to LLRs
// (mega_uplink_*): the uplink FFT stage (fft.hip) writing y into buffers the
// handle owns, which the equalize + demod (equalize*.hip) reads back.
//
// The reference runs, per uplink symbol on that symbol's stream,
// UplinkFFT -> Equalize::equalize -> LDPC decode, with a y buffer per symbol
// (src/baseband/process.cc:175-186). Here, by default, the slot's S symbols
// go through one FFT launch and one equalize launch on the caller's stream,
// through one slot of y that the handle reuses every run (87 MB at 256 x 32:
// it stays in the 256 MB Infinity Cache from the FFT's stores to the
// equalize's loads). With chunk_syms = K < S the symbols go in chunks of K
// through a ring of two chunk buffers: the FFT of chunk c writes buffer
// c % 2 on the caller's stream, the equalize of chunk c reads it on the
// handle's stream, the FFT of chunk c + 2 waits for the equalize of chunk c
// (events), so the FFT of one chunk runs beside the equalize of the previous
// and y needs only 2 K B N complex. Measured at 256 x 32 (FFT 81 us, equalize
// 48 us alone): one launch each 118.5 us, the staged calls with their own y
// 124.7; chunks of 5 symbols 185 us on two streams (166 graph-captured) and
// 146 us on one; chunks of 7 164 / 145 / 134. Both kernels are bound by
// latency with every CU busy, so running them side by side gains nothing,
// while smaller launches lose efficiency and each cross-stream wait costs
// several microseconds. MEGA_UPLINK_STREAMS=1 keeps chunked runs on one
// stream. The kernels are the standalone ones and every column's arithmetic
// is the same in any chunk, so the LLRs equal mega_uplink_fft of the slot
// followed by mega_equalize_slots bit for bit (FP32 form).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/megadetect.h"
#include "internal.h"

struct mega_uplink {
  mega::detail::FftArgs f;        // one slot's FFT (in / out filled per chunk)
  mega::detail::EqualizeArgs a;   // one slot's equalize (buffers filled per run)
  int form;
  uint32_t K;                     // symbols per chunk
  uint32_t chunks;
  hipStream_t es = nullptr;       // equalize stream
  void* ring[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> fft_ev, eq_ev;
  hipEvent_t fork = nullptr, join = nullptr;
};

namespace {

int hip_err(hipError_t e, const char* who, const char* why = nullptr) {
  std::string m = std::string(who) + ": " + hipGetErrorString(e);
  if (why) m += std::string(" (") + why + ")";
  return mega::detail::set_error(MEGA_ERR_HIP, m);
}

void release(mega_uplink* u) {
  for (void* b : u->ring)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : u->fft_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : u->eq_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {u->fork, u->join})
    if (e) (void)hipEventDestroy(e);
  if (u->es) (void)hipStreamDestroy(u->es);
  delete u;
}

}  // namespace

extern "C" {

int mega_uplink_create(const mega_fft_desc* fft, const mega_equalize_desc* shape,
                       uint32_t chunk_syms, mega_uplink_t* out) {
  const char* who = "mega_uplink_create";
  using mega::detail::set_error;
  if (!out) return set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": out is NULL");
  *out = nullptr;
  if (!fft || !shape) return set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL desc");
  mega::detail::EqualizeArgs a;
  bool empty = false;
  int r = mega::detail::equalize_args_checked(shape, a, empty, who, true, false);
  if (r != MEGA_OK) return r;
  if (empty) return set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": empty cell");
  if (a.slots != 1)
    return set_error(MEGA_ERR_UNSUPPORTED, std::string(who) + ": one slot per run");
  if (fft->bs_num != a.B || fft->ofdm_num != a.N || fft->num_syms != a.S)
    return set_error(MEGA_ERR_INVALID_ARG,
                     std::string(who) + ": the FFT's bs_num, ofdm_num and num_syms must be "
                                        "the equalize's num_bs, num_sc and num_syms");
  if (fft->quant_bits != 16 && fft->quant_bits != 8)
    return set_error(MEGA_ERR_UNSUPPORTED, std::string(who) + ": quant_bits must be 16 or 8");
  if (fft->ofdm_ca < 16 || fft->ofdm_ca > 4096 || (fft->ofdm_ca & (fft->ofdm_ca - 1)) ||
      fft->ofdm_num > fft->ofdm_ca)
    return set_error(MEGA_ERR_UNSUPPORTED,
                     std::string(who) + ": ofdm_ca must be a power of two from 16 to 4096, "
                                        "ofdm_num at most ofdm_ca");
  const uint32_t K = chunk_syms == 0 || chunk_syms > a.S ? a.S : chunk_syms;
  mega_uplink* u = new mega_uplink();
  u->a = a;
  u->a.y_sym = (uint64_t)a.B * a.N;  // the ring's layout
  if (a.llr_sym == 0) u->a.llr_sym = (uint64_t)a.U * a.ue_stride;
  u->form = shape->form;
  u->K = K;
  u->chunks = (a.S + K - 1) / K;
  mega::detail::FftArgs& f = u->f;
  f = mega::detail::FftArgs{};
  f.B = fft->bs_num;
  f.ca = fft->ofdm_ca;
  f.num = fft->ofdm_num;
  f.start = fft->ofdm_start;
  f.qbits = (uint32_t)fft->quant_bits;
  f.in_sym = fft->in_sym_stride ? fft->in_sym_stride : (uint64_t)fft->bs_num * fft->ofdm_ca;
  f.out_sym = (uint64_t)a.B * a.N;
  hipError_t e = hipStreamCreateWithFlags(&u->es, hipStreamNonBlocking);
  const size_t rbytes = (size_t)K * a.B * a.N * sizeof(float2);
  for (uint32_t k = 0; k < (u->chunks > 1 ? 2u : 1u) && e == hipSuccess; ++k)
    e = hipMalloc(&u->ring[k], rbytes);
  u->fft_ev.assign(u->chunks, nullptr);
  u->eq_ev.assign(u->chunks, nullptr);
  for (uint32_t c = 0; c < u->chunks && e == hipSuccess; ++c) {
    e = hipEventCreateWithFlags(&u->fft_ev[c], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&u->eq_ev[c], hipEventDisableTiming);
  }
  for (hipEvent_t* ev : {&u->fork, &u->join})
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    release(u);
    return hip_err(e, who);
  }
  *out = u;
  return MEGA_OK;
}

int mega_uplink_run(mega_uplink_t u, const void* samples, const void* w, void* llr,
                    void* xhat, mega_stream_t stream) {
  const char* who = "mega_uplink_run";
  if (!u) return mega::detail::set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL handle");
  if (!samples || !w || !llr)
    return mega::detail::set_error(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // one chunk, or MEGA_UPLINK_STREAMS=1: every launch on `stream` in chunk
  // order, no events
  const bool one = u->chunks == 1 || mega::detail::tune(mega::detail::kUplinkStreams) == 1;
  hipStream_t es = one ? st : u->es;
  hipError_t e = hipSuccess;
  // the equalize stream starts after the caller's earlier work (W, inputs)
  if (!one) {
    e = hipEventRecord(u->fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(u->es, u->fork, 0);
  }
  if (e != hipSuccess) return hip_err(e, who);
  const mega::detail::EqualizeArgs& a0 = u->a;
  for (uint32_t c = 0; c < u->chunks; ++c) {
    const uint32_t s0 = c * u->K;
    const uint32_t ns = a0.S - s0 < u->K ? a0.S - s0 : u->K;
    void* ring = u->ring[c & 1];
    // the FFT of chunk c overwrites the ring buffer chunk c - 2 was read from
    if (c >= 2 && !one) e = hipStreamWaitEvent(st, u->eq_ev[c - 2], 0);
    if (e != hipSuccess) return hip_err(e, who);
    mega::detail::FftArgs f = u->f;
    f.S = ns;
    f.in = static_cast<const char*>(samples) +
           (uint64_t)s0 * f.in_sym * (f.qbits == 16 ? 2 * sizeof(int16_t) : 2 * sizeof(int8_t));
    f.out = static_cast<float2*>(ring);
    const char* why = nullptr;
    e = mega::detail::launch_fft(f, 0, st, &why);
    if (e != hipSuccess) return hip_err(e, who, why);
    if (!one) {
      e = hipEventRecord(u->fft_ev[c], st);
      if (e == hipSuccess) e = hipStreamWaitEvent(es, u->fft_ev[c], 0);
      if (e != hipSuccess) return hip_err(e, who);
    }
    mega::detail::EqualizeArgs a = a0;
    a.S = ns;
    a.csi = static_cast<const float2*>(w);
    a.y = static_cast<const float2*>(ring);
    a.llr = static_cast<_Float16*>(llr) + (uint64_t)s0 * a0.llr_sym;
    a.xhat = xhat ? static_cast<float2*>(xhat) + (uint64_t)s0 * a0.U * a0.N : nullptr;
    e = u->form == MEGA_FORM_MIXED ? mega::detail::launch_equalize_mixed(a, w, es, &why)
                                   : mega::detail::launch_equalize(a, es, &why);
    if (e != hipSuccess) return hip_err(e, who, why);
    if (!one) e = hipEventRecord(u->eq_ev[c], es);
    if (e != hipSuccess) return hip_err(e, who);
  }
  // the caller's stream continues after the last equalize
  if (!one) {
    e = hipEventRecord(u->join, es);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, u->join, 0);
  }
  if (e != hipSuccess) return hip_err(e, who);
  return MEGA_OK;
}

int mega_uplink_destroy(mega_uplink_t u) {
  if (!u) return MEGA_OK;
  hipError_t e = hipStreamSynchronize(u->es);
  release(u);
  if (e != hipSuccess) return hip_err(e, "mega_uplink_destroy");
  return MEGA_OK;
}

}  // extern "C"
