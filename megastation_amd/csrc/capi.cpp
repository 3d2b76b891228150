// C ABI of libmegadetect: argument checks, error reporting, dispatch to the
// HIP launchers. See include/megadetect.h for the contract of each call and
// the reference interface it replaces.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/megadetect.h"
#include "mega/demod_core.h"
#include "batch.h"
#include "internal.h"
#include "mlp.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what, const char* why = nullptr) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (why) m += std::string(" (") + why + ")";
  return fail(MEGA_ERR_HIP, m);
}

bool valid_q(uint64_t q) { return q == 2 || q == 4 || q == 6 || q == 8; }

}  // namespace

int mega::detail::set_error(int code, const std::string& msg) { return fail(code, msg); }

struct mega_beamform {
  uint32_t num_bs, num_ue, batch_count;
  int device;
  int32_t* info;  // device, [batch_count]
  void* work;     // device, the three-kernel ZF's workspace for batch_count (or NULL)
  uint64_t work_bytes;
};

extern "C" {

const char* mega_version(void) { return "megadetect 0.1 (gfx950)"; }

const char* mega_last_error(void) { return g_err.c_str(); }

int mega_beamform_create(uint32_t num_bs, uint32_t num_ue, uint32_t batch_count,
                         mega_beamform_t* out) {
  if (!out) return fail(MEGA_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (num_bs == 0 || num_ue == 0 || batch_count == 0)
    return fail(MEGA_ERR_INVALID_ARG, "num_bs, num_ue and batch_count must be > 0");
  if (num_ue > mega::detail::kMaxUsersFp32)
    return fail(MEGA_ERR_UNSUPPORTED, "num_ue > 128 is not supported");
  auto* bf = new mega_beamform{num_bs, num_ue, batch_count, 0, nullptr, nullptr, 0};
  hipError_t e = hipGetDevice(&bf->device);
  if (e == hipSuccess) e = hipMalloc(&bf->info, sizeof(int32_t) * batch_count);
  if (e == hipSuccess) e = hipMemset(bf->info, 0, sizeof(int32_t) * batch_count);
  // the workspace of the three-kernel ZF (few wide groups), like the
  // reference's Gram workspace of batch_count groups (beamform.cu:19-31)
  const uint64_t wb = mega::detail::zf_wide_work_bytes(num_ue, num_bs, batch_count);
  if (e == hipSuccess && wb > 0 && mega::detail::zf_wide_pays(num_ue, num_bs, batch_count)) {
    e = hipMalloc(&bf->work, wb);
    if (e == hipSuccess) bf->work_bytes = wb;
  }
  if (e != hipSuccess) {
    if (bf->info) (void)hipFree(bf->info);
    if (bf->work) (void)hipFree(bf->work);
    delete bf;
    return hip_fail(e, "mega_beamform_create");
  }
  *out = bf;
  return MEGA_OK;
}

int mega_beamform_destroy(mega_beamform_t bf) {
  if (!bf) return MEGA_OK;
  hipError_t e = hipFree(bf->info);
  if (bf->work) {
    const hipError_t e2 = hipFree(bf->work);
    if (e == hipSuccess) e = e2;
  }
  delete bf;
  if (e != hipSuccess) return hip_fail(e, "mega_beamform_destroy");
  return MEGA_OK;
}

static int zf_checked(uint32_t num_bs, uint32_t num_ue, uint32_t batch,
                      const void* in, void* out, int32_t* info, int form,
                      mega_stream_t stream, const char* who, void* work = nullptr,
                      uint64_t work_bytes = 0) {
  if (form != MEGA_FORM_FP32 && form != MEGA_FORM_MIXED)
    return fail(MEGA_ERR_UNSUPPORTED, std::string(who) + ": unsupported form " +
                                          std::to_string(form));
  if (batch == 0) return MEGA_OK;
  if (!in || !out) return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL buffer");
  if (num_ue == 0 || num_bs == 0)
    return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": zero size");
  if (num_ue > (form == MEGA_FORM_FP32 ? mega::detail::kMaxUsersFp32 : mega::detail::kMaxUsers))
    return fail(MEGA_ERR_UNSUPPORTED, std::string(who) + (form == MEGA_FORM_FP32
                                                              ? ": num_ue > 128"
                                                              : ": num_ue > 64 (mixed form)"));
  const char* why = nullptr;
  hipError_t e =
      form == MEGA_FORM_MIXED
          ? mega::detail::launch_zf_mixed(num_ue, num_bs, batch, static_cast<const float2*>(in),
                                          static_cast<float2*>(out), nullptr, info, nullptr,
                                          static_cast<hipStream_t>(stream), &why)
          : mega::detail::launch_zf(num_ue, num_bs, batch, static_cast<const float2*>(in),
                                    static_cast<float2*>(out), info,
                                    static_cast<hipStream_t>(stream), &why, work, work_bytes);
  if (e != hipSuccess) return hip_fail(e, who, why);
  return MEGA_OK;
}

int mega_zf_split(uint32_t num_bs, uint32_t num_ue, uint32_t batch,
                  const void* csi_in, void* w_split, void* w_out,
                  int32_t* info_dev, int32_t* route_dev, mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!csi_in || !w_split) return fail(MEGA_ERR_INVALID_ARG, "mega_zf_split: NULL buffer");
  if (num_ue == 0 || num_bs == 0) return fail(MEGA_ERR_INVALID_ARG, "mega_zf_split: zero size");
  if (num_ue > mega::detail::kMaxUsers)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_zf_split: num_ue > 64");
  const char* why = nullptr;
  hipError_t e = mega::detail::launch_zf_mixed(
      num_ue, num_bs, batch, static_cast<const float2*>(csi_in), static_cast<float2*>(w_out),
      w_split, info_dev, route_dev, static_cast<hipStream_t>(stream), &why);
  if (e != hipSuccess) return hip_fail(e, "mega_zf_split", why);
  return MEGA_OK;
}

int mega_beamform_uplink(mega_beamform_t bf, void* csi, uint32_t batch,
                         mega_stream_t stream) {
  if (!bf) return fail(MEGA_ERR_INVALID_ARG, "beamform handle is NULL");
  if (batch > bf->batch_count)
    return fail(MEGA_ERR_BATCH, "Batch size " + std::to_string(batch) +
                                    " exceeds the maximum batch size " +
                                    std::to_string(bf->batch_count));
  return zf_checked(bf->num_bs, bf->num_ue, batch, csi, csi, bf->info,
                    MEGA_FORM_FP32, stream, "mega_beamform_uplink", bf->work, bf->work_bytes);
}

int mega_beamform_downlink(const void* csi, void* precode, uint32_t num_ue,
                           uint32_t num_bs, uint32_t batch,
                           mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!csi || !precode) return fail(MEGA_ERR_INVALID_ARG, "NULL buffer");
  if (num_ue == 0 || num_bs == 0) return fail(MEGA_ERR_INVALID_ARG, "zero size");
  hipError_t e = mega::detail::launch_scale_absmax(
      num_ue, num_bs, batch, static_cast<const float2*>(csi),
      static_cast<float2*>(precode), static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "mega_beamform_downlink");
  return MEGA_OK;
}

int mega_beamformer(mega_beamform_t bf, void* csi, void* precode,
                    uint32_t batch, mega_stream_t stream) {
  if (!bf) return fail(MEGA_ERR_INVALID_ARG, "beamform handle is NULL");
  if (batch > bf->batch_count)
    return fail(MEGA_ERR_BATCH, "Batch size exceeds the maximum batch size");
  int r = mega_beamform_uplink(bf, csi, batch, stream);
  if (r != MEGA_OK) return r;
  return mega_beamform_downlink(csi, precode, bf->num_ue, bf->num_bs, batch,
                                stream);
}

int mega_beamform_info(mega_beamform_t bf, int32_t* host_info, uint32_t batch,
                       mega_stream_t stream) {
  if (!bf || !host_info) return fail(MEGA_ERR_INVALID_ARG, "NULL argument");
  if (batch > bf->batch_count) return fail(MEGA_ERR_BATCH, "batch > batch_count");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(host_info, bf->info, sizeof(int32_t) * batch,
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "mega_beamform_info");
  int failed = 0;
  for (uint32_t i = 0; i < batch; ++i) failed += host_info[i] != 0;
  return failed;
}

int mega_zf(uint32_t num_bs, uint32_t num_ue, uint32_t batch,
            const void* csi_in, void* w_out, int32_t* info_dev, int form,
            mega_stream_t stream) {
  return zf_checked(num_bs, num_ue, batch, csi_in, w_out, info_dev, form,
                    stream, "mega_zf");
}

static int eq_args_checked(const mega_equalize_desc* d, mega::detail::EqualizeArgs& a,
                           bool& empty, const char* who, bool allow_mixed = false,
                           bool need_buffers = true) {
  empty = false;
  if (!d) return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": desc is NULL");
  if (!valid_q(d->mod_order))
    return fail(MEGA_ERR_UNSUPPORTED,
                "unsupported modulation order " + std::to_string(d->mod_order));
  if (d->form != MEGA_FORM_FP32 && !(allow_mixed && d->form == MEGA_FORM_MIXED))
    return fail(MEGA_ERR_UNSUPPORTED, std::string(who) + ": unsupported form " +
                                          std::to_string(d->form));
  if (d->num_ue == 0 || d->num_bs == 0 || d->num_sc == 0 || d->num_syms == 0 ||
      d->num_slots == 0) {
    empty = true;
    return MEGA_OK;
  }
  if (need_buffers && (!d->csi || !d->y || !d->llr))
    return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL buffer");
  if (d->sc_group == 0 || d->cb_size == 0)
    return fail(MEGA_ERR_INVALID_ARG, "sc_group and cb_size must be > 0");
  if (d->num_ue > (d->form == MEGA_FORM_FP32 ? mega::detail::kMaxUsersFp32
                                             : mega::detail::kMaxUsers))
    return fail(MEGA_ERR_UNSUPPORTED, d->form == MEGA_FORM_FP32 ? "num_ue > 128"
                                                                : "num_ue > 64 (mixed form)");
  if (d->num_slots > 65535 || d->num_syms > 65535)
    return fail(MEGA_ERR_UNSUPPORTED, "more than 65535 slots or symbols per launch");
  a.csi = static_cast<const float2*>(d->csi);
  a.y = static_cast<const float2*>(d->y);
  a.llr = static_cast<_Float16*>(d->llr);
  a.xhat = static_cast<float2*>(d->xhat);
  a.U = d->num_ue;
  a.B = d->num_bs;
  a.N = d->num_sc;
  a.sg = d->sc_group;
  a.G = (d->num_sc + d->sc_group - 1) / d->sc_group;
  a.S = d->num_syms;
  a.slots = d->num_slots;
  a.Q = d->mod_order;
  a.ue_stride = d->ue_stride;
  a.cb_size = d->cb_size;
  a.cb_stride = d->cb_stride;
  a.zeros = d->zeros;
  a.valid_bits = d->valid_bits;
  a.csi_slot = d->csi_slot_stride;
  a.y_sym = d->y_sym_stride;
  a.y_slot = d->y_slot_stride;
  a.llr_sym = d->llr_sym_stride;
  a.llr_slot = d->llr_slot_stride;
  return MEGA_OK;
}

}  // extern "C"

int mega::detail::equalize_args_checked(const mega_equalize_desc* d, EqualizeArgs& a,
                                        bool& empty, const char* who, bool allow_mixed,
                                        bool need_buffers) {
  return eq_args_checked(d, a, empty, who, allow_mixed, need_buffers);
}

extern "C" {

int mega_equalize_slots(const mega_equalize_desc* d, mega_stream_t stream) {
  mega::detail::EqualizeArgs a;
  bool empty;
  int r = eq_args_checked(d, a, empty, "mega_equalize_slots", true);
  if (r != MEGA_OK || empty) return r;
  const char* why = nullptr;
  hipError_t e = d->form == MEGA_FORM_MIXED
                     ? mega::detail::launch_equalize_mixed(a, d->csi,
                                                           static_cast<hipStream_t>(stream), &why)
                     : mega::detail::launch_equalize(a, static_cast<hipStream_t>(stream), &why);
  if (e != hipSuccess) return hip_fail(e, "mega_equalize", why);
  return MEGA_OK;
}

static int detect_checked(const mega_detect_desc* d, void* work, mega_stream_t stream,
                          const char* who, bool may_alloc = true) {
  mega::detail::EqualizeArgs a;
  bool empty;
  int r = eq_args_checked(&d->eq, a, empty, who, true);
  if (r != MEGA_OK || empty) return r;
  if (d->eq.form == MEGA_FORM_MIXED) {
    if (!work) return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": the mixed form needs work");
    const char* why = nullptr;
    hipError_t e = mega::detail::launch_detect_mixed(
        a, static_cast<float2*>(d->w_out), work, d->info_dev, d->route_dev,
        static_cast<hipStream_t>(stream), &why);
    if (e != hipSuccess) return hip_fail(e, who, why);
    return MEGA_OK;
  }
  if (!d->w_out && !mega::detail::detect_fused_supported(a.U, a.B))
    return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": U > 16 or B > 256 needs w_out");
  const char* why = nullptr;
  hipError_t e = mega::detail::launch_detect(a, static_cast<float2*>(d->w_out),
                                             d->info_dev,
                                             static_cast<hipStream_t>(stream), &why,
                                             may_alloc);
  if (e != hipSuccess) return hip_fail(e, who, why);
  return MEGA_OK;
}

int mega_detect_slots(const mega_detect_desc* d, mega_stream_t stream) {
  if (!d) return fail(MEGA_ERR_INVALID_ARG, "mega_detect_slots: desc is NULL");
  return detect_checked(d, d->work, stream, "mega_detect_slots");
}

int mega_detect_slots_noalloc(const mega_detect_desc* d, mega_stream_t stream) {
  static const char* who = "mega_detect_slots_noalloc";
  if (!d) return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": desc is NULL");
  if (d->work)
    return fail(MEGA_ERR_INVALID_ARG,
                std::string(who) + ": work must be NULL (the library's scratch is used)");
  // the FP32 form needs no scratch; its ZF takes the one-workgroup-per-group
  // kernel here (the three-kernel form's workspace would be an allocation)
  if (d->eq.form != MEGA_FORM_MIXED) return detect_checked(d, nullptr, stream, who, false);
  const uint64_t need = mega_detect_work_bytes(d);
  if (need > mega::detail::kDetectScratchBytes)
    return fail(MEGA_ERR_UNSUPPORTED,
                std::string(who) + ": needs " + std::to_string(need) +
                    " bytes of scratch, more than its maximum of " +
                    std::to_string(mega::detail::kDetectScratchBytes) +
                    " (fewer slots per call, or mega_detect_slots with a work buffer)");
  void* work = nullptr;
  if (need > 0) {
    const char* why = nullptr;
    hipError_t e = mega::detail::detect_scratch(&work, &why);
    if (e != hipSuccess) return hip_fail(e, who, why);
  }
  return detect_checked(d, work, stream, who);
}

uint64_t mega_detect_work_bytes(const mega_detect_desc* d) {
  if (!d || d->eq.form != MEGA_FORM_MIXED || d->eq.sc_group == 0) return 0;
  const uint64_t G = (d->eq.num_sc + d->eq.sc_group - 1) / d->eq.sc_group;
  const uint64_t Bp = (d->eq.num_bs + 15) / 16 * 16;  // split planes' row length
  return (uint64_t)d->eq.num_slots * G * Bp * d->eq.num_ue * 8;
}

int mega_equalize(uint64_t row, uint64_t col, uint64_t batch_count,
                  const void* csi, const void* bsvec, void* uevec,
                  uint64_t ue_stride, uint64_t csi_skip, uint8_t mod_order,
                  uint64_t cb_size, uint64_t cb_stride, uint64_t zeros,
                  uint64_t valid_bits, mega_stream_t stream) {
  if (row > 0xffffffffull || col > 0xffffffffull || batch_count > 0xffffffffull ||
      csi_skip > 0xffffffffull)
    return fail(MEGA_ERR_UNSUPPORTED, "dimension exceeds 32 bits");
  mega_equalize_desc d;
  std::memset(&d, 0, sizeof(d));
  d.num_ue = (uint32_t)row;
  d.num_bs = (uint32_t)col;
  d.num_sc = (uint32_t)batch_count;
  d.sc_group = (uint32_t)csi_skip;
  d.num_syms = 1;
  d.num_slots = 1;
  d.mod_order = mod_order;
  d.form = MEGA_FORM_FP32;
  d.ue_stride = ue_stride;
  d.cb_size = cb_size;
  d.cb_stride = cb_stride;
  d.zeros = zeros;
  d.valid_bits = valid_bits;
  d.csi = csi;
  d.y = bsvec;
  d.llr = uevec;
  return mega_equalize_slots(&d, stream);
}

static int fft_checked(const mega_fft_desc* d, int mode, uint32_t ue_num, uint32_t sc_group,
                       uint32_t ue_start, const void* pilot_sign, mega_stream_t stream,
                       const char* who) {
  const bool pilot = mode == 1;
  if (!d) return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": desc is NULL");
  if (d->bs_num == 0 || d->num_syms == 0 || (mode != 2 && d->ofdm_num == 0)) return MEGA_OK;
  if (!d->in || !d->out) return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL buffer");
  if (pilot && !pilot_sign)
    return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": NULL pilot sign table");
  if (pilot && (sc_group == 0 || ue_num == 0))
    return fail(MEGA_ERR_INVALID_ARG, std::string(who) + ": zero users or group size");
  mega::detail::FftArgs a;
  std::memset(&a, 0, sizeof a);
  a.in = d->in;
  a.out = static_cast<float2*>(d->out);
  a.qout = d->out;
  a.pilot = static_cast<const float2*>(pilot_sign);
  a.B = d->bs_num;
  a.S = d->num_syms;
  a.ca = d->ofdm_ca;
  a.num = d->ofdm_num;
  a.start = d->ofdm_start;
  a.qbits = (uint32_t)d->quant_bits;
  a.U = ue_num;
  a.sg = sc_group;
  a.ue_start = ue_start;
  a.in_sym = d->in_sym_stride ? d->in_sym_stride : (uint64_t)d->bs_num * d->ofdm_ca;
  // pilot symbols of one launch fill one CSI buffer (their users differ)
  // unless a stride is given
  a.out_sym = d->out_sym_stride ? d->out_sym_stride
              : pilot     ? 0
              : mode == 2 ? (uint64_t)d->bs_num * d->ofdm_ca
                          : (uint64_t)d->bs_num * d->ofdm_num;
  const char* why = nullptr;
  hipError_t e = mega::detail::launch_fft(a, mode, static_cast<hipStream_t>(stream), &why);
  if (e == hipErrorInvalidValue && why) return fail(MEGA_ERR_UNSUPPORTED, std::string(who) + ": " + why);
  if (e != hipSuccess) return hip_fail(e, who, why);
  return MEGA_OK;
}

int mega_uplink_fft(const mega_fft_desc* d, mega_stream_t stream) {
  return fft_checked(d, 0, 0, 0, 0, nullptr, stream, "mega_uplink_fft");
}

int mega_downlink_ifft(const mega_fft_desc* d, mega_stream_t stream) {
  return fft_checked(d, 2, 0, 0, 0, nullptr, stream, "mega_downlink_ifft");
}

int mega_pilot_fft(const mega_fft_desc* d, uint32_t ue_num, uint32_t sc_group, uint32_t ue_start,
                   const void* pilot_sign, mega_stream_t stream) {
  return fft_checked(d, 1, ue_num, sc_group, ue_start, pilot_sign, stream, "mega_pilot_fft");
}

// Largest prime <= n among the primes up to 2039, else 2039: the length rule
// of the reference's Zadoff-Chu generator (utils.h:54-64), whose prime table
// ends at 2039.
static uint32_t zc_length(uint64_t n) {
  uint32_t best = 0;
  for (uint32_t p = 2; p <= 2039; ++p) {
    bool prime = true;
    for (uint32_t d = 2; d * d <= p; ++d)
      if (p % d == 0) {
        prime = false;
        break;
      }
    if (!prime) continue;
    if (p <= n) best = p;
    else break;
  }
  // n >= 2039 or no prime <= n (n < 2): the table's last entry
  return (best == 0 || n >= 2039) ? 2039 : best;
}

int mega_pilot_sign(uint32_t ofdm_num, float* host_out) {
  // every product and sum rounded on its own, as the reference's host code
#pragma clang fp contract(off)
  if (!host_out && ofdm_num) return fail(MEGA_ERR_INVALID_ARG, "mega_pilot_sign: NULL output");
  const uint32_t L = zc_length(ofdm_num);
  const double r = std::floor(L * 2.0 / 31.0 + 0.5);
  const float alpha = (float)(M_PI / 4);
  for (uint32_t i = 0; i < ofdm_num; ++i) {
    // Zadoff-Chu root r (utils.h:66-72): the phase formed in double, its
    // sine and cosine in float
    const uint32_t m = i % L;
    float zs, zc;
    sincosf((float)(-M_PI * r * m * (m + 1) / L), &zs, &zc);
    // cyclic shift by alpha * i (utils.h:77-95), product as mega::Complex
    float cs, cc;
    sincosf(alpha * (float)i, &cs, &cc);
    const float re = zc * cc - zs * cs, im = zc * cs + zs * cc;
    // divided by its squared magnitude (fft_op.cu:183-186)
    const float d = re * re + im * im;
    host_out[2 * i] = re / d;
    host_out[2 * i + 1] = im / d;
  }
  return MEGA_OK;
}

int mega_precode_syms(const mega_precode_desc* d, mega_stream_t stream) {
  if (!d) return fail(MEGA_ERR_INVALID_ARG, "mega_precode: desc is NULL");
  if (d->num_bs == 0 || d->num_ue == 0 || d->num_syms == 0 || d->sc_group == 0 ||
      d->num_sc < d->sc_group)
    return MEGA_OK;  // no full group: nothing to write (as the reference)
  if (!d->precode || !d->uevec || !d->bsvec)
    return fail(MEGA_ERR_INVALID_ARG, "mega_precode: NULL buffer");
  if (d->bs_stride == 0) return fail(MEGA_ERR_INVALID_ARG, "mega_precode: bs_stride is 0");
  if (d->num_ue > mega::detail::kMaxUsers)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_precode: num_ue > 64");
  mega::detail::PrecodeArgs a;
  a.prec = static_cast<const float2*>(d->precode);
  a.ue = static_cast<const float2*>(d->uevec);
  a.out = static_cast<float2*>(d->bsvec);
  a.B = d->num_bs;
  a.U = d->num_ue;
  a.N = d->num_sc;
  a.sg = d->sc_group;
  a.S = d->num_syms;
  a.row = d->bs_stride;
  a.bs_start = d->bs_start;
  a.ue_sym = d->ue_sym_stride;
  a.out_sym = d->bs_sym_stride;
  const char* why = nullptr;
  hipError_t e = mega::detail::launch_precode(a, static_cast<hipStream_t>(stream), &why);
  if (e != hipSuccess) return hip_fail(e, "mega_precode", why);
  return MEGA_OK;
}

int mega_ldpc_decode(const mega_ldpc_desc* d, mega_stream_t stream) {
  if (!d) return fail(MEGA_ERR_INVALID_ARG, "mega_ldpc_decode: desc is NULL");
  if (d->base_graph != 1 && d->base_graph != 2)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_ldpc_decode: base_graph must be 1 or 2");
  if (mega::detail::ldpc_set_index(d->lifting_size) < 0)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_ldpc_decode: lifting size " +
                                          std::to_string(d->lifting_size) +
                                          " is not a 5G NR lifting size");
  const uint32_t kb = d->base_graph == 1 ? 22u : 10u;
  const uint32_t max_mb = d->base_graph == 1 ? 46u : 42u;
  if (d->parity_nodes < 4 || d->parity_nodes > max_mb)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_ldpc_decode: parity_nodes out of range");
  if (d->max_iterations < 1 || d->max_iterations > 100)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_ldpc_decode: max_iterations must be 1..100");
  if (d->llr_bits != 16 && d->llr_bits != 32)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_ldpc_decode: llr_bits must be 16 or 32");
  if (!(d->normalization >= 0.f && d->normalization <= 1.f))
    return fail(MEGA_ERR_INVALID_ARG, "mega_ldpc_decode: normalization must be in [0, 1]");
  if (d->num_cb == 0) return MEGA_OK;
  if (d->num_cb > 0x7fffffffu) return fail(MEGA_ERR_UNSUPPORTED, "mega_ldpc_decode: num_cb too large");
  if (!d->llr || !d->out) return fail(MEGA_ERR_INVALID_ARG, "mega_ldpc_decode: NULL buffer");
  const uint64_t nv = (uint64_t)(kb + d->parity_nodes) * d->lifting_size;
  if (d->llr_stride < nv)
    return fail(MEGA_ERR_INVALID_ARG, "mega_ldpc_decode: llr_stride < (Kb + mb) * Zc");
  if (d->out_stride < ((uint64_t)kb * d->lifting_size + 31) / 32)
    return fail(MEGA_ERR_INVALID_ARG, "mega_ldpc_decode: out_stride < ceil(Kb * Zc / 32)");
  mega::detail::LdpcArgs a;
  a.bg = d->base_graph;
  a.zc = d->lifting_size;
  a.mb = d->parity_nodes;
  a.iters = d->max_iterations;
  a.num_cb = d->num_cb;
  a.llr_bits = d->llr_bits;
  a.alpha = d->normalization > 0.f ? d->normalization : 0.8125f;
  a.llr = d->llr;
  a.llr_stride = d->llr_stride;
  a.out = d->out;
  a.out_stride = d->out_stride;
  const char* why = nullptr;
  hipError_t e = mega::detail::launch_ldpc_decode(a, static_cast<hipStream_t>(stream), &why);
  if (e != hipSuccess) return hip_fail(e, "mega_ldpc_decode", why);
  return MEGA_OK;
}

int mega_precode(uint64_t row, uint64_t col, uint64_t batch_count, const void* precode,
                 const void* uevec, void* bsvec, uint64_t bs_stride, uint64_t precode_skip,
                 uint64_t bs_start, mega_stream_t stream) {
  if (row > 0xffffffffull || col > 0xffffffffull || batch_count > 0xffffffffull ||
      precode_skip > 0xffffffffull)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_precode: size exceeds 32 bits");
  mega_precode_desc d;
  std::memset(&d, 0, sizeof d);
  d.num_bs = (uint32_t)row;
  d.num_ue = (uint32_t)col;
  d.num_sc = (uint32_t)batch_count;
  d.sc_group = (uint32_t)precode_skip;
  d.num_syms = 1;
  d.precode = precode;
  d.uevec = uevec;
  d.bsvec = bsvec;
  d.bs_stride = bs_stride;
  d.bs_start = bs_start;
  return mega_precode_syms(&d, stream);
}

static int hip_status(hipError_t e, const char* who, const char* why = nullptr) {
  return e == hipSuccess ? MEGA_OK : hip_fail(e, who, why);
}

int mega_gram_inv(uint32_t num_bs, uint32_t num_ue, uint32_t batch, const void* csi,
                  void* G_out, void* X_out, int32_t* info_dev, mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!csi || (!X_out && !G_out)) return fail(MEGA_ERR_INVALID_ARG, "mega_gram_inv: NULL buffer");
  if (num_ue == 0 || num_bs == 0) return fail(MEGA_ERR_INVALID_ARG, "mega_gram_inv: zero size");
  if (num_ue > mega::detail::kMaxUsers) return fail(MEGA_ERR_UNSUPPORTED, "mega_gram_inv: num_ue > 64");
  return hip_status(mega::detail::launch_gram_inv(num_ue, num_bs, batch, static_cast<const float2*>(csi),
                                                  static_cast<float2*>(G_out), static_cast<float2*>(X_out),
                                                  info_dev, static_cast<hipStream_t>(stream)),
                    "mega_gram_inv");
}

int mega_learned_features(uint32_t num_ue, uint32_t batch, const void* G, const void* G_prev,
                          const void* X_prev, void* feat_bf16, float* scale, mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!G || !G_prev || !X_prev || !feat_bf16 || !scale)
    return fail(MEGA_ERR_INVALID_ARG, "mega_learned_features: NULL buffer");
  if (num_ue == 0 || num_ue > mega::detail::kMaxUsers)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_learned_features: num_ue outside 1..64");
  return hip_status(mega::detail::launch_learned_features(
                        num_ue, batch, static_cast<const float2*>(G), static_cast<const float2*>(G_prev),
                        static_cast<const float2*>(X_prev), static_cast<__bf16*>(feat_bf16), scale,
                        static_cast<hipStream_t>(stream)),
                    "mega_learned_features");
}

int mega_learned_target(uint32_t num_ue, uint32_t batch, const void* X_true, const void* X_base,
                        const float* scale, float* target, mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!X_true || !X_base || !scale || !target)
    return fail(MEGA_ERR_INVALID_ARG, "mega_learned_target: NULL buffer");
  if (num_ue == 0 || num_ue > mega::detail::kMaxUsers)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_learned_target: num_ue outside 1..64");
  return hip_status(mega::detail::launch_learned_target(
                        num_ue, batch, static_cast<const float2*>(X_true),
                        static_cast<const float2*>(X_base), scale, target,
                        static_cast<hipStream_t>(stream)),
                    "mega_learned_target");
}

int mega_learned_refine(uint32_t num_ue, uint32_t batch, const void* G, const void* G_prev,
                        const void* X_prev, const float* corr, uint32_t corr_ld, const float* scale,
                        void* X_out, float* res_out, int32_t* flag_out, float bound,
                        uint32_t ns_iters, mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!G || !G_prev || !X_prev || !X_out || !res_out || !flag_out || (corr && !scale))
    return fail(MEGA_ERR_INVALID_ARG, "mega_learned_refine: NULL buffer");
  if (num_ue == 0 || num_ue > mega::detail::kMaxUsers)
    return fail(MEGA_ERR_UNSUPPORTED, "mega_learned_refine: num_ue outside 1..64");
  if (corr && corr_ld < num_ue * (num_ue + 1))
    return fail(MEGA_ERR_INVALID_ARG, "mega_learned_refine: corr_ld < U*(U+1)");
  return hip_status(mega::detail::launch_learned_refine(
                        num_ue, batch, static_cast<const float2*>(G), static_cast<const float2*>(G_prev),
                        static_cast<const float2*>(X_prev), corr, corr_ld, scale,
                        static_cast<float2*>(X_out), res_out, flag_out, bound, ns_iters,
                        static_cast<hipStream_t>(stream)),
                    "mega_learned_refine");
}

int mega_learned_finish(uint32_t num_bs, uint32_t num_ue, uint32_t batch, const void* csi,
                        const void* G, void* X, const int32_t* flag, void* w_out,
                        int32_t* info_dev, mega_stream_t stream) {
  if (batch == 0) return MEGA_OK;
  if (!csi || !X || !flag || !w_out) return fail(MEGA_ERR_INVALID_ARG, "mega_learned_finish: NULL buffer");
  if (num_ue == 0 || num_bs == 0) return fail(MEGA_ERR_INVALID_ARG, "mega_learned_finish: zero size");
  if (num_ue > mega::detail::kMaxUsers) return fail(MEGA_ERR_UNSUPPORTED, "mega_learned_finish: num_ue > 64");
  return hip_status(mega::detail::launch_learned_finish(
                        num_ue, num_bs, batch, static_cast<const float2*>(csi),
                        static_cast<const float2*>(G), static_cast<float2*>(X), flag,
                        static_cast<float2*>(w_out), info_dev, static_cast<hipStream_t>(stream)),
                    "mega_learned_finish");
}

int mega_mlp_gemm(const mega_mlp_gemm_desc* d, mega_stream_t stream) {
  if (!d || !d->A || !d->B || (!d->Cf && !d->Cb))
    return fail(MEGA_ERR_INVALID_ARG, "mega_mlp_gemm: NULL buffer");
  if (d->act != 0 && d->act != 1) return fail(MEGA_ERR_UNSUPPORTED, "mega_mlp_gemm: activation");
  if (d->M > 65535u * 64u) return fail(MEGA_ERR_UNSUPPORTED, "mega_mlp_gemm: M too large");
  mega::detail::MlpGemm g{static_cast<const __bf16*>(d->A), d->sai, d->sak,
                          static_cast<const __bf16*>(d->B), d->sbk, d->sbj, d->M, d->N, d->K,
                          d->bias, d->act, d->Cf, static_cast<__bf16*>(d->Cb), d->ldc, d->beta};
  return hip_status(mega::detail::mlp_gemm(g, static_cast<hipStream_t>(stream)), "mega_mlp_gemm");
}

int mega_mlp_act_backward(const float* dY, const float* Z, void* dZ_bf16, float* dZf, uint64_t n,
                          int32_t act, mega_stream_t stream) {
  if (n && (!dY || !Z || !dZ_bf16)) return fail(MEGA_ERR_INVALID_ARG, "mega_mlp_act_backward: NULL buffer");
  return hip_status(mega::detail::mlp_act_backward(dY, Z, static_cast<__bf16*>(dZ_bf16), dZf, n, act,
                                                   static_cast<hipStream_t>(stream)),
                    "mega_mlp_act_backward");
}

int mega_mlp_col_sum(const float* dZ, float* db, uint32_t M, uint32_t N, mega_stream_t stream) {
  if (N && (!dZ || !db)) return fail(MEGA_ERR_INVALID_ARG, "mega_mlp_col_sum: NULL buffer");
  return hip_status(mega::detail::mlp_col_sum(dZ, db, M, N, static_cast<hipStream_t>(stream)),
                    "mega_mlp_col_sum");
}

int mega_mlp_mse_grad(const float* Y, const float* T, float* dY, uint64_t n, uint32_t M,
                      mega_stream_t stream) {
  if (n && (!Y || !T || !dY || M == 0)) return fail(MEGA_ERR_INVALID_ARG, "mega_mlp_mse_grad: bad argument");
  return hip_status(mega::detail::mlp_mse_grad(Y, T, dY, n, M, static_cast<hipStream_t>(stream)),
                    "mega_mlp_mse_grad");
}

int mega_mlp_adam(float* w, void* wb, const float* g, float* m, float* v, uint64_t n, float lr,
                  float beta1, float beta2, float eps, uint32_t step, mega_stream_t stream) {
  if (n && (!w || !g || !m || !v || step == 0))
    return fail(MEGA_ERR_INVALID_ARG, "mega_mlp_adam: bad argument");
  return hip_status(mega::detail::mlp_adam(w, static_cast<__bf16*>(wb), g, m, v, n, lr, beta1, beta2,
                                           eps, step, static_cast<hipStream_t>(stream)),
                    "mega_mlp_adam");
}

int mega_mlp_to_bf16(const float* x, void* y, uint64_t n, mega_stream_t stream) {
  if (n && (!x || !y)) return fail(MEGA_ERR_INVALID_ARG, "mega_mlp_to_bf16: NULL buffer");
  return hip_status(mega::detail::mlp_to_bf16(x, static_cast<__bf16*>(y), n,
                                              static_cast<hipStream_t>(stream)),
                    "mega_mlp_to_bf16");
}

int mega_demod_host(const float* xhat_c64, uint64_t n, uint8_t mod_order,
                    uint16_t* llr_f16) {
  if (!valid_q(mod_order))
    return fail(MEGA_ERR_UNSUPPORTED,
                "unsupported modulation order " + std::to_string(mod_order));
  if (n && (!xhat_c64 || !llr_f16)) return fail(MEGA_ERR_INVALID_ARG, "NULL buffer");
  for (uint64_t i = 0; i < n; ++i) {
    _Float16 v[8];
    mega::demod::demod_any(xhat_c64[2 * i], xhat_c64[2 * i + 1], mod_order, v);
    std::memcpy(llr_f16 + i * mod_order, v, sizeof(_Float16) * mod_order);
  }
  return MEGA_OK;
}

}  // extern "C"
