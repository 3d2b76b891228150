// Launchers shared between the kernel translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/megadetect.h"

namespace mega {
namespace detail {

// One equalize + demod launch over [slots][syms] OFDM symbols.
struct EqualizeArgs {
  const float2* csi;
  const float2* y;
  _Float16* llr;
  float2* xhat;  // optional [slots][syms][U][N]
  uint32_t U, B, N, sg, G, S, slots, Q;
  uint64_t ue_stride, cb_size, cb_stride, zeros, valid_bits;
  uint64_t csi_slot, y_sym, y_slot, llr_sym, llr_slot;
};

// Returns hipSuccess or the launch error; *why gets a static message for
// argument combinations the kernels cannot take.
hipError_t launch_equalize(const EqualizeArgs& a, hipStream_t stream,
                           const char** why);

// FP32 ZF. Where zf_wide_pays (no shape at present: measured no faster) or
// MEGA_ZF_WIDE=2 asks for it, 17 to 64 users take the three-kernel form
// (zf_wide.hip) when a workspace of zf_wide_work_bytes is at hand:
// `work` (work_bytes) if given, else the stream's library workspace when
// may_alloc (never while the stream is being captured); otherwise, and for
// every other shape, one workgroup per group (zf.hip). W and the status are
// the same bits either way.
hipError_t launch_zf(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                     float2* out, int32_t* info, hipStream_t stream,
                     const char** why, void* work = nullptr, uint64_t work_bytes = 0,
                     bool may_alloc = true);
uint64_t zf_wide_work_bytes(uint32_t U, uint32_t B, uint32_t batch);
bool zf_wide_pays(uint32_t U, uint32_t B, uint32_t batch);
hipError_t launch_zf_wide(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                          float2* out, int32_t* info, void* work, hipStream_t stream,
                          const char** why);
// The library's device workspace of stream st (at least `need` bytes; see
// workspace.cpp). hipErrorStreamCaptureUnsupported while st is capturing.
hipError_t stream_workspace(hipStream_t st, size_t need, void** out);

// ZF + equalize + demod of every symbol of every slot: fused in one kernel for
// U <= 16 and B <= 256 (w_out optional), otherwise the ZF and equalize kernels
// back to back through w_out (required then).
hipError_t launch_detect(const EqualizeArgs& a, float2* w_out, int32_t* info,
                         hipStream_t stream, const char** why, bool may_alloc = true);
bool detect_fused_supported(uint32_t U, uint32_t B);

// Mixed form (split-BF16 matrix-core products; zf_mixed.hip): W as FP32
// (w_out, optional) and as split BF16 words (wp, optional, 8 bytes per
// weight, CSI indexing); route[g] = 1 where the group took the FP64 route.
hipError_t launch_zf_mixed(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                           float2* w_out, void* wp, int32_t* info, int32_t* route,
                           hipStream_t stream, const char** why);
// Equalize + demod from the split words (a.csi unused; wp with CSI indexing
// and slot stride a.csi_slot).
hipError_t launch_equalize_mixed(const EqualizeArgs& a, const void* wp,
                                 hipStream_t stream, const char** why);
// Mixed-form detection: zf_mixed into `work` (slots*G*B*U*8 bytes), then the
// mixed equalize.
hipError_t launch_detect_mixed(const EqualizeArgs& a, float2* w_out, void* work,
                               int32_t* info, int32_t* route, hipStream_t stream,
                               const char** why);
// Device address, on the current device, of the library's static scratch for
// mega_detect_slots_noalloc (kDetectScratchBytes; looked up once per device).
constexpr uint64_t kDetectScratchBytes = 32ull << 20;
hipError_t detect_scratch(void** out, const char** why);

// Learned form (learned.hip): Gram and plain inverse per group; network
// features; warm start + Newton-Schulz + residual check; final W with the
// plain-sweep fallback. Matrices row-major [batch][U][U] complex.
hipError_t launch_gram_inv(uint32_t U, uint32_t B, uint32_t batch, const float2* in, float2* G,
                           float2* X, int32_t* info, hipStream_t stream);
hipError_t launch_learned_features(uint32_t U, uint32_t batch, const float2* G, const float2* Gp,
                                   const float2* Xp, __bf16* feat, float* scale,
                                   hipStream_t stream);
hipError_t launch_learned_target(uint32_t U, uint32_t batch, const float2* Xt, const float2* Xb,
                                 const float* scale, float* tgt, hipStream_t stream);
hipError_t launch_learned_refine(uint32_t U, uint32_t batch, const float2* G, const float2* Gp,
                                 const float2* Xp, const float* corr, uint32_t corr_ld,
                                 const float* scale, float2* X_out, float* res, int32_t* flag,
                                 float bound, uint32_t iters, hipStream_t stream);
hipError_t launch_learned_finish(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                                 const float2* Gm, float2* X, const int32_t* flag, float2* w,
                                 int32_t* info, hipStream_t st);

// Downlink precode of S symbols (precode.hip): prec [G][U][B], ue [U][N] per
// symbol (ue_sym apart), out [B][row] per symbol (out_sym apart); strides in
// complex elements.
struct PrecodeArgs {
  const float2* prec;
  const float2* ue;
  float2* out;
  uint32_t B, U, N, sg, S;
  uint64_t row, bs_start, ue_sym, out_sym;
};
hipError_t launch_precode(const PrecodeArgs& a, hipStream_t stream, const char** why);

// Uplink / pilot FFT of S symbols x B antennas (fft.hip): in [S][B][ca]
// integer (re, im) pairs (in_sym samples apart), out y [S][B][num] or CSI
// [G][B][U] (out_sym complex elements apart).
struct FftArgs {
  const void* in;
  float2* out;
  void* qout;           // downlink IFFT only: integer pairs
  const float2* pilot;  // pilot only: [num]
  uint32_t B, S, ca, num, start, qbits;
  uint32_t U, sg, ue_start;  // pilot only
  uint64_t in_sym, out_sym;
  uint32_t log2ca;  // set by the launcher
  float qscale;     // set by the launcher
  uint32_t sg_inv;  // set by the launcher: div_inv(sg), pilot only
};
// mode: 0 uplink, 1 pilot, 2 downlink IFFT
hipError_t launch_fft(const FftArgs& a, int mode, hipStream_t stream, const char** why);

// LDPC decoding (ldpc.hip): num_cb code blocks of base graph bg (1 or 2),
// lifting size zc, mb parity block rows, iters layered min-sum iterations
// with scale alpha; llr binary16 (llr_bits 16) or float (32), llr_stride
// elements between code blocks; out bit words, out_stride words apart.
struct LdpcArgs {
  uint32_t bg, zc, mb, iters, num_cb;
  int32_t llr_bits;
  float alpha;
  const void* llr;
  uint64_t llr_stride;
  uint32_t* out;
  uint64_t out_stride;
};
int ldpc_set_index(uint32_t zc);
hipError_t launch_ldpc_decode(const LdpcArgs& a, hipStream_t stream, const char** why);

hipError_t launch_scale_absmax(uint32_t U, uint32_t B, uint32_t batch,
                               const float2* in, float2* out,
                               hipStream_t stream);

// Tuning overrides (tuning.cpp): MEGA_* environment variables parsed once
// (and on mega_tuning_reload), sanitized to each key's accepted range; 0 is
// the launcher's own choice except for kDetPf (default 1).
enum TuneKey {
  kEqKs, kEqMode, kEqSch, kEqSmt, kEqTpw, kEqXmt, kEqXreg, kEqXt, kEqXks, kPrecAcc, kDetPf,
  kEqWtc, kZfNth, kZfWide, kUplinkStreams, kEqWnw, kEqWp,
  kTuneCount
};
int tune(TuneKey k);
void tuning_reload();

// Limits of the kernels (checked by the C ABI before launching).
constexpr uint32_t kMaxUsers = 64;       // the chosen forms, precode, batch
constexpr uint32_t kMaxUsersFp32 = 128;  // the FP32 ZF and equalize

// Records a message for mega_last_error and returns code (capi.cpp).
int set_error(int code, const std::string& msg);

// Checks an equalize descriptor and fills a from it (capi.cpp): empty is set
// for a zero-sized launch; allow_mixed admits MEGA_FORM_MIXED; need_buffers
// requires csi, y and llr.
int equalize_args_checked(const mega_equalize_desc* d, EqualizeArgs& a, bool& empty,
                          const char* who, bool allow_mixed, bool need_buffers);

}  // namespace detail
}  // namespace mega
