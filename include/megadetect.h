/*
 * megadetect.h — C ABI of the MI355X uplink detection path (ZF beamforming
 * weights + fused equalize / max-log demod), libmegadetect.so.
 *
 * Plain C: fixed-width integers, raw device pointers, no exceptions. Every
 * call runs asynchronously on the given stream of the current HIP device
 * (stream = a hipStream_t; NULL is the device's null stream) unless it says it
 * synchronises. Functions return MEGA_OK (0) or a negative MEGA_ERR_* code;
 * mega_last_error() then holds a message for the calling thread.
 *
 * Buffer layouts are the reference's (src/operators/mega_complex.h:25-27,
 * src/operators/matrix/matrix.h:95-117, src/baseband/buffers.cc:77-103):
 *   complex      interleaved {float re, float im} (complex64)
 *   CSI / W      [G][B][U]: element (u,b,g) at u + U*b + U*B*g, i.e. one
 *                column-major U x B matrix A_g = H_g^H per subcarrier group
 *   precoder     [G][U][B]: one column-major B x U matrix per group
 *   y (bsVec)    [B][N] per OFDM symbol, subcarrier fastest
 *   LLR (ueVec)  binary16, one row of ue_stride halves per user per symbol;
 *                bit j of subcarrier k at
 *                (kQ / cb_size)*cb_stride + zeros + (kQ % cb_size) + j,
 *                written when kQ + j < valid_bits and the offset < ue_stride.
 * Subcarrier k uses group k / sg; a partial last group (N % sg != 0) is
 * allowed, and the CSI buffer then holds ceil(N / sg) groups.
 */
#ifndef MEGADETECT_H
#define MEGADETECT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MEGA_OK 0
#define MEGA_ERR_INVALID_ARG (-1)  /* null pointer, zero or mismatched size */
#define MEGA_ERR_UNSUPPORTED (-2)  /* modulation order, shape beyond limits */
#define MEGA_ERR_BATCH (-3)        /* batch exceeds the handle's batch_count */
#define MEGA_ERR_HIP (-4)          /* a HIP runtime call failed */

/* Numerical form of the ZF / equalize kernels. */
#define MEGA_FORM_FP32 0   /* plain FP32 arithmetic: the drop-in form */
/* Products on the BF16 matrix cores with FP32 operands split into BF16
 * pieces, FP32 accumulation: the Gram matrix and W = X A on three-piece
 * splits (FP32-accurate products, because the condition number multiplies
 * their errors), the equalize on hi + lo pairs (hi*hi + hi*lo + lo*hi);
 * FP32 sweep inverse with one refinement step whose residual is formed in
 * FP64; groups whose Gram matrix has 1-norm condition number above 1e3 are
 * computed in FP64 throughout. mega_zf and mega_detect_slots only. */
#define MEGA_FORM_MIXED 1

typedef void *mega_stream_t; /* hipStream_t */
typedef struct mega_beamform *mega_beamform_t;

/* Library version string and the calling thread's last error message. */
const char *mega_version(void);
const char *mega_last_error(void);

/* Re-reads the MEGA_* tuning overrides from the environment. The launchers
 * parse them once, at their first use, and keep them; this makes a later
 * change of the environment take effect (tuning sweeps, tests). Values
 * outside a variable's accepted range mean the launcher's own choice. No
 * reference counterpart. Always MEGA_OK. */
int mega_tuning_reload(void);

/* ---- Beamform ------------------------------------------------------------ */

/* Replaces Beamform::Beamform(num_bs, num_ue, batch_count)
 * (src/operators/beamform/beamform.cu:19-31). Allocates the per-group status
 * array on the current device. Users: up to 128 for the FP32 form, up to 64
 * for MEGA_FORM_MIXED (the reference's cuSOLVER path takes any number);
 * beyond that MEGA_ERR_UNSUPPORTED. */
int mega_beamform_create(uint32_t num_bs, uint32_t num_ue, uint32_t batch_count,
                         mega_beamform_t *out);

/* Replaces Beamform::~Beamform (beamform.cu:33-48). NULL is a no-op. */
int mega_beamform_destroy(mega_beamform_t bf);

/* Replaces Beamform::beamform_uplink(const Matrix& csi, stream)
 * (beamform.cu:144-152) and beamformer_uplink (:74-122): in place,
 * W_g = (A_g A_g^H)^-1 A_g for g < batch. batch > batch_count is
 * MEGA_ERR_BATCH (the reference throws std::runtime_error). A group whose
 * Gram matrix is not positive definite gets W_g = 0 and a nonzero status
 * (mega_beamform_info); the reference never reads its potrf info. */
int mega_beamform_uplink(mega_beamform_t bf, void *csi, uint32_t batch,
                         mega_stream_t stream);

/* Replaces Beamform::beamformer(const Matrix& csi, const Matrix& precode,
 * stream) (beamform.cu:124-142): beamform_uplink, then precode_g =
 * W_g^T / max|W_g| (beam_utils.cuh:40-80). */
int mega_beamformer(mega_beamform_t bf, void *csi, void *precode,
                    uint32_t batch, mega_stream_t stream);

/* Replaces static Beamform::beamform_downlink(csi, precode, stream)
 * (beamform.cu:154-168): precode_g = csi_g^T / max|csi_g| only. */
int mega_beamform_downlink(const void *csi, void *precode, uint32_t num_ue,
                           uint32_t num_bs, uint32_t batch,
                           mega_stream_t stream);

/* Per-group status of the last beamform_uplink / beamformer on this handle:
 * 0 = ok, k+1 = Cholesky pivot k was not positive. Synchronises `stream`,
 * copies batch values to host_info and returns the number of failed groups
 * (>= 0) or an error code. No reference counterpart (the reference leaves
 * cuSOLVER's info unread, beamform.cu:110-111). */
int mega_beamform_info(mega_beamform_t bf, int32_t *host_info, uint32_t batch,
                       mega_stream_t stream);

/* Out-of-place ZF without a handle: w_out may equal csi_in. info_dev
 * (optional, device, int32[batch]) receives the per-group status. Used by
 * callers that keep the channel estimate (the beamformer's in-place contract
 * overwrites it). */
int mega_zf(uint32_t num_bs, uint32_t num_ue, uint32_t batch,
            const void *csi_in, void *w_out, int32_t *info_dev, int form,
            mega_stream_t stream);

/* Mixed-form ZF (MEGA_FORM_MIXED) emitting the weights as split BF16 words
 * for mega_equalize_slots with form MEGA_FORM_MIXED. w_split receives, per
 * group, a plane of hi words then a plane of lo words, each U rows of
 * Bp = B rounded up to a multiple of 16 words; a word packs (re, im) as two
 * BF16 (re in the low half), W = hi + lo, zeros past B: batch * U * Bp * 8
 * bytes. w_out (optional) receives W in FP32 (CSI layout), route_dev
 * (optional) 1 for groups computed on the FP64 route. No reference
 * counterpart: the first half of mega_detect_slots with eq.form =
 * MEGA_FORM_MIXED. */
int mega_zf_split(uint32_t num_bs, uint32_t num_ue, uint32_t batch,
                  const void *csi_in, void *w_split, void *w_out,
                  int32_t *info_dev, int32_t *route_dev, mega_stream_t stream);

/* ---- Equalize + demod -------------------------------------------------- */

/* Replaces the raw overload Equalize::equalize(problem_size_row = U,
 * problem_size_col = B, batch_count = N, bCSI, bBsVec, bUeVec, ue_stride,
 * csi_skip = sg, mod_order = Q, cb_size, cb_stride, zeros, valid_bits, stream)
 * (src/operators/sgemv/equalize.cu:20-51): one OFDM symbol,
 * x_k = W_{k/sg} y_k, then max-log demod to binary16 LLRs
 * (demodulation.h:19-67). Q outside {2,4,6,8} is MEGA_ERR_UNSUPPORTED and
 * writes nothing (the reference silently writes nothing). The Matrix overload
 * (equalize.cu:53-67) maps to this with cb_stride = ueVec.stride(1),
 * ue_stride = cb_stride * cb_blocks, valid_bits = cb_size * cb_blocks. */
int mega_equalize(uint64_t row, uint64_t col, uint64_t batch_count,
                  const void *csi, const void *bsvec, void *uevec,
                  uint64_t ue_stride, uint64_t csi_skip, uint8_t mod_order,
                  uint64_t cb_size, uint64_t cb_stride, uint64_t zeros,
                  uint64_t valid_bits, mega_stream_t stream);

/* Many OFDM symbols of many slots in one launch: W of a group is staged once
 * and applied to every symbol. Element strides (complex elements for csi / y /
 * xhat, halves for llr):
 *   csi  slot s at csi + s*csi_slot_stride            ([G][B][U] each)
 *   y    symbol t of slot s at y + s*y_slot_stride + t*y_sym_stride ([B][N])
 *   llr  at llr + s*llr_slot_stride + t*llr_sym_stride ([U][ue_stride])
 *   xhat optional debug output, complex64 [slots][syms][U][N] (NULL = none).
 * Extends equalize.cu:20-51, which handles one symbol per launch.
 * With form = MEGA_FORM_MIXED, csi holds the split planes of mega_zf_split
 * (csi_slot_stride in 8-byte units: G * U * Bp for consecutive slots) and
 * the products run on the BF16 matrix cores with split operands. */
typedef struct {
  uint32_t num_ue, num_bs, num_sc, sc_group; /* U, B, N, sg */
  uint32_t num_syms, num_slots;
  uint32_t mod_order;                        /* Q */
  int32_t form;                              /* MEGA_FORM_* */
  uint64_t ue_stride, cb_size, cb_stride, zeros, valid_bits;
  const void *csi;
  uint64_t csi_slot_stride;
  const void *y;
  uint64_t y_sym_stride, y_slot_stride;
  void *llr;
  uint64_t llr_sym_stride, llr_slot_stride;
  void *xhat;
} mega_equalize_desc;

int mega_equalize_slots(const mega_equalize_desc *desc, mega_stream_t stream);

/* ---- Uplink FFT stage: received samples -> y and CSI ------------------------
 * The reference's UplinkFFT<QuanT> and PilotFFT<QuanT> operators
 * (src/operators/fft/fft_op.h:52-113, fft_op.cu:143-172: a load kernel,
 * cuFFT C2C forward and a store kernel) as one fused kernel. For symbol s
 * and antenna b, x[n] = (re, im) / (max + 1) of the integer pairs at
 * in + s*in_sym_stride + b*ofdm_ca (quant_bits 16: int16 as `short`, 8:
 * int8 as `char`), X = forward DFT of size ofdm_ca (unnormalised; a power
 * of two from 16 to 4096, else MEGA_ERR_UNSUPPORTED), and bin
 * col(i) = (i + ofdm_start + ofdm_ca/2) % ofdm_ca is subcarrier i < ofdm_num.
 * The constructor arguments of the reference become descriptor fields. */
typedef struct {
  uint32_t ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms;
  int32_t quant_bits;     /* 16 or 8 */
  const void *in;         /* [S][B][ofdm_ca] (re, im) integer pairs */
  uint64_t in_sym_stride; /* samples between symbols (0 = bs_num * ofdm_ca) */
  void *out;              /* see each call */
  uint64_t out_sym_stride; /* complex elements between symbols (0 = dense) */
} mega_fft_desc;

/* UplinkFFT::operator() (fft_op.cu:143-150, fft_store_uplink_kernel
 * fft_callback.cuh:43-74): out = y, [S][B][ofdm_num] complex64,
 * y[s][b][i] = X[col(i)] -- the bsVec layout mega_equalize reads. */
int mega_uplink_fft(const mega_fft_desc *desc, mega_stream_t stream);

/* PilotFFT::operator()(in, out, ue_start) (fft_op.cu:152-172,
 * fft_store_pilot_kernel fft_callback.cuh:76-108): pilot subcarrier i of
 * antenna b in pilot symbol s gives the CSI element of user
 * ue_start + s * sc_group + i % sc_group in group i / sc_group (s = 0 for
 * one symbol, the reference's case; num_syms > 1 handles the pilot symbols
 * of a slot in one launch, each carrying the next sc_group users, all into
 * the same CSI unless out_sym_stride is set), conj(X[col(i)]) * pilot_sign[i]
 * computed as
 * {e.re p.re + e.im p.im, e.re p.im - e.im p.re}, written to out =
 * [G][B][ue_num] complex64 (G = ceil(ofdm_num / sc_group); A = H^H, the
 * layout mega_zf / mega_beamform_uplink read); users >= ue_num are skipped
 * and other entries left as they are. pilot_sign: device, ofdm_num complex64
 * (mega_pilot_sign's table). */
int mega_pilot_fft(const mega_fft_desc *desc, uint32_t ue_num, uint32_t sc_group,
                   uint32_t ue_start, const void *pilot_sign, mega_stream_t stream);

/* DownlinkIFFT::operator() (fft_op.cu:168-174, ifft_store_downlink_kernel
 * ifft_callback.cuh:21-59): in = complex64 rows [S][B][ofdm_ca] (the
 * precode output, in_sym_stride in complex elements), out = integer
 * (re, im) pairs [S][B][ofdm_ca] (out_sym_stride in pairs): the unnormalised
 * inverse DFT, each part times (max + 1) / ofdm_ca, rounded to nearest even
 * and saturated. ofdm_start and ofdm_num are not used. */
int mega_downlink_ifft(const mega_fft_desc *desc, mega_stream_t stream);

/* PilotSign::init's table (fft_op.cu:176-194), on the host: the
 * Zadoff-Chu sequence of ofdm_num (utils.h:54-75) times the pi/4 cyclic
 * shift (utils.h:77-95), each element divided by its squared magnitude;
 * host_out receives ofdm_num (re, im) float pairs. Synchronous, no GPU. */
int mega_pilot_sign(uint32_t ofdm_num, float *host_out);

/* ---- Uplink data path: samples -> FFT -> equalize + demod -------------------
 * One slot's uplink symbols from received samples to LLRs through y buffers
 * the handle owns. By default (chunk_syms 0, or >= num_syms) one FFT launch
 * (mega_uplink_fft's kernel) and one equalize + demod launch
 * (mega_equalize_slots' kernels) on `stream`, through one slot of y reused
 * every run (it stays in the Infinity Cache between the two). chunk_syms =
 * K < S runs the symbols in chunks of K through a ring of two chunk buffers:
 * the FFT of chunk c on `stream`, the equalize of chunk c on the handle's
 * own stream, the FFT of chunk c + 2 after the equalize of chunk c, so one
 * chunk's FFT runs beside the previous chunk's equalize and y needs only
 * 2 K B N complex (slower at 256 x 32: both kernels already fill the GPU;
 * MEGA_UPLINK_STREAMS=1 keeps the chunks on `stream`). The reference runs
 * UplinkFFT, Equalize::equalize and the decoder per symbol with a y buffer
 * per symbol (src/baseband/process.cc:175-186). fft: ofdm_start, ofdm_ca,
 * ofdm_num (= shape->num_sc), bs_num (= num_bs), num_syms (= num_syms),
 * quant_bits, in_sym_stride (its buffers are ignored); shape: as
 * mega_equalize_slots for one slot (num_slots = 1; form FP32 with W in the
 * CSI layout, or MIXED with mega_zf_split's planes), its buffers and y
 * strides ignored. mega_uplink_run(samples = [S][B][ofdm_ca] integer pairs,
 * w, llr, xhat optional [S][U][N]) is asynchronous on `stream`, which ends
 * after the last equalize; LLRs and x_hat equal mega_uplink_fft of the slot
 * followed by mega_equalize_slots, bit for bit in the FP32 form. One run at
 * a time per handle (runs on one stream follow each other). */
typedef struct mega_uplink *mega_uplink_t;
int mega_uplink_create(const mega_fft_desc *fft, const mega_equalize_desc *shape,
                       uint32_t chunk_syms, mega_uplink_t *out);
int mega_uplink_run(mega_uplink_t uplink, const void *samples, const void *w, void *llr,
                    void *xhat, mega_stream_t stream);
int mega_uplink_destroy(mega_uplink_t uplink);

/* ---- Downlink precode ----------------------------------------------------- */

/* Replaces the raw overload Precode::precode(problem_size_row = B,
 * problem_size_col = U, batch_count = N, bPrecoded, bUeVec, bBsVec,
 * bs_stride, precode_skip = sg, bs_start, stream)
 * (src/operators/sgemv/precode.cu:20-57, precode.h:38-42): one OFDM symbol,
 *   bsvec[b][(k + bs_start + bs_stride/2) % bs_stride] = sum_u P_g[b][u] ue[u][k]
 * for g = k / sg, with P [G][U][B] (a column-major B x U precoder per group,
 * e.g. mega_beamformer's output), ue [U][N] (user rows, subcarrier fastest)
 * and bsvec [B][bs_stride]. As in the reference only the N / sg full groups
 * are computed; the other entries of bsvec are left as they are. The Matrix
 * overload (precode.cu:59-69) maps to this with B = precoded.dim(0),
 * U = precoded.dim(1), N = ueVec.dim(0), bs_stride = bsVec.stride(1).
 * num_ue > 64 is MEGA_ERR_UNSUPPORTED. */
int mega_precode(uint64_t row, uint64_t col, uint64_t batch_count,
                 const void *precode, const void *uevec, void *bsvec,
                 uint64_t bs_stride, uint64_t precode_skip, uint64_t bs_start,
                 mega_stream_t stream);

/* Many OFDM symbols in one launch: symbol t reads uevec + t*ue_sym_stride
 * and writes bsvec + t*bs_sym_stride (complex elements); each workgroup's
 * precoder block is read once for its share of the symbols. Extends
 * precode.cu:20-57, which handles one symbol per launch. */
typedef struct {
  uint32_t num_bs, num_ue, num_sc, sc_group, num_syms; /* B, U, N, sg, S */
  uint64_t bs_stride, bs_start;
  const void *precode;
  const void *uevec;
  uint64_t ue_sym_stride;
  void *bsvec;
  uint64_t bs_sym_stride;
} mega_precode_desc;

int mega_precode_syms(const mega_precode_desc *desc, mega_stream_t stream);

/* ---- Fused detection ------------------------------------------------------
 * ZF weights and equalize + demod of every symbol of every slot in one call:
 * eq.csi holds the channel estimates A_g = H_g^H (not W). For U <= 16 and
 * B <= 256 one kernel does both per (slot, group) and W never touches HBM
 * unless w_out is given; other shapes run the ZF and equalize kernels back to
 * back through w_out, which is then required. Results equal
 * mega_zf(eq.csi -> w_out) followed by mega_equalize_slots(w_out, ...).
 * No single reference counterpart: it fuses Beamform::beamform_uplink
 * (beamform.cu:144-152) with the S Equalize::equalize calls of a slot
 * (process.cc:172-185). With eq.form = MEGA_FORM_MIXED the mixed ZF kernel
 * writes split BF16 weights into `work` (and FP32 W into w_out if given)
 * and the mixed equalize kernel reads them, for any shape. */
typedef struct {
  mega_equalize_desc eq;
  void *w_out;       /* optional: W, same layout and strides as eq.csi */
  int32_t *info_dev; /* optional: device int32[num_slots * G] group status */
  /* MEGA_FORM_MIXED only: */
  void *work;         /* device scratch of mega_detect_work_bytes(desc) bytes */
  int32_t *route_dev; /* optional: device int32[num_slots * G], 1 = FP64 route */
} mega_detect_desc;

int mega_detect_slots(const mega_detect_desc *desc, mega_stream_t stream);

/* Bytes of device scratch mega_detect_slots needs for desc (0 for
 * MEGA_FORM_FP32; num_slots * G * U * Bp * 8 for MEGA_FORM_MIXED, Bp = B
 * rounded up to a multiple of 16). */
uint64_t mega_detect_work_bytes(const mega_detect_desc *desc);

/* mega_detect_slots that allocates nothing and takes no work buffer: the
 * mixed form's split weight planes go to a static device array of the
 * library, MEGA_DETECT_SCRATCH_BYTES per GPU, held until the process ends
 * (its address is looked up at the first call on each device). desc->work
 * must be NULL; a desc whose mega_detect_work_bytes exceeds the maximum is
 * MEGA_ERR_UNSUPPORTED and launches nothing. The FP32 form needs no scratch
 * and runs as mega_detect_slots. Calls that overlap on one GPU share the
 * array and race: keep them on one stream. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). */
#define MEGA_DETECT_SCRATCH_BYTES (32ull << 20)
int mega_detect_slots_noalloc(const mega_detect_desc *desc, mega_stream_t stream);

/* ---- Batching form: many uneven cells, one graph launch ----------------------
 * One cell of a batch: its own antenna x user size, subcarriers, group size,
 * symbols, modulation order and LLR layout (as mega_equalize_slots for one
 * slot). csi holds the channel estimates A_g = H_g^H ([G][B][U] complex64);
 * w receives W ([G][B][U], required: the ZF output the equalize reads);
 * info_dev (optional, int32[G]) the group status; xhat (optional) the
 * equalized symbols ([S][U][N] complex64). No reference counterpart: the
 * reference runs one Beamform + S Equalize calls per cell
 * (process.cc:172-185); this runs all cells of a batch in one launch. */
typedef struct {
  uint32_t num_ue, num_bs, num_sc, sc_group, num_syms, mod_order;
  uint64_t ue_stride, cb_size, cb_stride, zeros, valid_bits;
  const void *csi;
  void *w;
  const void *y;
  uint64_t y_sym_stride;  /* complex elements between symbols ([B][N] each) */
  void *llr;
  uint64_t llr_sym_stride; /* halves between symbols ([U][ue_stride] each) */
  void *xhat;
  int32_t *info_dev;
} mega_cell_desc;

typedef struct mega_batch *mega_batch_t;

/* Plans a batch: cells are grouped into kernel classes (ZF by users rounded
 * to 16-user tiles; equalize by antenna count), each equalize class's tile
 * plan and the launch schedule are chosen by timing the candidates on these
 * buffers (the autotuner: runs the batch several times, so outputs are
 * overwritten), and the chosen launches are captured into one hipGraph.
 * tune_path (optional) caches tuned plans across processes (a text file,
 * appended). Buffers are bound: run the handle on the same buffers. */
int mega_batch_create(const mega_cell_desc *cells, uint32_t num_cells,
                      const char *tune_path, mega_batch_t *out);
/* ZF + equalize + demod of every cell: one hipGraphLaunch on stream. */
int mega_batch_run(mega_batch_t batch, mega_stream_t stream);
/* Human-readable plan (classes, chosen tile plans, graph time). */
const char *mega_batch_plan(mega_batch_t batch);
int mega_batch_destroy(mega_batch_t batch);

/* ---- Slot pipeline: ZF of the next slot beside the current equalize ---------
 * Consecutive slots of one cell, one call per slot. push(slot i) issues the
 * ZF of slot i on the pipeline's own stream, then the equalize + demod of
 * slot i-1 on `stream` (bound at create), so the two kernels run together.
 * Weights live in `depth` (>= 2) device buffers the pipeline owns; every
 * dependency is an event: the equalize of a slot waits for its ZF, and a
 * ZF waits for the equalize that last read its weight buffer. csi is read on
 * the ZF stream (mega_pipeline_zf_stream): produce it there, or make that
 * stream wait for its producer; y is read on `stream`. Results equal
 * mega_zf (FP32 form) or
 * mega_zf_split (mixed form) followed by mega_equalize_slots of each slot,
 * bit for bit. No single reference counterpart: it schedules the
 * Beamform::beamform_uplink + Equalize::equalize calls of consecutive frames
 * (process.cc:172-185, gpu_engine.cc:289-343). `shape` gives the cell:
 * num_ue .. valid_bits, form (MEGA_FORM_FP32 or MEGA_FORM_MIXED), the y /
 * llr symbol strides and num_slots, the slots per push (>= 1): a push takes
 * that many consecutive slots (csi dense, [slots][G][B][U]; y and llr at
 * y_slot_stride / llr_slot_stride, 0 = dense; xhat dense) and runs their ZF
 * as one launch and their equalize as one launch (more groups per launch:
 * at 256x32 two slots per push cost 22% less per slot than one). Its buffer
 * pointers and csi_slot_stride are ignored. A push's csi, y, llr and xhat
 * (optional) must stay valid until its
 * equalize has run on `stream` (issued by the next push or by flush). One
 * thread per handle, on the device current at create; destroy waits for the
 * pipeline's work. Group failures are not reported (use mega_zf for that). */
typedef struct mega_pipeline *mega_pipeline_t;

int mega_pipeline_create(const mega_equalize_desc *shape, uint32_t depth,
                         mega_stream_t stream, mega_pipeline_t *out);
int mega_pipeline_push(mega_pipeline_t pipe, const void *csi, const void *y,
                       void *llr, void *xhat);
/* Issues the equalize of the last pushed slot. */
int mega_pipeline_flush(mega_pipeline_t pipe);
/* The pipeline's ZF stream, where csi is read (to order its producer). */
mega_stream_t mega_pipeline_zf_stream(mega_pipeline_t pipe);
int mega_pipeline_destroy(mega_pipeline_t pipe);

/* n consecutive slots of one cell with their buffers bound (csi[i], y[i],
 * llr[i], and xhat[i] if xhat is not NULL; each holds shape->num_slots
 * slots), scheduled as n pushes and a flush of the slot pipeline above and
 * captured once into a graph (weights
 * in three buffers the handle owns). mega_pipeline_graph_launch runs all n
 * slots with one launch on `stream`, after the work issued there before it;
 * results equal those of the pipeline (and of the sequential calls) bit for
 * bit. The host issues nothing per slot, so the two streams' kernels follow
 * each other without host gaps. Launches of one handle must not overlap
 * (they share the weight buffers); destroy waits for the last launch. No
 * single reference counterpart (see the slot pipeline). */
typedef struct mega_pipeline_graph *mega_pipeline_graph_t;

int mega_pipeline_graph_create(const mega_equalize_desc *shape, uint32_t n,
                               const void *const *csi, const void *const *y,
                               void *const *llr, void *const *xhat,
                               mega_pipeline_graph_t *out);
int mega_pipeline_graph_launch(mega_pipeline_graph_t graph, mega_stream_t stream);
int mega_pipeline_graph_destroy(mega_pipeline_graph_t graph);

/* ---- Learned form: network warm start for the ZF inverse --------------------
 * Building blocks of the learned ZF of a slot sequence (no reference
 * counterpart; the reference recomputes every slot by Cholesky). Matrices
 * are row-major [batch][U][U] complex64; U <= 64. */

/* Gram matrix G_g = A_g A_g^H and its inverse X_g by the plain FP32 sweep
 * (the path mega_zf takes); G_out optional, X_out NULL = Gram only; a
 * singular group gets X = 0 and status k+1 in info_dev (optional). */
int mega_gram_inv(uint32_t num_bs, uint32_t num_ue, uint32_t batch, const void *csi,
                  void *G_out, void *X_out, int32_t *info_dev, mega_stream_t stream);

/* Network input rows (BF16, 2*U*(U+1) per group): the upper triangles of
 * (G - G_prev)/s and X_prev * s as (re, im) pairs, s = tr(G_prev)/U (also
 * written to scale, float[batch]). */
int mega_learned_features(uint32_t num_ue, uint32_t batch, const void *G, const void *G_prev,
                          const void *X_prev, void *feat_bf16, float *scale,
                          mega_stream_t stream);

/* Training target of the network: upper((X_true - X_base) * s) as (re, im)
 * pairs, FP32, U*(U+1) per group (X_base: the refine result with no
 * correction and no iterations). */
int mega_learned_target(uint32_t num_ue, uint32_t batch, const void *X_true, const void *X_base,
                        const float *scale, float *target, mega_stream_t stream);

/* X0 = X_prev - X_prev (G - G_prev) X_prev + C/s (C: the network's upper
 * triangle, corr_ld floats per group; NULL = no correction), then ns_iters
 * Newton-Schulz steps X <- X (2I - G X); X_out the result, res_out[g] =
 * |X G - I|_F, flag_out[g] = 1 where it exceeds bound (or is not finite). */
int mega_learned_refine(uint32_t num_ue, uint32_t batch, const void *G, const void *G_prev,
                        const void *X_prev, const float *corr, uint32_t corr_ld,
                        const float *scale, void *X_out, float *res_out, int32_t *flag_out,
                        float bound, uint32_t ns_iters, mega_stream_t stream);

/* W = X A per group (CSI layout), where flagged groups first replace X by
 * the plain sweep inverse of their Gram matrix (the fallback); X is updated
 * in place. G (optional, row-major [batch][U][U]) is that Gram matrix as
 * mega_gram_inv computed it for this slot; NULL recomputes it from csi.
 * info_dev (optional) gets the fallback's status. */
int mega_learned_finish(uint32_t num_bs, uint32_t num_ue, uint32_t batch, const void *csi,
                        const void *G, void *X, const int32_t *flag, void *w_out,
                        int32_t *info_dev, mega_stream_t stream);

/* Network kernels (BF16 operands, FP32 accumulation): C (+)= A B + bias,
 * activation (0 identity, 1 GELU tanh form), A(i,k) = A[i*sai + k*sak],
 * B(k,j) = B[k*sbk + j*sbj]; C row-major (ldc) in FP32 (Cf, pre-activation,
 * Cf = z + beta*Cf) and/or BF16 (Cb, after the activation). Small M x N
 * grids split K: the partial sums go through a library-owned workspace per
 * (device, stream), allocated on first need and grown with hipMalloc, so
 * GEMMs on different streams may overlap; such a split GEMM is refused
 * (MEGA_ERR_HIP) while its stream is being captured into a graph. */
typedef struct {
  const void *A;
  int64_t sai, sak;
  const void *B;
  int64_t sbk, sbj;
  uint32_t M, N, K;
  const float *bias;
  int32_t act;
  float *Cf;
  void *Cb;
  uint32_t ldc;
  float beta;
} mega_mlp_gemm_desc;

int mega_mlp_gemm(const mega_mlp_gemm_desc *d, mega_stream_t stream);
/* dZ = dY * act'(Z) (BF16 dZ, optional FP32 copy dZf). */
int mega_mlp_act_backward(const float *dY, const float *Z, void *dZ_bf16, float *dZf,
                          uint64_t n, int32_t act, mega_stream_t stream);
/* db[j] = sum_i dZ[i][j] over M rows of N. */
int mega_mlp_col_sum(const float *dZ, float *db, uint32_t M, uint32_t N, mega_stream_t stream);
/* dY = (Y - T) / M (gradient of the mean over M rows of 0.5 |y - t|^2). */
int mega_mlp_mse_grad(const float *Y, const float *T, float *dY, uint64_t n, uint32_t M,
                      mega_stream_t stream);
/* One Adam step (bias-corrected, step >= 1) on FP32 weights w with moments
 * m, v; wb (optional) receives the BF16 copy. */
int mega_mlp_adam(float *w, void *wb, const float *g, float *m, float *v, uint64_t n, float lr,
                  float beta1, float beta2, float eps, uint32_t step, mega_stream_t stream);
int mega_mlp_to_bf16(const float *x, void *y, uint64_t n, mega_stream_t stream);

/* ---- LDPC consumer: decode the equalizer's code blocks ---------------------
 * Replaces LDPCDecoder(llr_type, LDPCConfig(mb, bg, Zc, max_iter)) and its
 * operator()(tLLRs, tOut, stream) (src/operators/ldpc/decoder.h:40-86,
 * config.h:27-73), which hand the decoding to cuPHY. Per code block:
 * (Kb + mb) * Zc LLRs (Kb = 22 for base graph 1, 10 for base graph 2) with
 * the 2 * Zc punctured systematic bits at zero, LLR > 0 meaning bit 0 -- the
 * rows mega_equalize writes with zeros = 2 * Zc and cb_size = (Kb - 2 + mb) *
 * Zc; out: the Kb * Zc information bits, 32 per word, LSB first. Decoding is
 * layered normalized min-sum over the base graph of 3GPP TS 38.212 for the
 * lifting size's set, max_iterations passes over the mb block rows (no
 * early stop), check messages scaled by `normalization` (0 selects 0.8125).
 * cuPHY's own schedule and scale are not public, so the decoded bits equal
 * the reference's on codewords that both decode, not bit for bit in
 * general. */
typedef struct {
  uint32_t base_graph;     /* 1 or 2 */
  uint32_t lifting_size;   /* Zc, one of TS 38.212 Table 5.3.2-1 (2 .. 384) */
  uint32_t parity_nodes;   /* mb: 4 .. 46 (base graph 1) or 4 .. 42 (2) */
  uint32_t max_iterations; /* 1 .. 100 */
  float normalization;     /* in (0, 1]; 0 = 0.8125 */
  int32_t llr_bits;        /* 16: binary16 LLRs, 32: float */
  const void *llr;
  uint64_t llr_stride;     /* elements between code blocks, >= (Kb + mb) * Zc */
  uint32_t num_cb;
  uint32_t *out;
  uint64_t out_stride;     /* 32-bit words between code blocks, >= ceil(Kb * Zc / 32) */
} mega_ldpc_desc;

int mega_ldpc_decode(const mega_ldpc_desc *desc, mega_stream_t stream);

/* ---- host demodulation ----------------------------------------------------
 * Replaces the host side of demodQPSK / demod16QAM / demod64QAM /
 * demod256QAM (demodulation.h:30-67, __device__ __host__): n complex64
 * values -> n*Q binary16 LLRs (uint16 bit patterns), bit-identical to the
 * device path. */
int mega_demod_host(const float *xhat_c64, uint64_t n, uint8_t mod_order,
                    uint16_t *llr_f16);

#ifdef __cplusplus
}
#endif

#endif /* MEGADETECT_H */
