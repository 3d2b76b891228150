/*
 * oracle.h — CPU restatement of MegaStation's uplink detection path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * library (libmegadetect) never links, loads or calls it.
 *
 * Conventions (reference: src/operators/mega_complex.h:25-27,
 * src/operators/matrix/matrix.h:95-117, src/baseband/buffers.cc:77-88):
 *   complex values are interleaved {float re, float im};
 *   CSI / W buffer:  [G][B][U], element (u,b,g) at u + U*b + U*B*g
 *                    (one column-major U x B matrix per subcarrier group);
 *   y (one symbol):  [B][N], element (b,k) at k + N*b;
 *   LLR (one symbol): binary16 bit patterns, one row of ue_stride halves per
 *                    user.
 *
 * Every function names the reference file:line it restates. binary16 values are
 * carried as uint16_t bit patterns because gcc on x86-64 has no _Float16.
 */
#ifndef MEGA_ORACLE_H
#define MEGA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- IEEE binary16, round to nearest even ---- */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
uint16_t orc_f16_sub(uint16_t a, uint16_t b);

/* ---- max-log demodulation (src/operators/modulation/demodulation.h:19-67) ----
 * Writes Q binary16 LLRs for one equalized symbol. Returns 0, or -1 for an
 * unsupported Q. */
int orc_demod(float re, float im, int Q, uint16_t *llr);

/* ---- equalize + demod, naive specification
 * (src/operators/sgemv/equalize_kernel.cuh:415-480, equalize.cu:20-67).
 * One OFDM symbol. x_hat[u,k] = sum_b W_g[u,b] * y[b,k], g = k / sg, in FP32
 * with the operator order of mega::Complex (mega_complex.h:35-49, no FMA);
 * then demod and store bit j of subcarrier k at
 *   u*ue_stride + (kQ / cb_size)*cb_stride + zeros + (kQ % cb_size) + j
 * when kQ + j < valid_bits and that offset is below ue_stride.
 * The last group may be partial (N % sg != 0): it uses W_{N/sg}.
 * xhat (optional, may be NULL): [U][N] complex64 equalized symbols.
 * Returns 0, or -1 for an unsupported Q. */
int orc_equalize_f32(uint64_t U, uint64_t B, uint64_t N, const float *csi,
                     const float *y, uint16_t *llr, uint64_t ue_stride,
                     uint64_t sg, int Q, uint64_t cb_size, uint64_t cb_stride,
                     uint64_t zeros, uint64_t valid_bits, float *xhat);

/* Same in FP64 accumulation (the EVM reference); x_hat is rounded to float before
 * demod. xhat64 (optional): [U][N] complex128. */
int orc_equalize_f64(uint64_t U, uint64_t B, uint64_t N, const float *csi,
                     const float *y, uint16_t *llr, uint64_t ue_stride,
                     uint64_t sg, int Q, uint64_t cb_size, uint64_t cb_stride,
                     uint64_t zeros, uint64_t valid_bits, double *xhat64);

/* ---- zero-forcing weights (src/operators/beamform/beamform.cu:74-122) ----
 * For each group: Gram = A A^H, lower Cholesky, W = Gram^-1 A by forward and
 * backward substitution. in and out may alias (the reference works in place).
 * info[g] = 0 on success, else k+1 for the first pivot (squared Cholesky
 * diagonal) not above U * 2^-23 * Gram[k][k] (numerically singular in FP32);
 * the group's W is then all zeros. Returns the number of failed groups. */
int orc_zf_f64(uint32_t U, uint32_t B, uint32_t batch, const float *in,
               float *out, double *out64, int32_t *info);
int orc_zf_f32(uint32_t U, uint32_t B, uint32_t batch, const float *in,
               float *out, int32_t *info);

/* ---- downlink precoder normalisation (src/operators/beamform/
 * beam_utils.cuh:40-80): P_g = W_g^T / max|W_g|, out [G][U][B] (B x U
 * column-major per group). */
void orc_scale_absmax(uint32_t U, uint32_t B, uint32_t batch, const float *in,
                      float *out);

/* Downlink precode of one OFDM symbol (precode.cu:20-57): out[b][(k +
 * bs_start + bs_stride/2) % bs_stride] = sum_u P_{k/sg}[b][u] ue[u][k] for
 * the N/sg full groups; other output entries untouched. */
void orc_precode(uint32_t B, uint32_t U, uint32_t N, uint32_t sg, const float *prec,
                 const float *ue, float *out, uint64_t bs_stride, uint64_t bs_start);

/* ---- CPU baseline: the same operator math fused per subcarrier group,
 * OpenMP over groups (FP32 Gram -> Cholesky -> W, then S symbols of
 * equalize + demod). Slots are laid out back to back:
 *   csi  [slots][G][B][U], y [slots][S][B][N], llr [slots][S][U][ue_stride].
 * n_threads <= 0 uses omp_get_max_threads(). Returns the thread count used,
 * or -1 on a bad argument. */
int orc_detect_cpu(uint32_t slots, uint32_t S, uint32_t U, uint32_t B,
                   uint32_t N, uint32_t sg, int Q, const float *csi,
                   const float *y, uint16_t *llr, uint64_t ue_stride,
                   uint64_t cb_size, uint64_t cb_stride, uint64_t zeros,
                   uint64_t valid_bits, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
