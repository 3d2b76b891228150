/* CPU oracle of the LDPC consumer -- TEST INFRASTRUCTURE ONLY.
 *
 * Layered normalized min-sum decoding of 5G NR LDPC codewords, the
 * operation the reference's LDPCDecoder hands to cuPHY
 * (src/operators/ldpc/decoder.h:73-86: LLRs of (Kb + mb) * Zc bits per code
 * block with the 2 * Zc punctured prefix at zero, LLR > 0 meaning bit 0;
 * hard decisions of the Kb * Zc information bits out, bit-packed, LSB
 * first). cuPHY's own decoder is closed and absent here, so this states the
 * textbook layered normalized min-sum schedule, in FP32 with contraction
 * off, operation for operation as the HIP kernel does it:
 *
 *   L[v] = llr[v]; R = 0
 *   repeat max_iter times, for block row r = 0 .. mb-1, every z < Zc:
 *     for each block (c, P) of row r, v = c*Zc + (z + P mod Zc) mod Zc:
 *       Q_e = L[v] - R_e                                  (old message)
 *     min1 <= min2 the two smallest |Q_e| (first index of min1 kept, strict <),
 *     s = XOR of the signs (Q_e < 0)
 *     R_e = +-(alpha * (e == argmin ? min2 : min1)), sign s XOR sign(Q_e)
 *     L[v] = Q_e + R_e
 *   bit k = L[k] < 0 for k < Kb * Zc.
 *
 * The messages of a check are kept in compressed form (alpha*min1,
 * alpha*min2, argmin, sign bits), from which R_e is rebuilt exactly.
 *
 * Pinned by (tests/test_oracle_ldpc.py): the reference's BG1 codewords
 * (test/ldpc/encode.data, input.data) satisfy every parity check of the
 * table and decode to their information bits from noiseless LLRs +-127
 * (test_ldpc.cc:155-190); brute force on a tiny code (every codeword of a
 * 2-row base graph found by enumeration decodes to itself, and a single bit
 * error is corrected); codewords of the numpy encoder under AWGN decode
 * with fewer bit errors than hard decisions of the channel LLRs.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORC_LDPC_MAX_DEG 32

int orc_ldpc_decode(int kb, int mb, int zc, const int16_t* rs, const int8_t* col,
                    const int16_t* shift, int max_iter, float alpha, const float* llr,
                    uint64_t llr_stride, int n_cw, uint32_t* out, uint64_t out_stride,
                    float* app_out) {
  const int nv = (kb + mb) * zc;
  const int n_chk = mb * zc;
  float L[(68 + 0) * 384];
  float m1s[46 * 384], m2s[46 * 384];
  uint8_t amin[46 * 384];
  uint32_t sgn[46 * 384];
  if (nv > 68 * 384 || n_chk > 46 * 384 || zc <= 0) return -1;
  for (int r = 0; r < mb; ++r)
    if (rs[r + 1] - rs[r] > ORC_LDPC_MAX_DEG) return -1;
  for (int w = 0; w < n_cw; ++w) {
    const float* in = llr + (uint64_t)w * llr_stride;
    for (int v = 0; v < nv; ++v) L[v] = in[v];
    for (int i = 0; i < n_chk; ++i) {
      m1s[i] = 0.f;
      m2s[i] = 0.f;
      amin[i] = 0;
      sgn[i] = 0u;
    }
    for (int it = 0; it < max_iter; ++it) {
      for (int r = 0; r < mb; ++r) {
        const int e0 = rs[r], deg = rs[r + 1] - rs[r];
        for (int z = 0; z < zc; ++z) {
          const int ci = r * zc + z;
          float q[ORC_LDPC_MAX_DEG];
          int vi[ORC_LDPC_MAX_DEG];
          float min1 = INFINITY, min2 = INFINITY;
          int im = 0;
          uint32_t sb = 0u, sp = 0u;
          for (int e = 0; e < deg; ++e) {
            const int c = col[e0 + e];
            const int p = shift[e0 + e] % zc;
            const int v = c * zc + (z + p) % zc;
            const float mag = e == amin[ci] ? m2s[ci] : m1s[ci];
            const float rold = (sgn[ci] >> e) & 1u ? -mag : mag;
            const float qe = L[v] - rold;
            const float a = fabsf(qe);
            if (a < min1) {
              min2 = min1;
              min1 = a;
              im = e;
            } else if (a < min2) {
              min2 = a;
            }
            const uint32_t s = qe < 0.f ? 1u : 0u;
            sb |= s << e;
            sp ^= s;
            q[e] = qe;
            vi[e] = v;
          }
          const float n1 = alpha * min1, n2 = alpha * min2;
          uint32_t nsg = 0u;
          for (int e = 0; e < deg; ++e) {
            const float mag = e == im ? n2 : n1;
            const uint32_t s = sp ^ ((sb >> e) & 1u);
            const float rnew = s ? -mag : mag;
            nsg |= s << e;
            L[vi[e]] = q[e] + rnew;
          }
          m1s[ci] = n1;
          m2s[ci] = n2;
          amin[ci] = (uint8_t)im;
          sgn[ci] = nsg;
        }
      }
    }
    uint32_t* o = out + (uint64_t)w * out_stride;
    const int nbits = kb * zc;
    for (int k = 0; k < (nbits + 31) / 32; ++k) o[k] = 0u;
    for (int k = 0; k < nbits; ++k)
      if (L[k] < 0.f) o[k / 32] |= 1u << (k % 32);
    if (app_out) memcpy(app_out + (uint64_t)w * nv, L, sizeof(float) * nv);
  }
  return 0;
}
