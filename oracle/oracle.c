/*
 * oracle.c — CPU restatement of MegaStation's uplink detection path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product library and the
 * CPU baseline of the benchmark. Never linked into libmegadetect.
 *
 * Built with -ffp-contract=off so that every FP32 expression rounds exactly as
 * written (the reference's mega::Complex operators, mega_complex.h:35-49, and
 * demodulation.h:30-67 are restated operation by operation).
 *
 * Parity pins (see tests/test_oracle.py):
 *   orc_f32_to_f16 / orc_f16_sub  - numpy float16 on all 2^16 patterns and
 *                                    random floats (IEEE RN-even)
 *   orc_demod                      - reference fixture test/sgemv/equalize.data
 *                                    (QPSK, bit-exact), 64-QAM known-answer
 *                                    round trip (test/modulation/
 *                                    test_modulation.cc:11-57), closed-form
 *                                    special cases for 16/64/256-QAM
 *   orc_equalize_f32 / _f64        - reference fixture (bit-exact), brute force
 *                                    against numpy complex128 products
 *   orc_zf_f64 / _f32              - reference fixture test/beamform/csi_out.data
 *                                    (1e-5), numpy.linalg.pinv, W A^H = I,
 *                                    ZF(ZF(A)) = A
 *   orc_scale_absmax               - reference fixture test/beamform/precoder.data
 *   orc_precode                    - reference fixture test/sgemv/precoded.data
 *                                    (1e-5), numpy complex128 products with the
 *                                    shifted placement, identity precoder,
 *                                    untouched partial group and guard columns
 *   orc_detect_cpu                 - equal, bit for bit, to orc_zf_f32 followed by
 *                                    orc_equalize_f32
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* binary16                                                                  */
/* ------------------------------------------------------------------------ */

/* Round an exactly representable double to binary16, nearest-even. */
static uint16_t d_to_f16(double v) {
  uint16_t sign = signbit(v) ? 0x8000u : 0u;
  double a = fabs(v);
  if (isnan(v)) return (uint16_t)(sign | 0x7e00u);
  if (isinf(v)) return (uint16_t)(sign | 0x7c00u);
  if (a < 0x1p-14) {
    /* subnormal range: units of 2^-24; q == 1024 encodes the smallest
     * normal, which is the correct carry */
    double q = nearbyint(a * 0x1p24);
    return (uint16_t)(sign | (uint16_t)q);
  }
  int e;
  double m = frexp(a, &e); /* a = m * 2^e, m in [0.5, 1) */
  e -= 1;                  /* a = (2m) * 2^e, 2m in [1, 2) */
  double q = nearbyint(m * 2.0 * 1024.0); /* 1024..2048 */
  uint32_t bits = (uint32_t)((e + 15) << 10) + (uint32_t)q - 1024u;
  if (e + 15 >= 31 || bits >= 0x7c00u) return (uint16_t)(sign | 0x7c00u);
  return (uint16_t)(sign | bits);
}

uint16_t orc_f32_to_f16(float f) { return d_to_f16((double)f); }

float orc_f16_to_f32(uint16_t h) {
  int sign = (h >> 15) & 1;
  int e = (h >> 10) & 0x1f;
  int m = h & 0x3ff;
  double v;
  if (e == 0x1f) {
    v = m ? NAN : INFINITY;
  } else if (e == 0) {
    v = ldexp((double)m, -24);
  } else {
    v = ldexp((double)(m + 1024), e - 25);
  }
  return (float)(sign ? -v : v);
}

/* a - b of two binary16 values, rounded once: the exact difference of two
 * binary16 values fits in a double's 53-bit significand. */
uint16_t orc_f16_sub(uint16_t a, uint16_t b) {
  return d_to_f16((double)orc_f16_to_f32(a) - (double)orc_f16_to_f32(b));
}

static uint16_t f16_abs(uint16_t h) { return (uint16_t)(h & 0x7fffu); }

/* ------------------------------------------------------------------------ */
/* demod (demodulation.h:19-67)                                               */
/* ------------------------------------------------------------------------ */

int orc_demod(float re, float im, int Q, uint16_t *llr) {
  switch (Q) {
    case 2: { /* demodulation.h:30-33: half(c * -20 * sqrtf(2)) */
      float s = sqrtf(2.f);
      llr[0] = orc_f32_to_f16(re * -20.f * s);
      llr[1] = orc_f32_to_f16(im * -20.f * s);
      return 0;
    }
    case 4: { /* demodulation.h:35-40 */
      uint16_t t = orc_f32_to_f16(200.f / sqrtf(10.f));
      llr[0] = orc_f32_to_f16(100.f * re);
      llr[1] = orc_f32_to_f16(100.f * im);
      llr[2] = orc_f16_sub(t, f16_abs(llr[0]));
      llr[3] = orc_f16_sub(t, f16_abs(llr[1]));
      return 0;
    }
    case 6: { /* demodulation.h:42-52 */
      uint16_t t1 = orc_f32_to_f16(400.f / sqrtf(42.f));
      uint16_t t2 = orc_f32_to_f16(200.f / sqrtf(42.f));
      llr[0] = orc_f32_to_f16(100.f * re);
      llr[1] = orc_f32_to_f16(100.f * im);
      llr[2] = orc_f16_sub(t1, f16_abs(llr[0]));
      llr[3] = orc_f16_sub(t1, f16_abs(llr[1]));
      llr[4] = orc_f16_sub(t2, f16_abs(llr[2]));
      llr[5] = orc_f16_sub(t2, f16_abs(llr[3]));
      return 0;
    }
    case 8: { /* demodulation.h:54-67 */
      uint16_t t1 = orc_f32_to_f16(800.f / sqrtf(170.f));
      uint16_t t2 = orc_f32_to_f16(400.f / sqrtf(170.f));
      uint16_t t3 = orc_f32_to_f16(200.f / sqrtf(170.f));
      llr[0] = orc_f32_to_f16(100.f * re);
      llr[1] = orc_f32_to_f16(100.f * im);
      llr[2] = orc_f16_sub(t1, f16_abs(llr[0]));
      llr[3] = orc_f16_sub(t1, f16_abs(llr[1]));
      llr[4] = orc_f16_sub(t2, f16_abs(llr[2]));
      llr[5] = orc_f16_sub(t2, f16_abs(llr[3]));
      llr[6] = orc_f16_sub(t3, f16_abs(llr[4]));
      llr[7] = orc_f16_sub(t3, f16_abs(llr[5]));
      return 0;
    }
    default:
      return -1;
  }
}

/* ------------------------------------------------------------------------ */
/* LLR placement (equalize_kernel.cuh:461-478, and the tiled kernel's         */
/* row-bound predicate at :286,319)                                           */
/* ------------------------------------------------------------------------ */

static void store_llr(uint16_t *row, uint64_t k, int Q, const uint16_t *v,
                      uint64_t ue_stride, uint64_t cb_size, uint64_t cb_stride,
                      uint64_t zeros, uint64_t valid_bits) {
  uint64_t bit_count = k * (uint64_t)Q;
  uint64_t off = (bit_count / cb_size) * cb_stride + zeros + bit_count % cb_size;
  for (int j = 0; j < Q; ++j) {
    if (bit_count + (uint64_t)j >= valid_bits) break;
    if (off + (uint64_t)j >= ue_stride) break;
    row[off + j] = v[j];
  }
}

static int valid_q(int Q) { return Q == 2 || Q == 4 || Q == 6 || Q == 8; }

/* ------------------------------------------------------------------------ */
/* equalize (equalize_kernel.cuh:415-480)                                     */
/* ------------------------------------------------------------------------ */

int orc_equalize_f32(uint64_t U, uint64_t B, uint64_t N, const float *csi,
                     const float *y, uint16_t *llr, uint64_t ue_stride,
                     uint64_t sg, int Q, uint64_t cb_size, uint64_t cb_stride,
                     uint64_t zeros, uint64_t valid_bits, float *xhat) {
  if (!valid_q(Q) || sg == 0 || cb_size == 0) return -1;
  for (uint64_t k = 0; k < N; ++k) {
    const float *w = csi + 2 * (k / sg) * U * B;
    for (uint64_t u = 0; u < U; ++u) {
      float acc_re = 0.f, acc_im = 0.f;
      for (uint64_t b = 0; b < B; ++b) {
        float a_re = w[2 * (u + U * b)], a_im = w[2 * (u + U * b) + 1];
        float y_re = y[2 * (k + N * b)], y_im = y[2 * (k + N * b) + 1];
        /* Complex::operator* then operator+= (mega_complex.h:35-38,45-49) */
        float p_re = a_re * y_re - a_im * y_im;
        float p_im = a_re * y_im + a_im * y_re;
        acc_re += p_re;
        acc_im += p_im;
      }
      if (xhat) {
        xhat[2 * (u * N + k)] = acc_re;
        xhat[2 * (u * N + k) + 1] = acc_im;
      }
      uint16_t v[8];
      orc_demod(acc_re, acc_im, Q, v);
      store_llr(llr + u * ue_stride, k, Q, v, ue_stride, cb_size, cb_stride,
                zeros, valid_bits);
    }
  }
  return 0;
}

int orc_equalize_f64(uint64_t U, uint64_t B, uint64_t N, const float *csi,
                     const float *y, uint16_t *llr, uint64_t ue_stride,
                     uint64_t sg, int Q, uint64_t cb_size, uint64_t cb_stride,
                     uint64_t zeros, uint64_t valid_bits, double *xhat64) {
  if (!valid_q(Q) || sg == 0 || cb_size == 0) return -1;
  for (uint64_t k = 0; k < N; ++k) {
    const float *w = csi + 2 * (k / sg) * U * B;
    for (uint64_t u = 0; u < U; ++u) {
      double acc_re = 0., acc_im = 0.;
      for (uint64_t b = 0; b < B; ++b) {
        double a_re = w[2 * (u + U * b)], a_im = w[2 * (u + U * b) + 1];
        double y_re = y[2 * (k + N * b)], y_im = y[2 * (k + N * b) + 1];
        acc_re += a_re * y_re - a_im * y_im;
        acc_im += a_re * y_im + a_im * y_re;
      }
      if (xhat64) {
        xhat64[2 * (u * N + k)] = acc_re;
        xhat64[2 * (u * N + k) + 1] = acc_im;
      }
      if (llr) {
        uint16_t v[8];
        orc_demod((float)acc_re, (float)acc_im, Q, v);
        store_llr(llr + u * ue_stride, k, Q, v, ue_stride, cb_size, cb_stride,
                  zeros, valid_bits);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* zero forcing (beamform.cu:74-122: Cgemm(N,C) Gram, potrf LOWER, potrs)     */
/* ------------------------------------------------------------------------ */

/* One group in double. a: U x B column-major complex (float); w: output
 * (double, U x B column-major); L: U*U*2 scratch. Returns info. */
static int zf_group_f64(uint32_t U, uint32_t B, const float *a, double *w,
                        double *L) {
  /* Gram lower triangle: G[i][j] = sum_b a[i,b] conj(a[j,b]) */
  for (uint32_t i = 0; i < U; ++i)
    for (uint32_t j = 0; j <= i; ++j) {
      double re = 0., im = 0.;
      for (uint32_t b = 0; b < B; ++b) {
        double xr = a[2 * (i + U * b)], xi = a[2 * (i + U * b) + 1];
        double yr = a[2 * (j + U * b)], yi = a[2 * (j + U * b) + 1];
        re += xr * yr + xi * yi;
        im += xi * yr - xr * yi;
      }
      L[2 * (i * U + j)] = re;
      L[2 * (i * U + j) + 1] = im;
    }
  /* Cholesky, lower, left-looking. A pivot (Schur-complement diagonal) not
   * above U * 2^-23 * G_kk marks the group numerically singular, the
   * product's FP32 criterion. */
  const double tol = (double)U * 0x1p-23;
  for (uint32_t k = 0; k < U; ++k) {
    const double gkk = L[2 * (k * U + k)];
    double d = gkk;
    for (uint32_t j = 0; j < k; ++j) {
      double r = L[2 * (k * U + j)], m = L[2 * (k * U + j) + 1];
      d -= r * r + m * m;
    }
    if (!(d > tol * gkk) || !isfinite(d)) return (int)k + 1;
    d = sqrt(d);
    L[2 * (k * U + k)] = d;
    L[2 * (k * U + k) + 1] = 0.;
    for (uint32_t i = k + 1; i < U; ++i) {
      double re = L[2 * (i * U + k)], im = L[2 * (i * U + k) + 1];
      for (uint32_t j = 0; j < k; ++j) {
        double xr = L[2 * (i * U + j)], xi = L[2 * (i * U + j) + 1];
        double yr = L[2 * (k * U + j)], yi = L[2 * (k * U + j) + 1];
        re -= xr * yr + xi * yi; /* L[i,j] * conj(L[k,j]) */
        im -= xi * yr - xr * yi;
      }
      L[2 * (i * U + k)] = re / d;
      L[2 * (i * U + k) + 1] = im / d;
    }
  }
  /* per column: z = L^-1 a, w = L^-H z */
  for (uint32_t b = 0; b < B; ++b) {
    double *col = w + 2 * (size_t)U * b;
    for (uint32_t i = 0; i < U; ++i) {
      double re = a[2 * (i + U * b)], im = a[2 * (i + U * b) + 1];
      for (uint32_t j = 0; j < i; ++j) {
        double lr = L[2 * (i * U + j)], li = L[2 * (i * U + j) + 1];
        double zr = col[2 * j], zi = col[2 * j + 1];
        re -= lr * zr - li * zi;
        im -= lr * zi + li * zr;
      }
      double dd = L[2 * (i * U + i)];
      col[2 * i] = re / dd;
      col[2 * i + 1] = im / dd;
    }
    for (uint32_t ii = U; ii-- > 0;) {
      double re = col[2 * ii], im = col[2 * ii + 1];
      for (uint32_t j = ii + 1; j < U; ++j) {
        double lr = L[2 * (j * U + ii)], li = -L[2 * (j * U + ii) + 1];
        double zr = col[2 * j], zi = col[2 * j + 1];
        re -= lr * zr - li * zi;
        im -= lr * zi + li * zr;
      }
      double dd = L[2 * (ii * U + ii)];
      col[2 * ii] = re / dd;
      col[2 * ii + 1] = im / dd;
    }
  }
  return 0;
}

int orc_zf_f64(uint32_t U, uint32_t B, uint32_t batch, const float *in,
               float *out, double *out64, int32_t *info) {
  int failed = 0;
  double *L = (double *)malloc(sizeof(double) * 2 * U * U);
  double *w = (double *)malloc(sizeof(double) * 2 * (size_t)U * B);
  for (uint32_t g = 0; g < batch; ++g) {
    const float *a = in + 2 * (size_t)U * B * g;
    int r = zf_group_f64(U, B, a, w, L);
    if (r) {
      memset(w, 0, sizeof(double) * 2 * (size_t)U * B);
      failed++;
    }
    if (info) info[g] = r;
    for (size_t e = 0; e < 2 * (size_t)U * B; ++e) {
      if (out) out[2 * (size_t)U * B * g + e] = (float)w[e];
      if (out64) out64[2 * (size_t)U * B * g + e] = w[e];
    }
  }
  free(L);
  free(w);
  return failed;
}

/* One group in float, same algorithm. w may alias a only if the caller copies
 * a first (we read a in the Gram and in the solve). */
static int zf_group_f32(uint32_t U, uint32_t B, const float *a, float *w,
                        float *L) {
  for (uint32_t i = 0; i < U; ++i)
    for (uint32_t j = 0; j <= i; ++j) {
      float re = 0.f, im = 0.f;
      for (uint32_t b = 0; b < B; ++b) {
        float xr = a[2 * (i + U * b)], xi = a[2 * (i + U * b) + 1];
        float yr = a[2 * (j + U * b)], yi = a[2 * (j + U * b) + 1];
        re += xr * yr + xi * yi;
        im += xi * yr - xr * yi;
      }
      L[2 * (i * U + j)] = re;
      L[2 * (i * U + j) + 1] = im;
    }
  const float tol = (float)U * 0x1p-23f;
  for (uint32_t k = 0; k < U; ++k) {
    const float gkk = L[2 * (k * U + k)];
    float d = gkk;
    for (uint32_t j = 0; j < k; ++j) {
      float r = L[2 * (k * U + j)], m = L[2 * (k * U + j) + 1];
      d -= r * r + m * m;
    }
    if (!(d > tol * gkk) || !isfinite(d)) return (int)k + 1;
    d = sqrtf(d);
    L[2 * (k * U + k)] = d;
    L[2 * (k * U + k) + 1] = 0.f;
    for (uint32_t i = k + 1; i < U; ++i) {
      float re = L[2 * (i * U + k)], im = L[2 * (i * U + k) + 1];
      for (uint32_t j = 0; j < k; ++j) {
        float xr = L[2 * (i * U + j)], xi = L[2 * (i * U + j) + 1];
        float yr = L[2 * (k * U + j)], yi = L[2 * (k * U + j) + 1];
        re -= xr * yr + xi * yi;
        im -= xi * yr - xr * yi;
      }
      L[2 * (i * U + k)] = re / d;
      L[2 * (i * U + k) + 1] = im / d;
    }
  }
  for (uint32_t b = 0; b < B; ++b) {
    float *col = w + 2 * (size_t)U * b;
    for (uint32_t i = 0; i < U; ++i) {
      float re = a[2 * (i + U * b)], im = a[2 * (i + U * b) + 1];
      for (uint32_t j = 0; j < i; ++j) {
        float lr = L[2 * (i * U + j)], li = L[2 * (i * U + j) + 1];
        float zr = col[2 * j], zi = col[2 * j + 1];
        re -= lr * zr - li * zi;
        im -= lr * zi + li * zr;
      }
      float dd = L[2 * (i * U + i)];
      col[2 * i] = re / dd;
      col[2 * i + 1] = im / dd;
    }
    for (uint32_t ii = U; ii-- > 0;) {
      float re = col[2 * ii], im = col[2 * ii + 1];
      for (uint32_t j = ii + 1; j < U; ++j) {
        float lr = L[2 * (j * U + ii)], li = -L[2 * (j * U + ii) + 1];
        float zr = col[2 * j], zi = col[2 * j + 1];
        re -= lr * zr - li * zi;
        im -= lr * zi + li * zr;
      }
      float dd = L[2 * (ii * U + ii)];
      col[2 * ii] = re / dd;
      col[2 * ii + 1] = im / dd;
    }
  }
  return 0;
}

int orc_zf_f32(uint32_t U, uint32_t B, uint32_t batch, const float *in,
               float *out, int32_t *info) {
  int failed = 0;
  float *L = (float *)malloc(sizeof(float) * 2 * U * U);
  float *w = (float *)malloc(sizeof(float) * 2 * (size_t)U * B);
  for (uint32_t g = 0; g < batch; ++g) {
    const float *a = in + 2 * (size_t)U * B * g;
    int r = zf_group_f32(U, B, a, w, L);
    if (r) {
      memset(w, 0, sizeof(float) * 2 * (size_t)U * B);
      failed++;
    }
    if (info) info[g] = r;
    memcpy(out + 2 * (size_t)U * B * g, w, sizeof(float) * 2 * (size_t)U * B);
  }
  free(L);
  free(w);
  return failed;
}

/* ------------------------------------------------------------------------ */
/* precoder normalisation (beam_utils.cuh:40-80)                              */
/* ------------------------------------------------------------------------ */

void orc_scale_absmax(uint32_t U, uint32_t B, uint32_t batch, const float *in,
                      float *out) {
  size_t n = (size_t)U * B;
  for (uint32_t g = 0; g < batch; ++g) {
    const float *x = in + 2 * n * g;
    float *o = out + 2 * n * g;
    float mx = 0.f;
    for (size_t e = 0; e < n; ++e) {
      float re = x[2 * e], im = x[2 * e + 1];
      mx = fmaxf(mx, sqrtf(re * re + im * im)); /* Complex::abs, :68-70 */
    }
    float s = 1.0f / mx;
    for (size_t e = 0; e < n; ++e) {
      size_t r = e % U, c = e / U;
      size_t t = r * B + c;
      float re = x[2 * e], im = x[2 * e + 1];
      /* input * Complex(1/max, 0) */
      o[2 * t] = re * s - im * 0.f;
      o[2 * t + 1] = re * 0.f + im * s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* downlink precode (precode_kernel.cuh:418-452, the naive specification;      */
/* layouts of the tiled kernel, :180-187 and :292-300; grid precode.cu:25-28) */
/* ------------------------------------------------------------------------ */

/* out[b][(k + bs_start + bs_stride/2) % bs_stride] = sum_u P_g[b][u] ue[u][k]
 * for g = k / sg and k < (N / sg) * sg: the reference launches N / sg
 * groups (floor), so the subcarriers of a partial last group and the columns
 * no subcarrier maps to are left as they are. P is [G][U][B] (column-major
 * B x U per group), ue [U][N], out [B][bs_stride]; complex64 interleaved.
 * Products and sums in the order of mega::Complex (mega_complex.h:35-49),
 * users in increasing order. */
void orc_precode(uint32_t B, uint32_t U, uint32_t N, uint32_t sg, const float *prec,
                 const float *ue, float *out, uint64_t bs_stride, uint64_t bs_start) {
  const uint32_t G = N / sg;
  for (uint32_t k = 0; k < G * sg; ++k) {
    const float *P = prec + 2 * (size_t)(k / sg) * U * B;
    const uint64_t col = (k + bs_start + bs_stride / 2) % bs_stride;
    for (uint32_t b = 0; b < B; ++b) {
      float re = 0.f, im = 0.f;
      for (uint32_t u = 0; u < U; ++u) {
        const float pr = P[2 * ((size_t)u * B + b)], pi = P[2 * ((size_t)u * B + b) + 1];
        const float xr = ue[2 * ((size_t)u * N + k)], xi = ue[2 * ((size_t)u * N + k) + 1];
        const float tr = pr * xr - pi * xi, ti = pr * xi + pi * xr;
        re = re + tr;
        im = im + ti;
      }
      out[2 * (b * bs_stride + col)] = re;
      out[2 * (b * bs_stride + col) + 1] = im;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* CPU baseline                                                               */
/* ------------------------------------------------------------------------ */

/* Equalize + demod of one group's subcarriers for one symbol, b-outer order so
 * the inner loops stream y rows; each (u,k) still sums b = 0..B-1 in order
 * with the same operations as orc_equalize_f32, so results are identical. */
__attribute__((target_clones("avx2", "default"))) static void group_symbol(
    uint32_t U, uint32_t B, uint32_t N, uint32_t k0, uint32_t nk, int Q,
    const float *w, const float *y, uint16_t *llr, uint64_t ue_stride,
    uint64_t cb_size, uint64_t cb_stride, uint64_t zeros, uint64_t valid_bits,
    float *acc /* 2*U*nk */) {
  memset(acc, 0, sizeof(float) * 2 * U * nk);
  for (uint32_t b = 0; b < B; ++b) {
    const float *yr = y + 2 * ((size_t)N * b + k0);
    const float *wc = w + 2 * (size_t)U * b;
    for (uint32_t kk = 0; kk < nk; ++kk) {
      float y_re = yr[2 * kk], y_im = yr[2 * kk + 1];
      float *ak = acc + 2 * (size_t)U * kk;
      for (uint32_t u = 0; u < U; ++u) {
        float a_re = wc[2 * u], a_im = wc[2 * u + 1];
        float p_re = a_re * y_re - a_im * y_im;
        float p_im = a_re * y_im + a_im * y_re;
        ak[2 * u] += p_re;
        ak[2 * u + 1] += p_im;
      }
    }
  }
  for (uint32_t kk = 0; kk < nk; ++kk)
    for (uint32_t u = 0; u < U; ++u) {
      uint16_t v[8];
      orc_demod(acc[2 * (U * kk + u)], acc[2 * (U * kk + u) + 1], Q, v);
      store_llr(llr + u * ue_stride, k0 + kk, Q, v, ue_stride, cb_size,
                cb_stride, zeros, valid_bits);
    }
}

int orc_detect_cpu(uint32_t slots, uint32_t S, uint32_t U, uint32_t B,
                   uint32_t N, uint32_t sg, int Q, const float *csi,
                   const float *y, uint16_t *llr, uint64_t ue_stride,
                   uint64_t cb_size, uint64_t cb_stride, uint64_t zeros,
                   uint64_t valid_bits, int n_threads) {
  if (!valid_q(Q) || sg == 0 || cb_size == 0 || U == 0 || B == 0) return -1;
  uint32_t G = (N + sg - 1) / sg;
  int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
  long total = (long)slots * G;
#pragma omp parallel num_threads(nt)
  {
    float *w = (float *)malloc(sizeof(float) * 2 * (size_t)U * B);
    float *L = (float *)malloc(sizeof(float) * 2 * U * U);
    float *acc = (float *)malloc(sizeof(float) * 2 * (size_t)U * sg);
#pragma omp for schedule(static)
    for (long t = 0; t < total; ++t) {
      uint32_t s = (uint32_t)(t / G), g = (uint32_t)(t % G);
      const float *a = csi + 2 * (size_t)U * B * ((size_t)s * G + g);
      if (zf_group_f32(U, B, a, w, L))
        memset(w, 0, sizeof(float) * 2 * (size_t)U * B);
      uint32_t k0 = g * sg, nk = (k0 + sg <= N) ? sg : N - k0;
      for (uint32_t sym = 0; sym < S; ++sym) {
        size_t si = (size_t)s * S + sym;
        group_symbol(U, B, N, k0, nk, Q, w, y + 2 * (size_t)B * N * si,
                     llr + (size_t)U * ue_stride * si, ue_stride, cb_size,
                     cb_stride, zeros, valid_bits, acc);
      }
    }
    free(w);
    free(L);
    free(acc);
  }
  return nt;
}
