// Runs every oracle function under AddressSanitizer and UndefinedBehavior-
// Sanitizer -- TEST INFRASTRUCTURE ONLY (built by tests/sanitize_build.py, run
// by tests/test_sanitize.py). Each buffer is a heap allocation of exactly
// the size the function may touch, so a read or write past it is reported;
// the shapes are the edge cases: a partial last group (N % sg != 0), LLR
// rows with code blocks, a punctured prefix and padding, a singular group,
// the smallest and largest LDPC codes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../oracle/oracle.h"
#include "../../megastation_amd/csrc/ldpc_tables.h"

extern "C" int orc_ldpc_decode(int kb, int mb, int zc, const int16_t* rs, const int8_t* col,
                               const int16_t* shift, int max_iter, float alpha, const float* llr,
                               uint64_t llr_stride, int n_cw, uint32_t* out, uint64_t out_stride,
                               float* app_out);

template <class T>
static T* exact(size_t n) {
  T* p = static_cast<T*>(std::malloc(n * sizeof(T)));
  for (size_t i = 0; i < n; ++i) p[i] = T(0);
  return p;
}

static float rnd(unsigned& s) {
  s = s * 1664525u + 1013904223u;
  return ((s >> 8) & 0xffff) / 32768.0f - 1.0f;
}

int main() {
  unsigned seed = 7;
  int fails = 0;
  // demod, every order and a bad one
  {
    uint16_t* llr = exact<uint16_t>(8);
    for (int Q : {2, 4, 6, 8}) fails += orc_demod(0.3f, -0.7f, Q, llr) != 0;
    fails += orc_demod(0.f, 0.f, 5, llr) != -1;
    std::free(llr);
  }
  // equalize: 3 users, 5 antennas, 37 subcarriers in groups of 7 (partial
  // last group), 64-QAM, two code blocks of 100 bits with a 10-half prefix
  // in rows of 128
  {
    const uint64_t U = 3, B = 5, N = 37, sg = 7, G = (N + sg - 1) / sg;
    const uint64_t cb = 100, zeros = 10, cbs = 128, ue_stride = 2 * cbs, valid = 2 * cb;
    float* w = exact<float>(2 * U * B * G);
    float* y = exact<float>(2 * B * N);
    for (uint64_t i = 0; i < 2 * U * B * G; ++i) w[i] = rnd(seed);
    for (uint64_t i = 0; i < 2 * B * N; ++i) y[i] = rnd(seed);
    uint16_t* llr = exact<uint16_t>(U * ue_stride);
    float* xh = exact<float>(2 * U * N);
    double* xh64 = exact<double>(2 * U * N);
    fails += orc_equalize_f32(U, B, N, w, y, llr, ue_stride, sg, 6, cb, cbs, zeros, valid, xh) != 0;
    fails += orc_equalize_f64(U, B, N, w, y, llr, ue_stride, sg, 6, cb, cbs, zeros, valid, xh64) != 0;
    fails += orc_equalize_f32(U, B, N, w, y, llr, ue_stride, sg, 3, cb, cbs, zeros, valid, xh) != -1;
    std::free(w);
    std::free(y);
    std::free(llr);
    std::free(xh);
    std::free(xh64);
  }
  // ZF with a singular group (a repeated user), precoder normalisation
  {
    const uint32_t U = 5, B = 9, G = 3;
    float* a = exact<float>(2 * U * B * G);
    for (uint32_t i = 0; i < 2 * U * B * G; ++i) a[i] = rnd(seed);
    for (uint32_t b = 0; b < B; ++b) {  // group 1: user 3 = user 1
      a[2 * (U * B + U * b + 3)] = a[2 * (U * B + U * b + 1)];
      a[2 * (U * B + U * b + 3) + 1] = a[2 * (U * B + U * b + 1) + 1];
    }
    float* w = exact<float>(2 * U * B * G);
    double* w64 = exact<double>(2 * U * B * G);
    int32_t* info = exact<int32_t>(G);
    fails += orc_zf_f64(U, B, G, a, w, w64, info) != 1;
    fails += info[1] == 0;
    fails += orc_zf_f32(U, B, G, a, w, info) != 1;
    float* p = exact<float>(2 * U * B * G);
    orc_scale_absmax(U, B, G, w, p);
    std::free(a);
    std::free(w);
    std::free(w64);
    std::free(info);
    std::free(p);
  }
  // precode: a partial last group, shifted placement in rows of 32
  {
    const uint32_t B = 6, U = 3, N = 20, sg = 6, G = N / sg;
    const uint64_t row = 32, start = 5;
    float* prec = exact<float>(2 * B * U * G);
    float* ue = exact<float>(2 * U * N);
    for (uint32_t i = 0; i < 2 * B * U * G; ++i) prec[i] = rnd(seed);
    for (uint32_t i = 0; i < 2 * U * N; ++i) ue[i] = rnd(seed);
    float* out = exact<float>(2 * B * row);
    orc_precode(B, U, N, sg, prec, ue, out, row, start);
    std::free(prec);
    std::free(ue);
    std::free(out);
  }
  // the OpenMP CPU baseline, two slots, a partial last group
  {
    const uint32_t slots = 2, S = 2, U = 4, B = 8, N = 30, sg = 7, G = (N + sg - 1) / sg;
    const uint64_t cb = N * 4, cbs = (cb + 63) / 64 * 64;
    float* csi = exact<float>(2ull * slots * G * B * U);
    float* y = exact<float>(2ull * slots * S * B * N);
    for (uint64_t i = 0; i < 2ull * slots * G * B * U; ++i) csi[i] = rnd(seed);
    for (uint64_t i = 0; i < 2ull * slots * S * B * N; ++i) y[i] = rnd(seed);
    uint16_t* llr = exact<uint16_t>(slots * S * U * cbs);
    fails += orc_detect_cpu(slots, S, U, B, N, sg, 4, csi, y, llr, cbs, cb, cbs, 0, cb, 2) <= 0;
    std::free(csi);
    std::free(y);
    std::free(llr);
  }
  // LDPC: the smallest code (BG2, Zc = 2, 4 parity rows) and the largest
  // (BG1, Zc = 384, 46 rows), two code blocks each
  {
    struct Case {
      int kb, mb, zc, set;
      const int16_t* rs;
      const int8_t* col;
      const int16_t* sh;
    } cases[] = {{10, 4, 2, 0, mega::ldpc::kBg2RowStart, mega::ldpc::kBg2Col,
                  mega::ldpc::kBg2Shift[0]},
                 {22, 46, 384, 1, mega::ldpc::kBg1RowStart, mega::ldpc::kBg1Col,
                  mega::ldpc::kBg1Shift[1]}};
    for (const Case& c : cases) {
      const int nv = (c.kb + c.mb) * c.zc, words = (c.kb * c.zc + 31) / 32;
      float* llr = exact<float>(2 * (size_t)nv);
      for (int i = 2 * c.zc; i < 2 * nv; ++i) llr[i] = 3.f * rnd(seed) + 1.f;
      uint32_t* out = exact<uint32_t>(2 * (size_t)words);
      float* app = exact<float>(2 * (size_t)nv);
      fails += orc_ldpc_decode(c.kb, c.mb, c.zc, c.rs, c.col, c.sh, 3, 0.8125f, llr, nv, 2, out,
                               words, app) != 0;
      std::free(llr);
      std::free(out);
      std::free(app);
    }
  }
  std::printf("oracle under sanitizers: %d failed\n", fails);
  return fails ? 1 : 0;
}
