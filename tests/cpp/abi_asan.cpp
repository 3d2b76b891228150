// The C ABI's host-side code under AddressSanitizer and UndefinedBehavior-
// Sanitizer (built by tests/sanitize_build.py against a copy of the library whose
// host code carries the sanitizers; tests/test_sanitize.py runs it, no GPU
// needed): every entry point's argument checks with bad arguments (they
// return before any device work), the host demod over a sweep of values,
// the pilot-sign table, the tuning overrides and the error strings.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "megadetect.h"

static int g_fail = 0;
#define EXPECT(c)                                                \
  do {                                                           \
    if (!(c)) {                                                  \
      ++g_fail;                                                  \
      std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
    }                                                            \
  } while (0)

int main() {
  void* fake = reinterpret_cast<void*>(0x1000);  // never dereferenced on an error path
  EXPECT(std::strstr(mega_version(), "gfx950") != nullptr);
  EXPECT(mega_tuning_reload() == MEGA_OK);

  // beamform / ZF
  mega_beamform_t h = nullptr;
  EXPECT(mega_beamform_create(0, 4, 2, &h) == MEGA_ERR_INVALID_ARG);
  EXPECT(mega_beamform_create(256, 129, 2, &h) == MEGA_ERR_UNSUPPORTED);
  EXPECT(mega_beamform_create(8, 4, 2, nullptr) == MEGA_ERR_INVALID_ARG);
  EXPECT(mega_beamform_uplink(nullptr, nullptr, 1, nullptr) == MEGA_ERR_INVALID_ARG);
  EXPECT(mega_beamform_destroy(nullptr) == MEGA_OK);
  EXPECT(mega_zf(8, 4, 1, nullptr, nullptr, nullptr, 0, nullptr) == MEGA_ERR_INVALID_ARG);
  EXPECT(mega_zf(8, 4, 1, fake, fake, nullptr, 7, nullptr) == MEGA_ERR_UNSUPPORTED);
  EXPECT(mega_zf_split(8, 65, 1, fake, fake, nullptr, nullptr, nullptr, nullptr) ==
         MEGA_ERR_UNSUPPORTED);
  EXPECT(mega_zf(256, 65, 1, fake, fake, nullptr, MEGA_FORM_MIXED, nullptr) ==
         MEGA_ERR_UNSUPPORTED);
  EXPECT(mega_zf(256, 129, 1, fake, fake, nullptr, MEGA_FORM_FP32, nullptr) ==
         MEGA_ERR_UNSUPPORTED);

  // equalize: bad modulation order, NULL buffers, zero group size
  EXPECT(mega_equalize(4, 8, 16, fake, fake, fake, 32, 16, 3, 32, 32, 0, 32, nullptr) ==
         MEGA_ERR_UNSUPPORTED);
  std::string msg = mega_last_error();
  EXPECT(msg.find("modulation order 3") != std::string::npos);
  mega_equalize_desc d;
  std::memset(&d, 0, sizeof(d));
  d.num_ue = 4, d.num_bs = 8, d.num_sc = 16, d.sc_group = 4, d.num_syms = 1, d.num_slots = 1;
  d.mod_order = 4, d.ue_stride = 64, d.cb_size = 64, d.cb_stride = 64, d.valid_bits = 64;
  EXPECT(mega_equalize_slots(&d, nullptr) == MEGA_ERR_INVALID_ARG);  // NULL buffers
  d.csi = d.y = fake;
  d.llr = fake;
  d.sc_group = 0;
  EXPECT(mega_equalize_slots(&d, nullptr) == MEGA_ERR_INVALID_ARG);
  d.sc_group = 4;
  d.form = 9;
  EXPECT(mega_equalize_slots(&d, nullptr) == MEGA_ERR_UNSUPPORTED);
  EXPECT(mega_equalize_slots(nullptr, nullptr) == MEGA_ERR_INVALID_ARG);

  // detect: the mixed form without its work buffer, its byte count
  mega_detect_desc dd;
  std::memset(&dd, 0, sizeof(dd));
  dd.eq = d;
  dd.eq.form = MEGA_FORM_MIXED;
  EXPECT(mega_detect_slots(&dd, nullptr) == MEGA_ERR_INVALID_ARG);
  EXPECT(mega_detect_work_bytes(&dd) == 1ull * 4 * 16 * 4 * 8);
  EXPECT(mega_detect_work_bytes(nullptr) == 0);

  // precode, FFT, LDPC, pipeline, batch
  mega_precode_desc pd;
  std::memset(&pd, 0, sizeof(pd));
  pd.num_bs = 8, pd.num_ue = 4, pd.num_sc = 16, pd.sc_group = 4, pd.num_syms = 1;
  EXPECT(mega_precode_syms(&pd, nullptr) == MEGA_ERR_INVALID_ARG);
  mega_fft_desc fd;
  std::memset(&fd, 0, sizeof(fd));
  fd.ofdm_ca = 24, fd.ofdm_num = 16, fd.bs_num = 2, fd.num_syms = 1, fd.quant_bits = 16;
  fd.in = fd.out = fake;
  EXPECT(mega_uplink_fft(&fd, nullptr) == MEGA_ERR_UNSUPPORTED);  // not a power of two
  mega_ldpc_desc ld;
  std::memset(&ld, 0, sizeof(ld));
  ld.base_graph = 1, ld.lifting_size = 17, ld.parity_nodes = 46, ld.max_iterations = 8;
  ld.llr_bits = 16, ld.llr = fake, ld.llr_stride = 1 << 20, ld.num_cb = 1, ld.out = (uint32_t*)fake;
  ld.out_stride = 1 << 10;
  EXPECT(mega_ldpc_decode(&ld, nullptr) == MEGA_ERR_UNSUPPORTED);
  ld.lifting_size = 384, ld.parity_nodes = 47;
  EXPECT(mega_ldpc_decode(&ld, nullptr) == MEGA_ERR_UNSUPPORTED);
  mega_pipeline_t pl = nullptr;
  d.form = MEGA_FORM_MIXED;
  EXPECT(mega_pipeline_create(&d, 1, nullptr, &pl) != MEGA_OK);
  mega_batch_t bt = nullptr;
  EXPECT(mega_batch_create(nullptr, 0, nullptr, &bt) == MEGA_ERR_INVALID_ARG);

  // host demod: every order over a sweep, an unsupported order
  std::vector<float> x(2 * 4096);
  for (size_t i = 0; i < x.size(); ++i) x[i] = std::sin(0.37f * i) * 1.7f;
  std::vector<uint16_t> l(8 * 4096);
  for (uint8_t Q : {2, 4, 6, 8}) EXPECT(mega_demod_host(x.data(), 4096, Q, l.data()) == MEGA_OK);
  EXPECT(mega_demod_host(x.data(), 4096, 5, l.data()) == MEGA_ERR_UNSUPPORTED);

  // pilot sign tables of several lengths (host only)
  for (uint32_t n : {1u, 16u, 31u, 1200u, 3276u}) {
    std::vector<float> p(2 * n);
    EXPECT(mega_pilot_sign(n, p.data()) == MEGA_OK);
    for (float v : p) EXPECT(std::isfinite(v));
  }
  EXPECT(mega_pilot_sign(16, nullptr) == MEGA_ERR_INVALID_ARG);

  std::printf("C ABI under sanitizers: %d failed\n", g_fail);
  return g_fail ? 1 : 0;
}
