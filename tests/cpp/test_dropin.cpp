// C++ drop-in tests: the reference's unit tests restated against the
// mega:: classes of include/mega/ (no gtest in this image: a few macros).
//
//   test_dropin host            descriptor tests + demod known answer (no GPU)
//   test_dropin gpu <dir>       all of the above with device buffers, plus the
//                               equalize, precode, FFT and beamform fixtures
//                               read from <dir>
//
// Mirrors test/matrix/test_matrix.cc, test/matrix/test_cuphy_tensor.cc,
// test/modulation/test_modulation.cc:11-57, test/sgemv/test_sgemv.cc:126-196,
// test/fft/test_fft.cc:56-160, test/beamform/test_beamform.cc:67-95 and
// test/ldpc/test_ldpc.cc:75-88,155-214 (configuration and decoder tests).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mega/beamform.h"
#include "mega/demodulation.h"
#include "mega/equalize.h"
#include "mega/fft.h"
#include "mega/ldpc.h"
#include "mega/matrix.h"
#include "mega/mega_complex.h"
#include "mega/padded_tensor.h"
#include "mega/precode.h"

static int g_fail = 0, g_checks = 0;
#define EXPECT(c)                                                        \
  do {                                                                   \
    ++g_checks;                                                          \
    if (!(c)) {                                                          \
      ++g_fail;                                                          \
      std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);         \
    }                                                                    \
  } while (0)
#define EXPECT_THROW(stmt, ex)                     \
  do {                                             \
    bool thrown = false;                           \
    try {                                          \
      stmt;                                        \
    } catch (const ex&) {                          \
      thrown = true;                               \
    }                                              \
    EXPECT(thrown);                                \
  } while (0)

using mega::Matrix;
using mega::CuphyTensor;

static Matrix::alloc_type_t g_dev = Matrix::kHost, g_pin = Matrix::kHost;

static void matrix_tests() {
  std::printf("Matrix\n");
  {
    Matrix m(sizeof(short), 3, Matrix::kHost);
    EXPECT(m.nDim() == 1);
    EXPECT(m.allocType() == Matrix::kHost);
    EXPECT(m.raw() != nullptr);
    EXPECT(m.dim(0) == 3 && m.stride(0) == 1);
    EXPECT(m.szBytes(0) == sizeof(short) && m.szBytes(1) == 3 * sizeof(short));
  }
  {
    Matrix m(sizeof(short), 3, 4, g_dev);
    EXPECT(m.nDim() == 2 && m.allocType() == g_dev && m.raw() != nullptr);
    EXPECT(m.dim(0) == 3 && m.dim(1) == 4 && m.stride(0) == 1 && m.stride(1) == 3);
    EXPECT(m.szBytes(0) == 2 && m.szBytes(1) == 6 && m.szBytes(2) == 24);
  }
  {
    Matrix m(sizeof(short), 3, 4, 5, g_dev);
    EXPECT(m.nDim() == 3 && m.dim(2) == 5 && m.stride(2) == 12);
    EXPECT(m.szBytes(2) == 24 && m.szBytes(3) == 120);
  }
  {
    Matrix m(sizeof(int), 5, 3, 4, g_pin);
    int ref[60];
    for (int i = 0; i < 60; ++i) ref[i] = i;
    std::memcpy(m.ptr(), ref, sizeof(ref));
    int id = 0;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 5; ++k) EXPECT(*(m[i][j][k].ptr<int>()) == ref[id++]);
    EXPECT_THROW(m[0][0][0][0], std::out_of_range);
    EXPECT_THROW(m[4], std::out_of_range);
    m[0][0][0].deref<int>(0) = -100;
    EXPECT(*(m[0][0][0].ptr<int>()) == -100);
    Matrix s = m.sub(1, 3);
    EXPECT(s.dim(2) == 2 && *(s[0][0][0].ptr<int>()) == 15);
    EXPECT_THROW(m.sub(2, 5), std::out_of_range);
  }
}

static void tensor_tests() {
  std::printf("CuphyTensor\n");
  {
    CuphyTensor t(CuphyTensor::kHalf, 4, Matrix::kHost);
    EXPECT(t.nDim() == 1 && t.raw() != nullptr && t.dim(0) == 4 && t.stride(0) == 1);
    EXPECT(t.szBytes(0) == 2 && t.szBytes(1) == 8);
  }
  {
    CuphyTensor t(CuphyTensor::kFloat, 4, 5, g_dev, CuphyTensor::kCoalesce);
    EXPECT(t.nDim() == 2 && t.allocType() == g_dev && t.raw() != nullptr);
    EXPECT(t.stride(0) == 1 && t.stride(1) == 32);
    EXPECT(t.szBytes(0) == 4 && t.szBytes(1) == 128 && t.szBytes(2) == 640);
  }
  {
    CuphyTensor t(CuphyTensor::kBit, 8, 5, 6, Matrix::kHost, CuphyTensor::kCoalesce);
    EXPECT(t.stride(1) == 1024 && t.stride(2) == 1024 * 5);
    EXPECT(t.szBytes(0) == 0 && t.szBytes(1) == 128 && t.szBytes(2) == 640 &&
           t.szBytes(3) == 3840);
  }
  {
    CuphyTensor t(CuphyTensor::kBit, 8, 5, 3, Matrix::kHost, CuphyTensor::kTight);
    uint32_t ref[15];
    for (int i = 0; i < 15; ++i) ref[i] = i;
    std::memcpy(t.ptr(), ref, sizeof(ref));
    int id = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 5; ++j) EXPECT(*(uint32_t*)(t[i][j].ptr()) == ref[id++]);
    EXPECT_THROW(t[3], std::out_of_range);
    EXPECT_THROW(t[(uint32_t)-1], std::out_of_range);
    EXPECT_THROW(t[0][0][0], std::underflow_error);
  }
  {
    CuphyTensor t(CuphyTensor::kBit, 8, Matrix::kHost);
    EXPECT(t.stride(0) == 1 && t.szBytes(0) == 0 && t.szBytes(1) == 4);
    CuphyTensor u(CuphyTensor::kBit, 8, 5, Matrix::kHost);
    EXPECT(u.stride(1) == 32 && u.szBytes(1) == 4 && u.szBytes(2) == 20);
  }
  {
    // the equalizer's LLR buffer of the reference pipeline (buffers.cc:97-103)
    CuphyTensor llr(CuphyTensor::kHalf, 19008 + 576, 32, 13, Matrix::kDesc,
                    CuphyTensor::kCoalesce);
    EXPECT(llr.stride(1) == 19584);
  }
}

static void demod_known_answer() {
  std::printf("Demod 64-QAM known answer\n");
  // 64-QAM table as the reference modulator builds it (modulation.cu:42-53):
  // index bits 5,3,1 -> real level, 4,2,0 -> imaginary level.
  const float s = 1.0f / std::sqrt(42.0f);
  const float lv[8] = {3 * s, 1 * s, 5 * s, 7 * s, -3 * s, -1 * s, -5 * s, -7 * s};
  uint8_t stream[48] = {0};
  for (int bit = 0; bit < 64 * 6; ++bit)
    stream[bit / 8] |= ((bit / 6 >> (bit % 6)) & 1) << (bit % 8);
  for (int c = 0; c < 64; ++c) {
    int idx = 0;
    for (int i = 0; i < 6; ++i) {
      const int m = c * 6 + i;
      idx = ((stream[m / 8] >> (m % 8)) & 1) | (idx << 1);
    }
    const int im_i = (((idx >> 4) & 1) << 2) + (((idx >> 2) & 1) << 1) + (idx & 1);
    const int re_i = (((idx >> 5) & 1) << 2) + (((idx >> 3) & 1) << 1) + ((idx >> 1) & 1);
    mega::half llr[6];
    mega::demod64QAM(mega::Complex(lv[re_i], lv[im_i]), llr);
    int v = 0;
    for (int b = 0; b < 6; ++b) v |= (llr[b] < (mega::half)0.0f) << b;
    EXPECT(v == c);
  }
}

template <typename T>
static std::vector<T> read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<T> v(n / sizeof(T));
  f.read(reinterpret_cast<char*>(v.data()), n);
  return v;
}

static void equalize_fixture(const std::string& dir) {
  std::printf("Equalize fixture\n");
  const uint32_t ue = 4, bs = 8, ofdm = 16;
  auto csi = read_file<mega::Complex>(dir + "/sgemv_csi.bin");
  auto bsv = read_file<mega::Complex>(dir + "/sgemv_bsvec_kb.bin");
  auto ref = read_file<uint16_t>(dir + "/sgemv_equalize_f16bits.bin");
  Matrix d_csi(sizeof(mega::Complex), ue, bs, Matrix::kDevice);
  Matrix d_bsv(sizeof(mega::Complex), ofdm, bs, Matrix::kDevice);
  Matrix d_eq(sizeof(mega::half), 2 * ofdm, ue, Matrix::kDevice);
  EXPECT(csi.size() * 8 == d_csi.szBytes() && bsv.size() * 8 == d_bsv.szBytes());
  std::vector<mega::Complex> t(bs * ofdm);
  for (uint32_t i = 0; i < ofdm; ++i)
    for (uint32_t j = 0; j < bs; ++j) t[j * ofdm + i] = bsv[i * bs + j];
  EXPECT(hipMemcpy(d_csi.ptr(), csi.data(), d_csi.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  EXPECT(hipMemcpy(d_bsv.ptr(), t.data(), d_bsv.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  mega::Equalize::equalize(d_csi, d_bsv, d_eq, ofdm, 2, 2 * ofdm, 0, 1);
  std::vector<uint16_t> out(2 * ue * ofdm);
  EXPECT(hipMemcpy(out.data(), d_eq.ptr(), d_eq.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  for (uint32_t i = 0; i < 2 * ue * ofdm; ++i) EXPECT(out[i] == ref[i]);
  // an unsupported modulation order writes nothing and does not throw
  EXPECT(hipMemset(d_eq.ptr(), 0, d_eq.szBytes()) == hipSuccess);
  mega::Equalize::equalize(d_csi, d_bsv, d_eq, ofdm, 3, 2 * ofdm, 0, 1);
  EXPECT(hipMemcpy(out.data(), d_eq.ptr(), d_eq.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  bool zero = true;
  for (auto v : out) zero = zero && v == 0;
  EXPECT(zero);
}

static void precode_fixture(const std::string& dir) {
  std::printf("Precode fixture\n");
  const uint32_t ue = 4, bs = 8, ofdm = 16, ofdmE = 32;
  auto prec = read_file<mega::Complex>(dir + "/precode_prec.bin");
  auto uev = read_file<mega::Complex>(dir + "/precode_uevec_ku.bin");
  auto ref = read_file<float>(dir + "/precode_out.bin");
  Matrix d_prec(sizeof(mega::Complex), bs, ue, Matrix::kDevice);
  Matrix d_ue(sizeof(mega::Complex), ofdm, ue, Matrix::kDevice);
  Matrix d_out(sizeof(mega::Complex), ofdmE, bs, Matrix::kDevice);
  EXPECT(prec.size() * 8 == d_prec.szBytes() && uev.size() * 8 == d_ue.szBytes());
  EXPECT(ref.size() * 4 == d_out.szBytes());
  std::vector<mega::Complex> t(ue * ofdm);
  for (uint32_t i = 0; i < ofdm; ++i)
    for (uint32_t j = 0; j < ue; ++j) t[j * ofdm + i] = uev[i * ue + j];
  EXPECT(hipMemcpy(d_prec.ptr(), prec.data(), d_prec.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  EXPECT(hipMemcpy(d_ue.ptr(), t.data(), d_ue.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  EXPECT(hipMemset(d_out.ptr(), 0, d_out.szBytes()) == hipSuccess);
  mega::Precode::precode(d_prec, d_ue, d_out, ofdm, 8);
  std::vector<float> out(2 * bs * ofdmE);
  EXPECT(hipMemcpy(out.data(), d_out.ptr(), d_out.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  for (uint32_t i = 0; i < 2 * bs * ofdmE; ++i) EXPECT(std::fabs(out[i] - ref[i]) < 1e-5f);
}

static void fft_fixtures(const std::string& dir) {
  std::printf("FFT fixtures\n");
  const uint32_t bs = 8, ca = 32, num = 16, ue = 4;
  // UplinkFFT(8, 32, 16, 8): reference output stored [subcarrier][antenna]
  auto in = read_file<short>(dir + "/fft_uplink_in.bin");
  auto ref = read_file<mega::Complex>(dir + "/fft_uplink_out_kb.bin");
  Matrix d_in(2 * sizeof(short), ca, bs, Matrix::kDevice);
  Matrix d_out(sizeof(mega::Complex), num, bs, Matrix::kDevice);
  EXPECT(in.size() * 2 == d_in.szBytes());
  EXPECT(hipMemcpy(d_in.ptr(), in.data(), d_in.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  mega::UplinkFFT<short> fft(8, ca, num, bs);
  fft(d_in, d_out);
  std::vector<mega::Complex> out(bs * num);
  EXPECT(hipMemcpy(out.data(), d_out.ptr(), d_out.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  for (uint32_t i = 0; i < num; ++i)
    for (uint32_t j = 0; j < bs; ++j) {
      EXPECT(std::fabs(out[j * num + i].re - ref[i * bs + j].re) < 1e-5f);
      EXPECT(std::fabs(out[j * num + i].im - ref[i * bs + j].im) < 1e-5f);
    }
  // PilotFFT(8, 32, 16, 8, 4, 4) with the pilot sign of 16 subcarriers
  auto pin = read_file<short>(dir + "/fft_pilot_in.bin");
  auto cref = read_file<mega::Complex>(dir + "/fft_pilot_csi.bin");
  mega::PilotSign::init(num);
  Matrix d_pin(2 * sizeof(short), ca, bs, Matrix::kDevice);
  Matrix d_csi(sizeof(mega::Complex), ue, bs, num / ue, Matrix::kDevice);
  EXPECT(hipMemcpy(d_pin.ptr(), pin.data(), d_pin.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  mega::PilotFFT<short> pfft(8, ca, num, bs, ue, ue);
  pfft(d_pin, d_csi);
  std::vector<mega::Complex> csi(bs * num);
  EXPECT(hipMemcpy(csi.data(), d_csi.ptr(), d_csi.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  for (uint32_t i = 0; i < bs * num; ++i) {
    EXPECT(std::fabs(csi[i].re - cref[i].re) < 1e-4f);
    EXPECT(std::fabs(csi[i].im - cref[i].im) < 1e-4f);
  }
  mega::PilotSign::destroy();
  EXPECT_THROW(pfft(d_pin, d_csi), std::runtime_error);  // no table after destroy
  mega::UplinkFFT<short> bad(0, 48, 16, 8);              // not a power of two
  EXPECT_THROW(bad(d_in, d_out), std::runtime_error);
  // DownlinkIFFT(8, 32, 16, 8): int16 output compared exactly (:162-206)
  auto iin = read_file<mega::Complex>(dir + "/ifft_in.bin");
  auto iref = read_file<short>(dir + "/ifft_out.bin");
  Matrix d_iin(sizeof(mega::Complex), ca, bs, Matrix::kDevice);
  Matrix d_iout(2 * sizeof(short), ca, bs, Matrix::kDevice);
  EXPECT(hipMemcpy(d_iin.ptr(), iin.data(), d_iin.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
  mega::DownlinkIFFT<short> ifft(8, ca, num, bs);
  ifft(d_iin, d_iout);
  std::vector<short> iout(2 * bs * ca);
  EXPECT(hipMemcpy(iout.data(), d_iout.ptr(), d_iout.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  for (uint32_t i = 0; i < 2 * bs * ca; ++i) EXPECT(iout[i] == iref[i]);
}

static void beamform_fixture(const std::string& dir) {
  std::printf("Beamform fixture\n");
  const uint32_t ue = 4, bs = 8;
  auto in = read_file<float>(dir + "/beam_csi_in.bin");
  auto w_ref = read_file<float>(dir + "/beam_csi_out.bin");
  auto p_ref = read_file<float>(dir + "/beam_precoder.bin");
  Matrix csi(sizeof(mega::Complex), ue, bs, 2, Matrix::kDevice);
  Matrix prec(sizeof(mega::Complex), bs, ue, 2, Matrix::kDevice);
  Matrix h_csi(sizeof(mega::Complex), ue, bs, 2, Matrix::kHost);
  Matrix h_prec(sizeof(mega::Complex), bs, ue, 2, Matrix::kHost);
  const size_t sz = in.size() * sizeof(float);
  EXPECT(sz == csi.szBytes() / 2);
  EXPECT(hipMemcpy(csi[0].ptr(), in.data(), sz, hipMemcpyHostToDevice) == hipSuccess);
  EXPECT(hipMemcpy(csi[1].ptr(), in.data(), sz, hipMemcpyHostToDevice) == hipSuccess);
  mega::Beamform beam(bs, ue, 2);
  mega::Beamform copy = beam;  // copies share the handle
  copy.beamformer(csi, prec);
  EXPECT(hipMemcpy(h_csi.ptr(), csi.ptr(), csi.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  EXPECT(hipMemcpy(h_prec.ptr(), prec.ptr(), prec.szBytes(), hipMemcpyDeviceToHost) == hipSuccess);
  for (uint32_t b = 0; b < 2; ++b) {
    const float* w = h_csi[b].ptr<float>();
    const float* p = h_prec[b].ptr<float>();
    for (uint32_t i = 0; i < 2 * ue * bs; ++i) {
      EXPECT(std::fabs(w[i] - w_ref[i]) < 1e-5f);
      EXPECT(std::fabs(p[i] - p_ref[i]) < 1e-5f);
    }
  }
  auto info = beam.info(2);
  EXPECT(info[0] == 0 && info[1] == 0);
  Matrix big(sizeof(mega::Complex), ue, bs, 3, Matrix::kDevice);
  EXPECT_THROW(beam.beamform_uplink(big), std::runtime_error);
  Matrix wrong(sizeof(mega::Complex), bs, ue, 3, Matrix::kDevice);
  EXPECT_THROW(beam.beamformer(csi, wrong), std::runtime_error);
  Matrix bad(sizeof(mega::Complex), ue, bs, 2, Matrix::kDevice);
  EXPECT_THROW(mega::Beamform::beamform_downlink(csi, bad), std::runtime_error);
}

// test_ldpc.cc:75-88: the configuration's derived sizes.
static void ldpc_config_test() {
  std::printf("LDPCConfig\n");
  mega::LDPCConfig config(46, 1, 64, 8);
  EXPECT(config.parity_nodes == 46 && config.base_graph == 1);
  EXPECT(config.lifting_factor == 64 && config.max_decode_iter == 8);
  EXPECT(config.info_nodes == 22);
  EXPECT(config.encoded_bits == 4224 && config.decoded_bits == 1408);
  EXPECT(config.punctured_bits == 128);
  mega::LDPCConfig c2(42, 2, 384, 1);
  EXPECT(c2.info_nodes == 10 && c2.encoded_bits == 50ull * 384 && c2.decoded_bits == 3840);
}

// test_ldpc.cc:155-214: noiseless float LLRs (+-127 after the punctured
// zero prefix) of encode.data decode to input.data, one code block and two.
static void ldpc_fixture(const std::string& dir) {
  std::printf("LDPCDecoder\n");
  const auto enc = read_file<unsigned char>(dir + "/ldpc_encoded_bits.bin");
  const auto inp = read_file<unsigned char>(dir + "/ldpc_input_bits.bin");
  mega::LDPCConfig config(46, 1, 64, 8);
  EXPECT(enc.size() == 2 * config.encoded_bits / 8 && inp.size() == 2 * config.decoded_bits / 8);
  mega::LDPCDecoder decoder(CuphyTensor::kFloat, config);
  const uint32_t nllr = (uint32_t)(config.encoded_bits + config.punctured_bits);
  for (uint32_t ncb = 1; ncb <= 2; ++ncb) {
    std::vector<float> llr((size_t)nllr * ncb, 0.f);
    for (uint32_t c = 0; c < ncb; ++c)
      for (uint64_t i = 0; i < config.encoded_bits; ++i) {
        const uint64_t bit = c * config.encoded_bits + i;
        const bool one = enc[bit / 8] & (1u << (bit % 8));
        llr[(size_t)c * nllr + config.punctured_bits + i] = one ? -127.f : 127.f;
      }
    CuphyTensor d_llr(CuphyTensor::kFloat, nllr, ncb, Matrix::kDevice);
    CuphyTensor d_out(CuphyTensor::kBit, (uint32_t)config.decoded_bits, ncb, Matrix::kDevice);
    EXPECT(d_out.szBytes() == ncb * config.decoded_bits / 8);
    EXPECT(hipMemcpy(d_llr.ptr(), llr.data(), d_llr.szBytes(), hipMemcpyHostToDevice) == hipSuccess);
    decoder(d_llr, d_out);
    std::vector<unsigned char> out(d_out.szBytes());
    EXPECT(hipMemcpy(out.data(), d_out.ptr(), out.size(), hipMemcpyDeviceToHost) == hipSuccess);
    EXPECT(std::memcmp(out.data(), inp.data(), out.size()) == 0);
  }
  // the LLR type must match the decoder's
  CuphyTensor half_llr(CuphyTensor::kHalf, nllr, Matrix::kDesc);
  CuphyTensor out1(CuphyTensor::kBit, (uint32_t)config.decoded_bits, Matrix::kDesc);
  EXPECT_THROW(decoder(half_llr, out1), std::invalid_argument);
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "host";
  try {
    if (mode == "gpu") {
      if (argc < 3) {
        std::printf("usage: test_dropin gpu <fixture dir>\n");
        return 2;
      }
      g_dev = Matrix::kDevice;
      g_pin = Matrix::kHPin;
    }
    matrix_tests();
    tensor_tests();
    demod_known_answer();
    ldpc_config_test();
    if (mode == "gpu") {
      ldpc_fixture(argv[2]);
      equalize_fixture(argv[2]);
      precode_fixture(argv[2]);
      fft_fixtures(argv[2]);
      beamform_fixture(argv[2]);
    }
  } catch (const std::exception& e) {
    std::printf("  EXCEPTION %s\n", e.what());
    ++g_fail;
  }
  std::printf("%d checks, %d failed\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
