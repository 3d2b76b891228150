"""Write tests/golden/reference_fixtures.npz from the data files that the
reference's own unit tests read.

Sources (binary fixtures next to the reference tests):
  test/sgemv/csi.data       U=4 x B=8 CSI, column-major complex64
                            (test/sgemv/test_sgemv.cc:34-46)
  test/sgemv/bsVec.data     y as [ofdm=16][bs=8] complex64; the test transposes
                            it to [bs][ofdm] before upload (test_sgemv.cc:136-146)
  test/sgemv/equalize.data  expected QPSK LLRs, binary16, [ue=4][2*ofdm=32]
                            (test_sgemv.cc:126-159, compared with EXPECT_EQ)
  test/beamform/csi_in.data   U=4 x B=8 CSI (test/beamform/test_beamform.cc:19-33)
  test/beamform/csi_out.data  expected ZF weights, same layout (1e-5, :91)
  test/beamform/precoder.data expected normalised precoder, B x U (1e-5, :92)
  test/sgemv/prec.data      precoder, bs=8 x ue=4 column-major complex64, one
                            group of 16 subcarriers (test_sgemv.cc:161-196)
  test/sgemv/ueVec.data     UE symbols as [ofdm=16][ue=4] complex64; the test
                            transposes them to [ue][ofdm] before upload
  test/sgemv/precoded.data  expected output, [bs=8][ofdmE=32] complex64, for
                            precode_skip = 16 and bs_start = 8 (1e-5)
  test/fft/fft_in.data      uplink FFT input, [bs=8][32] int16 (re, im) pairs
                            (test/fft/test_fft.cc:105-160)
  test/fft/fft_out.data     expected uplink output, [16][bs=8] complex64 for
                            ofdm_start 8, ofdm_ca 32, ofdm_num 16 (1e-5)
  test/fft/short_in.data    pilot FFT input, [bs=8][32] int16 pairs (:56-103)
  test/fft/complex_out.data expected CSI, [G=4][bs=8][ue=4] complex64 for
                            ofdm_start 8, ofdm_ca 32, 16 subcarriers, groups
                            of 4, pilot sign of 16 (1e-4)
  test/fft/zachu.data       Zadoff-Chu sequence of length 32, first 31 values
                            (complex64, 1e-5; :10-30)
  test/fft/zachu_shift.data Zadoff-Chu of 16 times the pi/4 cyclic shift (:32-54)
  test/fft/ifft_in.data     downlink IFFT input, [bs=8][32] complex64 (:162-206)
  test/fft/ifft_out.data    expected int16 (re, im) pairs, [bs=8][32] (exact)
  test/ldpc/input.data      two code blocks of information bits, BG1, Zc = 64,
                            22 * 64 = 1408 bits each, bit-packed LSB first
                            (test/ldpc/test_ldpc.cc:41-60)
  test/ldpc/encode.data     their codewords without the 2 * Zc punctured bits,
                            (22 - 2 + 46) * 64 = 4224 bits each, same packing
                            (:62-73, :155-190)

The arrays are stored unchanged (raw little-endian values); this script only
re-packages them so the tests run where the reference tree is absent.
Run:  python tests/golden/make_reference_fixtures.py [/root/reference]
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "reference_fixtures.npz")


def rd(rel, dtype):
    return np.fromfile(os.path.join(REF, rel), dtype=dtype)


def main():
    arrays = {
        "ldpc_input_bits": rd("test/ldpc/input.data", np.uint8),
        "ldpc_encoded_bits": rd("test/ldpc/encode.data", np.uint8),
        "sgemv_csi": rd("test/sgemv/csi.data", np.complex64),
        "sgemv_bsvec_kb": rd("test/sgemv/bsVec.data", np.complex64),
        "sgemv_equalize_f16bits": rd("test/sgemv/equalize.data", np.uint16),
        "beam_csi_in": rd("test/beamform/csi_in.data", np.complex64),
        "beam_csi_out": rd("test/beamform/csi_out.data", np.complex64),
        "beam_precoder": rd("test/beamform/precoder.data", np.complex64),
        "precode_prec": rd("test/sgemv/prec.data", np.complex64),
        "precode_uevec_ku": rd("test/sgemv/ueVec.data", np.complex64),
        "precode_out": rd("test/sgemv/precoded.data", np.complex64),
        "fft_uplink_in": rd("test/fft/fft_in.data", np.int16),
        "fft_uplink_out_kb": rd("test/fft/fft_out.data", np.complex64),
        "fft_pilot_in": rd("test/fft/short_in.data", np.int16),
        "fft_pilot_csi": rd("test/fft/complex_out.data", np.complex64),
        "zadoff_chu_32": rd("test/fft/zachu.data", np.complex64),
        "zadoff_chu_16_shift": rd("test/fft/zachu_shift.data", np.complex64),
        "ifft_in": rd("test/fft/ifft_in.data", np.complex64),
        "ifft_out": rd("test/fft/ifft_out.data", np.int16),
    }
    assert arrays["sgemv_csi"].size == 4 * 8
    assert arrays["sgemv_bsvec_kb"].size == 16 * 8
    assert arrays["sgemv_equalize_f16bits"].size == 4 * 32
    assert arrays["beam_csi_in"].size == 4 * 8
    assert arrays["precode_prec"].size == 8 * 4
    assert arrays["precode_uevec_ku"].size == 16 * 4
    assert arrays["precode_out"].size == 8 * 32
    assert arrays["fft_uplink_in"].size == 2 * 8 * 32
    assert arrays["fft_uplink_out_kb"].size == 16 * 8
    assert arrays["fft_pilot_in"].size == 2 * 8 * 32
    assert arrays["fft_pilot_csi"].size == 4 * 8 * 4
    assert arrays["zadoff_chu_32"].size == 31
    assert arrays["zadoff_chu_16_shift"].size == 16
    assert arrays["ifft_in"].size == 8 * 32
    assert arrays["ifft_out"].size == 2 * 8 * 32
    np.savez(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
