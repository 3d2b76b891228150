"""The 64-antenna x 16-user cell (1200 subcarriers in groups of 16, i.i.d.
CN(0,1) channels) at SNR 10-30 dB in 5 dB steps with 16- and 64-QAM, in the
plain FP32 form (mega_zf then mega_equalize_slots) and the mixed form
(mega_detect_slots, split BF16), against the FP64 oracle on the same inputs:
x_hat, LLRs and hard decisions by the comparison policy, EVM within 0.1 dB of
FP64's, and the bit errors against the transmitted bits equal to FP64's up
to the hard decisions within the policy's margin of the decision boundary
(hard decisions: LLR sign, LLR > 0 = bit 0, demodulation.h:35-67)."""
import numpy as np
import pytest

import megastation_amd as m
import oracle
from megastation_amd import synth

from gpu_helpers import (HARD_BIT_MARGIN, SENTINEL, check_symbol, dev_c64, f16,
                         gpu_equalize_slots, make_slots, oracle_symbol_f64, torch)
from test_gpu_mixed import _detect_mixed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["plain", "mixed"])
@pytest.mark.parametrize("Q", [4, 6])
@pytest.mark.parametrize("snr", [10.0, 15.0, 20.0, 25.0, 30.0])
def test_c2_cell_snr_qam_sweep(snr, Q, form):
    U, B, N, sg, S, slots = 16, 64, 1200, 16, 3, 2
    print(f"case: sweep {form} {1 << Q}-QAM {snr:.0f} dB U={U} B={B} N={N}", flush=True)
    G = N // sg
    sl = make_slots(int(snr) * 7 + Q, slots, U, B, N, sg, S, Q, snr, "iid")
    lay = synth.standalone_layout(N, Q)
    csi = np.stack([s.csi for s in sl])
    y = np.stack([s.y for s in sl])
    if form == "plain":
        t = torch()
        w = t.zeros(slots * G * B * U, dtype=t.complex64, device="cuda:0")
        m.zf(dev_c64(csi), w, B, U, slots * G)
        t.cuda.synchronize()
        w_h = w.cpu().numpy().reshape(slots, G, B, U)
        llr, xh = gpu_equalize_slots(w_h, y, U, B, N, sg, Q, lay, slots=slots, S=S)
    else:
        llr, xh, info, route = _detect_mixed(csi, y, U, B, N, sg, Q, lay, slots, S)
        assert not info.any() and not route.any()
    err_gpu = err_64 = near = nbits = 0
    for s_ in range(slots):
        w64, _, _ = oracle.zf(sl[s_].csi, U, B, G, "f64")
        for t_ in range(S):
            llr64, x64 = oracle_symbol_f64(w64, sl[s_].y[t_], U, B, N, sg, Q, lay)
            check_symbol(llr[s_, t_], xh[s_, t_], llr64, x64, N, sg, Q, lay)
            e_gpu = synth.evm_db(xh[s_, t_], sl[s_].x[t_])
            e_64 = synth.evm_db(x64, sl[s_].x[t_])
            assert abs(e_gpu - e_64) < 0.1, (e_gpu, e_64)
            bits = sl[s_].bits[t_]                                   # [U][N][Q]
            hg = synth.hard_bits_llr(llr[s_, t_][:, :N * Q].reshape(U, N, Q))
            h64 = synth.hard_bits_llr(llr64[:, :N * Q].reshape(U, N, Q))
            err_gpu += int((hg != bits).sum())
            err_64 += int((h64 != bits).sum())
            near += int((np.abs(f16(llr64[:, :N * Q])) < HARD_BIT_MARGIN).sum())
            nbits += bits.size
    assert abs(err_gpu - err_64) <= near, (err_gpu, err_64, near)
    print(f"  BER {form} {err_gpu / nbits:.3e}  FP64 {err_64 / nbits:.3e}  "
          f"({nbits} bits, {near} within the margin)", flush=True)
