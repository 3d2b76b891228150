"""CPU checks of the learned form's oracle and generator (no GPU).

The network oracle (oracle/mlp.py) is pinned by central finite differences
of its own loss in exact arithmetic (rounding off), by known BF16 roundings
and by Adam's first-step property; the channel generator by its defining
statistics (power-delay profile, RMS delay spread, Jakes correlation)."""
import math

import numpy as np
import pytest
import scipy.special

from megastation_amd import synth
from oracle import mlp


def _tiny(seed=0, dims=(6, 5, 4, 3), M=7):
    rng = np.random.default_rng(seed)
    Ws = [rng.standard_normal((dims[l], dims[l + 1])) * 0.7 for l in range(3)]
    bs = [rng.standard_normal(dims[l + 1]) * 0.1 for l in range(3)]
    x = rng.standard_normal((M, dims[0]))
    t = rng.standard_normal((M, dims[3]))
    return Ws, bs, x, t


def test_backward_matches_finite_differences():
    Ws, bs, x, t = _tiny()
    out, cache = mlp.forward(x, Ws, bs, rnd=False)
    _, d = mlp.mse(out, t)
    dWs, dbs = mlp.backward(d, Ws, cache, rnd=False)
    h = 1e-6

    def loss():
        o, _ = mlp.forward(x, Ws, bs, rnd=False)
        return mlp.mse(o, t)[0]

    for l in range(3):
        for P, dP in ((Ws[l], dWs[l]), (bs[l], dbs[l])):
            num = np.zeros_like(P)
            for idx in np.ndindex(P.shape):
                old = P[idx]
                P[idx] = old + h
                lp = loss()
                P[idx] = old - h
                lm = loss()
                P[idx] = old
                num[idx] = (lp - lm) / (2 * h)
            assert np.allclose(dP, num, rtol=1e-6, atol=1e-8), (l, np.abs(dP - num).max())


def test_gelu_grad_is_the_derivative():
    z = np.linspace(-4, 4, 81)
    h = 1e-6
    num = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.allclose(mlp.gelu_grad(z), num, atol=1e-8)


def test_bf16_rounding_known_values():
    ulp = 2.0 ** -7                      # spacing of BF16 just above 1
    vals = np.array([1.0, 1 + ulp / 2, 1 + 3 * ulp / 2, 1 + ulp / 2 + 2 ** -20, -2.5, 3.0e38])
    got = mlp.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0            # tie -> even (down)
    assert got[2] == 1 + 2 * ulp                      # tie -> even (up)
    assert got[3] == 1 + ulp                          # above the tie -> up
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8 and np.isfinite(got[5])


def test_adam_first_step_moves_by_lr():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([3.0, -0.01, 1e-3])
    m = np.zeros(3)
    v = np.zeros(3)
    w1, _, _ = mlp.adam(w, g, m, v, step=1, lr=0.1)
    assert np.allclose(w - w1, 0.1 * np.sign(g), rtol=1e-4)


def test_tdl_jakes_statistics():
    ch = synth.TdlJakes(8, 4, 624, 12, seed=3)
    assert abs(ch.power.sum() - 1) < 1e-12
    assert abs(ch.rms_delay_spread() - 300e-9) < 1e-12
    assert np.all(np.diff(ch.power) < 0)                  # exponential decay
    # slot-to-slot correlation of the path gains: J0(2 pi fd T) for Jakes
    rho = []
    for seed in range(4):
        c = synth.TdlJakes(16, 8, 624, 12, seed=seed)
        for t in range(0, 200, 20):
            h0, h1 = c.paths(t).ravel(), c.paths(t + 1).ravel()
            rho.append(np.real(np.vdot(h0, h1)) / np.linalg.norm(h0) / np.linalg.norm(h1))
    j0 = scipy.special.j0(2 * math.pi * 100 * 0.5e-3)
    assert abs(np.mean(rho) - j0) < 0.01, (np.mean(rho), j0)
    # mean channel power per (antenna, user) entry is 1
    a = np.stack([ch.slot(t) for t in range(0, 400, 40)])
    assert abs(np.mean(np.abs(a) ** 2) - 1) < 0.1
