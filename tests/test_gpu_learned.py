"""The learned form's kernels on cuda:0 against the CPU oracles.

Network kernels vs oracle/mlp.py (same BF16-rounded operands, FP64 there):
outputs and gradients within 1e-2 relative Frobenius error (FP32
accumulation order and rare BF16 rounding flips of the hidden activations),
Adam within 1e-5. Inverse kernels vs FP64 numpy: Gram within 1e-6, inverses
within 1e-4, the first-order warm start within 1e-4, W of fallback groups
bit-identical to mega_zf."""
import numpy as np
import pytest

import megastation_amd as m
from megastation_amd import learned, synth
from oracle import mlp

from gpu_helpers import assert_imag_counts, dev_c64, host_c64, rel_fro, torch

pytestmark = pytest.mark.gpu


def _bf(t_np):
    t = torch()
    return t.from_numpy(np.asarray(t_np, dtype=np.float32)).cuda().to(t.bfloat16)


_rel = rel_fro  # complex arrays compare as complex128


@pytest.mark.parametrize("M,N,K", [(70, 45, 33), (64, 64, 64), (5, 130, 300)])
def test_gemm_strided_transposes(M, N, K):
    print(f"case: gemm M={M} N={N} K={K}", flush=True)
    t = torch()
    rng = np.random.default_rng(M + N + K)
    A = mlp.bf16(rng.standard_normal((M, K)))
    B = mlp.bf16(rng.standard_normal((K, N)))
    bias = rng.standard_normal(N).astype(np.float32)
    ref = A @ B + bias
    for a_t in (False, True):
        for b_t in (False, True):
            Ad = _bf(A.T.copy() if a_t else A)
            Bd = _bf(B.T.copy() if b_t else B)
            sai, sak = (1, M) if a_t else (K, 1)
            sbk, sbj = (1, K) if b_t else (N, 1)
            C = t.empty(M, N, device="cuda")
            Cb = t.empty(M, N, dtype=t.bfloat16, device="cuda")
            learned.gemm(Ad, sai, sak, Bd, sbk, sbj, M, N, K, bias=t.from_numpy(bias).cuda(),
                         act=learned.GELU, Cf=C, Cb=Cb)
            t.cuda.synchronize()
            assert _rel(C.cpu().numpy(), ref) < 1e-6, (a_t, b_t)
            assert _rel(Cb.float().cpu().numpy(), mlp.gelu(ref)) < 1e-2


def test_mlp_forward_backward_adam_match_oracle():
    print("case: mlp fwd/bwd/adam", flush=True)
    t = torch()
    U = 4
    net = learned.WarmStartNet(U, hidden=64, seed=5)
    rng = np.random.default_rng(11)
    M = 96
    x = mlp.bf16(rng.standard_normal((M, net.dims[0])))
    tgt = rng.standard_normal((M, net.dims[3])).astype(np.float32) * 0.1
    Ws = [w.cpu().numpy().astype(np.float64) for w in net.w]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    out = net.forward(_bf(x), keep=True)
    ref_out, cache = mlp.forward(x, Ws, bs)
    assert _rel(out.cpu().numpy(), ref_out) < 1e-2
    d = net.mse_grad(out, t.from_numpy(tgt).cuda())
    _, dref = mlp.mse(ref_out, tgt)
    assert _rel(d.cpu().numpy(), dref) < 1e-2
    grads = net.backward(d)
    dWs, dbs = mlp.backward(d.cpu().numpy().astype(np.float64), Ws, cache)
    for l in range(3):
        assert _rel(grads[l].cpu().numpy(), dWs[l]) < 1e-2, l
        assert _rel(grads[3 + l].cpu().numpy(), dbs[l]) < 1e-2, l
    g0 = [g.cpu().numpy().astype(np.float64) for g in grads]
    net.adam(lr=1e-2)
    t.cuda.synchronize()
    for l in range(3):
        w1, _, _ = mlp.adam(Ws[l], g0[l], 0 * Ws[l], 0 * Ws[l], 1, 1e-2)
        assert _rel(net.w[l].cpu().numpy(), w1) < 1e-5
        assert np.array_equal(net.wb[l].float().cpu().numpy(),
                              mlp.bf16(net.w[l].cpu().numpy()).astype(np.float32))


def _groups(seed, G, B, U):
    rng = np.random.Generator(np.random.Philox(seed))
    return np.conj(synth.channel_iid(rng, G, B, U)).astype(np.complex64)   # A [G][B][U]


def _gram64(a, G, B, U):
    A = a.reshape(G, B, U).astype(np.complex128)
    return np.matmul(np.swapaxes(A, 1, 2), A.conj())        # [g][u][v] = sum_b A[u,b] conj(A[v,b])


@pytest.mark.parametrize("U,B", [(8, 24), (16, 64), (32, 128), (5, 40), (48, 96)])
def test_gram_inv_matches_fp64(U, B):
    print(f"case: gram_inv U={U} B={B}", flush=True)
    G = 9
    a = _groups(U * B, G, B, U)
    Gm, X = learned.gram_inv(dev_c64(a), U, B, G)
    Gn = host_c64(Gm).astype(np.complex128)
    ref = _gram64(a, G, B, U)
    assert_imag_counts(ref)
    assert _rel(Gn, ref) < 1e-6
    Xn = host_c64(X).astype(np.complex128)
    assert_imag_counts(np.linalg.inv(ref))
    assert _rel(Xn, np.linalg.inv(ref)) < 1e-4
    Gonly, none = learned.gram_inv(dev_c64(a), U, B, G, want_x=False)
    assert none is None and np.array_equal(host_c64(Gonly), host_c64(Gm))


@pytest.mark.parametrize("U,B", [(32, 128), (12, 48), (40, 160)])
def test_refine_first_order_newton_schulz_and_residual_check(U, B):
    print(f"case: refine U={U} B={B}", flush=True)
    G = 12
    ch = synth.TdlJakes(B, U, G * 12, 12, seed=2)
    a0, a1 = ch.slot(10), ch.slot(11)
    G0, X0 = learned.gram_inv(dev_c64(a0), U, B, G)
    G1, X1 = learned.gram_inv(dev_c64(a1), U, B, G)
    g0 = _gram64(a0, G, B, U)
    g1 = _gram64(a1, G, B, U)
    x0 = np.linalg.inv(g0)
    # no correction, no iterations: the first-order update
    Xfo, res, flag = learned.refine(G1, G0, X0, U, iters=0)
    fo = x0 - x0 @ (g1 - g0) @ x0
    assert_imag_counts(fo)
    assert _rel(host_c64(Xfo), fo) < 1e-4
    e_fo = np.linalg.norm(fo @ g1 - np.eye(U), axis=(1, 2))
    assert np.allclose(res.cpu().numpy(), e_fo, rtol=1e-3)
    # Newton-Schulz: the residual squares per step (contraction from e < 1)
    prev = e_fo
    for it in (1, 2, 3):
        _, r, _ = learned.refine(G1, G0, X0, U, iters=it)
        r = r.cpu().numpy()
        ok = prev < 0.5
        assert np.all(r[ok] <= prev[ok] ** 2 * U + 1e-4), it
        prev = r
    # exact previous inverse of an unchanged channel: converged, not flagged
    Xs, r, f = learned.refine(G0, G0, X0, U, iters=2, bound=1e-3)
    assert r.cpu().numpy().max() < 1e-3 and not f.cpu().numpy().any()
    # a bound of 0 flags everything
    _, _, f = learned.refine(G1, G0, X0, U, iters=2, bound=0.0)
    assert f.cpu().numpy().all()


@pytest.mark.parametrize("U,B", [(32, 128), (12, 48), (40, 160)])
def test_finish_fallback_and_apply(U, B):
    print(f"case: finish U={U} B={B}", flush=True)
    t = torch()
    G = 10
    a = _groups(5, G, B, U)
    d_a = dev_c64(a)
    Gm, X = learned.gram_inv(d_a, U, B, G)
    flag = t.from_numpy(np.array([1, 0] * (G // 2), dtype=np.int32)).cuda()
    W = t.empty(G * B * U, dtype=t.complex64, device="cuda")
    info = t.full((G,), -1, dtype=t.int32, device="cuda")
    Xs = X.clone()
    learned.finish(d_a, Xs, flag, U, B, W, info)
    W_zf = t.empty_like(W)
    m.zf(d_a, W_zf, B, U, G)
    t.cuda.synchronize()
    w = host_c64(W).reshape(G, B * U)
    wz = host_c64(W_zf).reshape(G, B * U)
    f = flag.cpu().numpy().astype(bool)
    assert np.array_equal(w[f], wz[f])                       # fallback = plain ZF
    # the slot's stored Gram matrix instead of a recomputed one: same W
    W2 = t.empty_like(W)
    learned.finish(d_a, X.clone(), flag, U, B, W2, info, G=Gm)
    t.cuda.synchronize()
    assert np.array_equal(host_c64(W2), host_c64(W))
    # unflagged: W = X A with the given X
    A = a.reshape(G, B, U).astype(np.complex128)
    xn = host_c64(X).astype(np.complex128)
    ref = np.einsum("guv,gbv->gbu", xn, A).reshape(G, B * U)
    assert_imag_counts(ref[~f])
    assert _rel(w[~f], ref[~f]) < 1e-5
    assert not info.cpu().numpy().any()


def test_checkpoint_loads_into_the_same_network(tmp_path):
    """save -> load gives a network whose forward pass is bit-identical."""
    print("case: learned checkpoint", flush=True)
    t = torch()
    U = 4
    net = learned.WarmStartNet(U, hidden=64, seed=3)
    x = _bf(np.random.default_rng(1).standard_normal((5, learned.n_features(U))))
    ref = net.forward(x).cpu().numpy()
    p = str(tmp_path / "n.npz")
    learned.save(net, p, {"note": "test"})
    net2, info = learned.load(p, seed=9)
    assert info["note"] == "test" and info["U"] == U
    assert np.array_equal(net2.forward(x).cpu().numpy(), ref)
    t.cuda.synchronize()


def test_trained_network_reduces_fallbacks_on_held_out_sequence():
    """The network trained on two generator sequences (the committed
    checkpoint of megastation_amd.train_learned: 200k slot pairs, the
    network's own kernels, loaded with its SHA-256 checked; trained here on
    200k group samples if the file is absent), run in the learned ZF over 40
    held-out slots of a third sequence, with and without the network: the
    network must cut both the median residual |X G - I|_F after two
    Newton-Schulz steps and the share of groups that fall back to the plain
    ZF; W of every group stays within the residual check's bound of the
    plain ZF's W."""
    import os
    from megastation_amd import train_learned
    print("case: learned gain", flush=True)
    t = torch()
    U, B, N, sg = 32, 128, 3276, 12
    test = learned.GpuChannel(synth.TdlJakes(B, U, N, sg, seed=99))
    if os.path.exists(train_learned.DEFAULT_OUT):
        net, info = learned.load(train_learned.DEFAULT_OUT)
        print(f"loaded {train_learned.DEFAULT_OUT}: {info}", flush=True)
        assert info["train_slot_pairs"] >= 200_000 and info["loss_last"] < 0.7 * info["loss_first"]
    else:
        chans = [learned.GpuChannel(synth.TdlJakes(B, U, N, sg, seed=s)) for s in (11, 12)]
        net = learned.WarmStartNet(U, seed=0)
        hist = learned.train(net, chans, samples=200_000)
        assert hist[-1][1] < 0.7 * hist[0][1]
    G = test.ch.G
    stats = {}
    for label, nx in (("without", None), ("with", net)):
        z = learned.LearnedZF(U, B, G, nx)
        res, fb, werr = [], [], []
        for ts in range(500, 541):
            A = test.slot(ts)
            W, info = z.step(A)
            if z.last.get("residual") is None:
                continue
            res.append(z.last["residual"].cpu().numpy())
            fb.append(z.last["flag"].cpu().numpy())
            if ts % 10 == 0:   # the residual check: plain ZF of the same slot
                Wz = t.empty_like(W)
                m.zf(A, Wz, B, U, G)
                a, b = host_c64(W).reshape(G, -1), host_c64(Wz).reshape(G, -1)
                werr.append(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1))
            assert not info.cpu().numpy().any()
        stats[label] = (np.median(np.concatenate(res)), np.concatenate(fb).mean(),
                        np.concatenate(werr).max())
    print(stats, flush=True)
    assert stats["with"][0] < 0.7 * stats["without"][0]
    assert stats["with"][1] < 0.5 * stats["without"][1]
    # |X G - I|_F <= 1e-3 bounds the relative error of W against the exact ZF
    assert stats["with"][2] < 5e-3 and stats["without"][2] < 5e-3
