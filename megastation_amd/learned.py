"""Learned form of the ZF weights for slot sequences: a small BF16 MLP
predicts a correction to the first-order update of the previous slot's
inverse, two Newton-Schulz steps refine it, and groups whose residual
|W H - I|_F = |X G - I|_F exceeds the bound fall back to the plain FP32 ZF
(which also serves as the measure of the error).

Network: features (2*U*(U+1): upper triangles of (G_t - G_{t-1})/s and
X_{t-1}*s) -> 512 -> 512 -> U*(U+1) (upper triangle of the correction * s),
GELU hidden layers, BF16 weights and activations with FP32 accumulation and
FP32 master weights (Adam). Every step of training and inference runs on the
library's kernels (mega_mlp_*, mega_learned_*, mega_gram_inv); torch only
provides device memory and streams.

The channel sequences come from the project's seeded generator
(synth.TdlJakes), evaluated on the GPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import MegaError, _check, _ptr, _stream, lib

HIDDEN = 512
GELU, IDENTITY = 1, 0


class GemmDesc(ctypes.Structure):
    """Mirror of ``mega_mlp_gemm_desc`` (include/megadetect.h)."""
    _fields_ = [("A", ctypes.c_void_p), ("sai", ctypes.c_int64), ("sak", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("sbk", ctypes.c_int64), ("sbj", ctypes.c_int64),
                ("M", ctypes.c_uint32), ("N", ctypes.c_uint32), ("K", ctypes.c_uint32),
                ("bias", ctypes.c_void_p), ("act", ctypes.c_int32), ("Cf", ctypes.c_void_p),
                ("Cb", ctypes.c_void_p), ("ldc", ctypes.c_uint32), ("beta", ctypes.c_float)]


_sig_done = False


def _L():
    global _sig_done
    L = lib()
    if not _sig_done:
        P, u32, u64, i32, f32 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                                 ctypes.c_int32, ctypes.c_float)
        sig = {
            "mega_gram_inv": [u32, u32, u32, P, P, P, P, P],
            "mega_learned_features": [u32, u32, P, P, P, P, P, P],
            "mega_learned_target": [u32, u32, P, P, P, P, P],
            "mega_learned_refine": [u32, u32, P, P, P, P, u32, P, P, P, P, f32, u32, P],
            "mega_learned_finish": [u32, u32, u32, P, P, P, P, P, P, P],
            "mega_mlp_gemm": [ctypes.POINTER(GemmDesc), P],
            "mega_mlp_act_backward": [P, P, P, P, u64, i32, P],
            "mega_mlp_col_sum": [P, P, u32, u32, P],
            "mega_mlp_mse_grad": [P, P, P, u64, u32, P],
            "mega_mlp_adam": [P, P, P, P, P, u64, f32, f32, f32, f32, u32, P],
            "mega_mlp_to_bf16": [P, P, u64, P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _sig_done = True
    return L


def gemm(A, sai, sak, B, sbk, sbj, M, N, K, *, bias=None, act=IDENTITY, Cf=None, Cb=None,
         ldc=None, beta=0.0, stream=None):
    """C = A B (+bias, activation) on the BF16 matrix cores (mega_mlp_gemm)."""
    d = GemmDesc(_ptr(A).value, sai, sak, _ptr(B).value, sbk, sbj, M, N, K,
                 _ptr(bias).value, act, _ptr(Cf).value, _ptr(Cb).value,
                 N if ldc is None else ldc, beta)
    _check(_L().mega_mlp_gemm(ctypes.byref(d), _stream(stream)))


def n_features(U: int) -> int:
    return 2 * U * (U + 1)


def n_outputs(U: int) -> int:
    return U * (U + 1)


class WarmStartNet:
    """The MLP with its training state, on cuda:0."""

    def __init__(self, U: int, hidden: int = HIDDEN, seed: int = 0):
        import torch
        self.U = U
        self.dims = [n_features(U), hidden, hidden, n_outputs(U)]
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w, self.b, self.wb = [], [], []
        for l in range(3):
            fan_in = self.dims[l]
            w = torch.randn(self.dims[l], self.dims[l + 1], generator=g) * math.sqrt(1.0 / fan_in)
            if l == 2:
                w *= 0.1  # start near the first-order update
            self.w.append(w.float().cuda().contiguous())
            self.b.append(torch.zeros(self.dims[l + 1], device="cuda"))
            self.wb.append(torch.empty(self.dims[l], self.dims[l + 1], dtype=torch.bfloat16,
                                       device="cuda"))
        self.params = self.w + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.step = 0
        self.sync_bf16()

    def sync_bf16(self):
        for w, wb in zip(self.w, self.wb):
            _check(_L().mega_mlp_to_bf16(_ptr(w), _ptr(wb), w.numel(), None))

    def forward(self, x_bf16, keep: bool = False, stream=None):
        """x: BF16 [M][dims[0]] -> FP32 [M][dims[3]] (the correction * s).
        keep=True saves what backward() needs."""
        import torch
        M = x_bf16.shape[0]
        h = x_bf16
        saved = [x_bf16]
        zs = []
        out = None
        bufs = None
        if not keep:  # inference: reuse buffers of this batch size
            bufs = self._ws.get(M) if hasattr(self, "_ws") else None
            if bufs is None:
                bufs = [(torch.empty(M, self.dims[l + 1], device="cuda"),
                         None if l == 2 else torch.empty(M, self.dims[l + 1], dtype=torch.bfloat16,
                                                         device="cuda")) for l in range(3)]
                if not hasattr(self, "_ws"):
                    self._ws = {}
                self._ws[M] = bufs
        for l in range(3):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            last = l == 2
            if bufs is not None:
                z, hb = bufs[l]
            else:
                z = torch.empty(M, n_out, device="cuda")
                hb = None if last else torch.empty(M, n_out, dtype=torch.bfloat16, device="cuda")
            gemm(h, n_in, 1, self.wb[l], n_out, 1, M, n_out, n_in, bias=self.b[l],
                 act=IDENTITY if last else GELU, Cf=z, Cb=hb, stream=stream)
            zs.append(z)
            if last:
                out = z
            else:
                saved.append(hb)
                h = hb
        if keep:
            self._saved = (saved, zs, M)
        return out

    def backward(self, dout, stream=None):
        """dout: FP32 [M][dims[3]] -> gradients in self.grads (FP32, like params)."""
        import torch
        saved, zs, M = self._saved
        gw, gb = [None] * 3, [None] * 3
        dy = dout
        for l in (2, 1, 0):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            dzb = torch.empty(M, n_out, dtype=torch.bfloat16, device="cuda")
            dzf = torch.empty(M, n_out, device="cuda")
            _check(_L().mega_mlp_act_backward(_ptr(dy), _ptr(zs[l]), _ptr(dzb), _ptr(dzf),
                                              M * n_out, IDENTITY if l == 2 else GELU,
                                              _stream(stream)))
            gw[l] = torch.empty(n_in, n_out, device="cuda")
            # dW = H^T dZ: A(i,k) = H[k][i]
            gemm(saved[l], 1, n_in, dzb, n_out, 1, n_in, n_out, M, Cf=gw[l], stream=stream)
            gb[l] = torch.empty(n_out, device="cuda")
            _check(_L().mega_mlp_col_sum(_ptr(dzf), _ptr(gb[l]), M, n_out, _stream(stream)))
            if l > 0:
                dy = torch.empty(M, n_in, device="cuda")
                # dH = dZ W^T: B(k,j) = W[j][k]
                gemm(dzb, n_out, 1, self.wb[l], 1, n_out, M, n_in, n_out, Cf=dy, stream=stream)
        self.grads = gw + gb
        return self.grads

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step += 1
        for k, (p, g, m, v) in enumerate(zip(self.params, self.grads, self.m, self.v)):
            wb = self.wb[k] if k < 3 else None
            _check(_L().mega_mlp_adam(_ptr(p), _ptr(wb), _ptr(g), _ptr(m), _ptr(v), p.numel(),
                                      lr, b1, b2, eps, self.step, _stream(stream)))

    def mse_grad(self, out, target, stream=None):
        import torch
        d = torch.empty_like(out)
        _check(_L().mega_mlp_mse_grad(_ptr(out), _ptr(target), _ptr(d), out.numel(),
                                      out.shape[0], _stream(stream)))
        return d

    def state_dict(self):
        return {f"w{l}": self.w[l].cpu().numpy() for l in range(3)} | \
               {f"b{l}": self.b[l].cpu().numpy() for l in range(3)}

    def load_state_dict(self, d):
        import torch
        for l in range(3):
            self.w[l].copy_(torch.from_numpy(np.asarray(d[f"w{l}"])))
            self.b[l].copy_(torch.from_numpy(np.asarray(d[f"b{l}"])))
        self.sync_bf16()


# ---------------------------------------------------------------------------
# Checkpoints: the FP32 master weights and biases of a WarmStartNet in a
# numpy .npz (no pickled objects) with the network's shape, what it was
# trained on, and a SHA-256 over every array's dtype, shape and bytes in a
# fixed order; read_checkpoint recomputes it and refuses a file that does not
# match.
# ---------------------------------------------------------------------------

CKPT_ARRAYS = ("w0", "w1", "w2", "b0", "b1", "b2")


class CheckpointError(ValueError):
    pass


def _digest(arrays: dict, meta: str) -> str:
    import hashlib
    h = hashlib.sha256()
    h.update(meta.encode())
    for k in CKPT_ARRAYS:
        a = np.ascontiguousarray(arrays[k])
        h.update(f"{k}:{a.dtype.str}:{a.shape}".encode())
        h.update(a.tobytes())
    return h.hexdigest()


def write_checkpoint(path: str, arrays: dict, info: dict) -> str:
    """Writes weights (w0..w2, b0..b2: float32) and info (JSON-able dict) to
    path (.npz); returns the SHA-256 stored with them."""
    import json
    meta = json.dumps(info, sort_keys=True)
    arrs = {k: np.ascontiguousarray(arrays[k], dtype=np.float32) for k in CKPT_ARRAYS}
    digest = _digest(arrs, meta)
    with open(path, "wb") as fh:
        np.savez_compressed(fh, meta=np.array(meta), sha256=np.array(digest), **arrs)
    return digest


def read_checkpoint(path: str) -> tuple[dict, dict]:
    """(arrays, info) of a checkpoint, after checking its SHA-256 (raises
    CheckpointError on a mismatch or a missing array). CPU only."""
    import json
    with np.load(path, allow_pickle=False) as z:
        missing = [k for k in CKPT_ARRAYS + ("meta", "sha256") if k not in z.files]
        if missing:
            raise CheckpointError(f"{path}: missing {missing}")
        arrs = {k: z[k] for k in CKPT_ARRAYS}
        meta = str(z["meta"])
        stored = str(z["sha256"])
    if _digest(arrs, meta) != stored:
        raise CheckpointError(f"{path}: SHA-256 mismatch (file corrupted or altered)")
    info = json.loads(meta)
    dims = info.get("dims")
    if dims is None or [arrs[f"w{l}"].shape for l in range(3)] != \
            [(dims[l], dims[l + 1]) for l in range(3)]:
        raise CheckpointError(f"{path}: weight shapes do not match dims {dims}")
    return arrs, info


def save(net: "WarmStartNet", path: str, info: dict | None = None) -> str:
    """Checkpoint of net (FP32 master weights) with its shape and `info`."""
    d = net.state_dict()
    meta = dict(info or {}, U=net.U, dims=net.dims)
    return write_checkpoint(path, d, meta)


def load(path: str, seed: int = 0) -> tuple["WarmStartNet", dict]:
    """A WarmStartNet with the checkpoint's weights (checksum verified)."""
    arrs, info = read_checkpoint(path)
    net = WarmStartNet(int(info["U"]), hidden=int(info["dims"][1]), seed=seed)
    if net.dims != list(info["dims"]):
        raise CheckpointError(f"{path}: dims {info['dims']} do not fit U={info['U']}")
    net.load_state_dict(arrs)
    return net, info


class GpuChannel:
    """synth.TdlJakes evaluated on the GPU: slot(t) -> A = H^H [G][B][U]
    complex64 on cuda:0 (same parameters and random draws as the CPU class)."""

    def __init__(self, ch):
        import torch
        self.ch = ch
        self.amp = torch.from_numpy(ch.amp).cuda()
        self.omega = torch.from_numpy(ch.omega).cuda()
        self.phi = torch.from_numpy(ch.phi).cuda()
        self.steer = torch.from_numpy(ch.steer.astype(np.complex128)).cuda()

    def slot(self, t: int):
        import torch
        ts = t * self.ch.slot_s
        ph = self.omega * ts + self.phi
        h = self.amp * torch.complex(torch.cos(ph), torch.sin(ph)).sum(dim=1)   # [P] c128
        h = h.reshape(self.ch.B * self.ch.U, self.ch.n_taps)
        H = (h @ self.steer).reshape(self.ch.B, self.ch.U, self.ch.G)          # [B][U][G]
        return torch.conj(H).permute(2, 0, 1).to(torch.complex64).contiguous()


def gram_inv(A, U: int, B: int, G: int, want_x: bool = True, stream=None, out=None):
    """(G_mat, X) per group, row-major [G][U][U] complex64 (X None if not
    wanted); out = (G_mat, X) buffers to fill."""
    import torch
    if out is not None:
        Gm, X = out
        X = X if want_x else None
    else:
        Gm = torch.empty(G, U, U, dtype=torch.complex64, device="cuda")
        X = torch.empty_like(Gm) if want_x else None
    _check(_L().mega_gram_inv(B, U, G, _ptr(A), _ptr(Gm), _ptr(X), None, _stream(stream)))
    return Gm, X


def features(Gm, Gp, Xp, U: int, stream=None, out=None):
    import torch
    n = Gm.shape[0]
    if out is not None:
        f, s = out
    else:
        f = torch.empty(n, n_features(U), dtype=torch.bfloat16, device="cuda")
        s = torch.empty(n, device="cuda")
    _check(_L().mega_learned_features(U, n, _ptr(Gm), _ptr(Gp), _ptr(Xp), _ptr(f), _ptr(s),
                                      _stream(stream)))
    return f, s


def refine(Gm, Gp, Xp, U: int, corr=None, scale=None, bound: float = 1e-3, iters: int = 2,
           stream=None, out=None):
    """Warm start + Newton-Schulz + residual check: (X, residual, flag)."""
    import torch
    n = Gm.shape[0]
    if out is not None:
        X, res, flag = out
    else:
        X = torch.empty_like(Gm)
        res = torch.empty(n, device="cuda")
        flag = torch.empty(n, dtype=torch.int32, device="cuda")
    _check(_L().mega_learned_refine(U, n, _ptr(Gm), _ptr(Gp), _ptr(Xp), _ptr(corr),
                                    n_outputs(U), _ptr(scale), _ptr(X), _ptr(res), _ptr(flag),
                                    bound, iters, _stream(stream)))
    return X, res, flag


def target(X_true, X_base, scale, U: int, stream=None):
    import torch
    n = X_true.shape[0]
    t = torch.empty(n, n_outputs(U), device="cuda")
    _check(_L().mega_learned_target(U, n, _ptr(X_true), _ptr(X_base), _ptr(scale), _ptr(t),
                                    _stream(stream)))
    return t


def finish(A, X, flag, U: int, B: int, w_out, info=None, stream=None, G=None):
    """W = X A; flagged groups take the plain sweep of their Gram matrix
    (G, this slot's from gram_inv, or recomputed from A when None)."""
    n = X.shape[0]
    _check(_L().mega_learned_finish(B, U, n, _ptr(A), _ptr(G), _ptr(X), _ptr(flag),
                                    _ptr(w_out), _ptr(info), _stream(stream)))


def training_pairs(chan: GpuChannel, t0: int, n_pairs: int, U: int, B: int):
    """Yield (features, target) for consecutive slot pairs t-1 -> t."""
    G = chan.ch.G
    Gp, Xp = gram_inv(chan.slot(t0), U, B, G)
    for t in range(t0 + 1, t0 + 1 + n_pairs):
        Gt, Xt = gram_inv(chan.slot(t), U, B, G)
        f, s = features(Gt, Gp, Xp, U)
        Xb, _, _ = refine(Gt, Gp, Xp, U, iters=0)
        yield f, target(Xt, Xb, s, U)
        Gp, Xp = Gt, Xt


def train(net: WarmStartNet, chans, samples: int = 200_000, batch_pairs: int = 4,
          lr: float = 1e-3, log=None, lr_at=None):
    """Trains on consecutive slot pairs of the given channels (round robin)
    until `samples` group samples have been seen. lr_at(seen) (optional)
    gives the learning rate as a function of the samples seen so far.
    Returns the loss history [(samples seen, loss)]."""
    import torch
    U, B = net.U, chans[0].ch.B
    G = chans[0].ch.G
    gens = [training_pairs(c, 0, 10 ** 9, U, B) for c in chans]
    losses = []
    seen, k = 0, 0
    while seen < samples:
        fs, ts = [], []
        for _ in range(batch_pairs):
            f, t = next(gens[k % len(gens)])
            k += 1
            fs.append(f)
            ts.append(t)
        x = torch.cat(fs)
        y = torch.cat(ts)
        out = net.forward(x, keep=True)
        d = net.mse_grad(out, y)
        net.backward(d)
        net.adam(lr=lr if lr_at is None else lr_at(seen))
        seen += x.shape[0]
        if len(losses) % 10 == 0 or seen >= samples:
            loss = float(0.5 * ((out - y) ** 2).sum(dim=1).mean())
            losses.append((seen, loss))
            if log:
                log(f"samples {seen}: loss {loss:.4g}")
    return losses


class LearnedZF:
    """Per-slot learned ZF over a sequence: call step(A) with each slot's CSI
    (cuda complex64 [G][B][U]); returns (W, info) and keeps the state. All
    buffers are allocated once (two state sets used alternately).
    net=None runs the same pipeline without the network correction."""

    def __init__(self, U: int, B: int, G: int, net: WarmStartNet | None, bound: float = 1e-3,
                 iters: int = 2):
        import torch
        self.U, self.B, self.G = U, B, G
        self.net, self.bound, self.iters = net, bound, iters
        c64 = dict(dtype=torch.complex64, device="cuda")
        self.Gs = [torch.empty(G, U, U, **c64) for _ in range(2)]
        self.Xs = [torch.empty(G, U, U, **c64) for _ in range(2)]
        self.feat = torch.empty(G, n_features(U), dtype=torch.bfloat16, device="cuda")
        self.scale = torch.empty(G, device="cuda")
        self.res = torch.empty(G, device="cuda")
        self.flag = torch.empty(G, dtype=torch.int32, device="cuda")
        self.ones = torch.ones(G, dtype=torch.int32, device="cuda")
        self.W = torch.empty(G * B * U, **c64)
        self.info = torch.empty(G, dtype=torch.int32, device="cuda")
        self.k = -1  # state slot of the previous step (-1: none)
        self.last = {}

    def step(self, A, stream=None):
        """W (the library-owned buffer, valid until the next step) and info."""
        U, B, G = self.U, self.B, self.G
        cur = 0 if self.k != 0 else 1
        Gt, X = self.Gs[cur], self.Xs[cur]
        if self.k < 0:  # first slot: the plain path
            gram_inv(A, U, B, G, stream=stream, out=(Gt, X))
            finish(A, X, self.ones, U, B, self.W, self.info, stream=stream)
            self.k = cur
            self.last = {"fallback": G, "residual": None}
            return self.W, self.info
        Gp, Xp = self.Gs[self.k], self.Xs[self.k]
        gram_inv(A, U, B, G, want_x=False, stream=stream, out=(Gt, X))
        corr = scale = None
        if self.net is not None:
            f, scale = features(Gt, Gp, Xp, U, stream=stream, out=(self.feat, self.scale))
            corr = self.net.forward(f, stream=stream)
        refine(Gt, Gp, Xp, U, corr, scale, self.bound, self.iters, stream=stream,
               out=(X, self.res, self.flag))
        finish(A, X, self.flag, U, B, self.W, self.info, stream=stream, G=Gt)
        self.k = cur
        self.last = {"flag": self.flag, "residual": self.res}
        return self.W, self.info
