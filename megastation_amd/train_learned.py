"""Trains the learned form's warm-start network on generator slots and writes
its checkpoint (weights + SHA-256, megastation_amd.learned.save).

    python -m megastation_amd.train_learned [--slots 200000] [--out PATH]

Cell: 128 antennas x 32 users, 3276 subcarriers in groups of 12 (273
groups), slot sequences of the project's channel generator (24-tap TDL,
exponential PDP with 300 ns RMS delay spread, Jakes Doppler at 100 Hz, 0.5 ms
slots; seeds 11 and 12, used round robin). A training sample is one group of
a pair of consecutive slots (features of slot t given slot t-1, target the
exact inverse's correction), so --slots slot pairs give slots x 273 samples.
Every step runs on the library's kernels on cuda:0 (gram_inv, features,
refine, target, the MLP's forward / backward / Adam).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "warmstart_128x32.npz")
CELL = dict(U=32, B=128, N=3276, sg=12)
SEEDS = (11, 12)


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--slots", type=int, default=200_000, help="slot pairs to train on")
    ap.add_argument("--batch-pairs", type=int, default=8, help="slot pairs per step")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0, help="initial weights")
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()

    import torch  # noqa: F401  (device memory and streams)
    from megastation_amd import learned, synth
    U, B, N, sg = (CELL[k] for k in ("U", "B", "N", "sg"))
    chans = [learned.GpuChannel(synth.TdlJakes(B, U, N, sg, seed=s)) for s in SEEDS]
    G = chans[0].ch.G
    samples = a.slots * G
    net = learned.WarmStartNet(U, seed=a.seed)
    # 1e-3 for the first half, then 3e-4, then 1e-4 for the last fifth
    lr_at = lambda seen: a.lr * (1.0 if seen < 0.5 * samples else
                                 0.3 if seen < 0.8 * samples else 0.1)
    t0 = time.perf_counter()
    last = [t0]

    def log(msg):
        now = time.perf_counter()
        if now - last[0] > 15:
            print(f"[train] {now - t0:.0f} s: {msg}", flush=True)
            last[0] = now
    hist = learned.train(net, chans, samples=samples, batch_pairs=a.batch_pairs, lr=a.lr,
                         log=log, lr_at=lr_at)
    secs = time.perf_counter() - t0
    info = {"cell": "128x32, 3276 subcarriers, groups of 12",
            "channel": "TDL-24, exponential PDP 300 ns RMS, Jakes 100 Hz, 0.5 ms slots",
            "channel_seeds": list(SEEDS), "train_slot_pairs": a.slots,
            "train_samples": hist[-1][0], "batch_pairs": a.batch_pairs, "lr": a.lr,
            "init_seed": a.seed, "loss_first": hist[0][1], "loss_last": hist[-1][1],
            "seconds": round(secs, 1)}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    digest = learned.save(net, a.out, info)
    print(f"[train] {a.slots} slot pairs ({hist[-1][0]} samples) in {secs:.1f} s, loss "
          f"{hist[0][1]:.4g} -> {hist[-1][1]:.4g}; wrote {a.out} (sha256 {digest})", flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(HERE))
    main()
