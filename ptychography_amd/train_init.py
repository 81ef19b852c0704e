"""Train the learned-initialiser network on the GPU and save its weights.

    python -m ptychography_amd.train_init --out init64.npz [--frames 60000] [--seconds 300]

Training frames come from seeded random smooth objects (amplitude in
[0.3, 1], phase in [-pi, pi]) seen through the fixed radius-16 defocused
pupil at 10^6 photons per frame with Poisson noise (``learned.training_frames``),
each sample drawn under a random exact reflection / transpose; forward,
backward and Adam run on the library's own BF16 MFMA kernels.
"""

import argparse
import time

import numpy as np
import torch

from .learned import InitNet, save_weights, train, training_frames


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--frames", type=int, default=240000)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--per-object", type=int, default=128,
                    help="frames cut from each random object (fewer -> more objects)")
    ap.add_argument("--seconds", type=float, default=300.0)
    ap.add_argument("--steps", type=int, default=10 ** 7)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--log-every", type=int, default=1000)
    ap.add_argument("--no-augment", action="store_true")
    a = ap.parse_args(argv)
    dev = torch.device("cuda", torch.cuda.current_device())
    t0 = time.time()
    x, t, P, w = training_frames(a.frames, seed=a.seed, device=dev, per_object=a.per_object)
    torch.cuda.synchronize()
    print(f"{x.shape[0]} training frames in {time.time() - t0:.1f}s", flush=True)
    net = InitNet(seed=a.seed, device=dev)
    t0 = time.time()
    hist = train(net, x, t, w, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                 log_every=a.log_every, time_budget=a.seconds, augment=not a.no_augment)
    dt = time.time() - t0
    k = max(1, min(200, len(hist) // 10))
    print(f"{len(hist)} steps in {dt:.1f}s ({len(hist) * a.batch / dt:.0f} frames/s); "
          f"loss {np.mean(hist[:k]):.4e} -> {np.mean(hist[-k:]):.4e}", flush=True)
    save_weights(net, a.out)
    print(f"saved {a.out}", flush=True)


if __name__ == "__main__":
    main()
