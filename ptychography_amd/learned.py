"""Learned initial object: a 3x3-conv encoder-decoder run on the GPU's own
BF16 MFMA kernels (``csrc/nn.hip``), trained on seeded synthetic frames.

No counterpart in the reference: this is the initial-guess stage that feeds
``ERA`` / ``DM``.  The network maps a frame's centred sqrt-intensity
(64 x 64) to the complex object patch seen by that frame (real / imaginary
channels, gauge-fixed so the illumination-weighted patch mean is real and
positive -- a diffraction pattern cannot tell an object from the same object
times a global phase).  Patches are stitched into O0 by a weighted overlap
average that rotates each patch onto what is already stitched.  The plain
start stays the fallback and the judge: ``initial_object`` keeps the learned
O0 only if its first-iteration eMod beats the flat start's.

Layers (NHWC, bf16 activations and weights, FP32 accumulate):
  enc  64^2: 1->32, 32->32 | pool | 32^2: 32->64, 64->64 | pool |
       16^2: 64->128, 128->128 | pool | 8^2: 128->256, 256->256
  dec  up | 16^2: 256->128, 128->128 | up | 32^2: 128->64, 64->64 | up |
       64^2: 64->32, 32->32, 32->2 (linear)
"""

import ctypes
import time

import numpy as np
import torch

from . import _abi, sim
from .engine import Engine, require_gpu

_vp = ctypes.c_void_p


def _p(t):
    return _vp(t.data_ptr()) if t is not None else None


LAYERS = [("conv", 1, 32), ("conv", 32, 32), ("pool",), ("conv", 32, 64), ("conv", 64, 64),
          ("pool",), ("conv", 64, 128), ("conv", 128, 128), ("pool",), ("conv", 128, 256),
          ("conv", 256, 256), ("up",), ("conv", 256, 128), ("conv", 128, 128), ("up",),
          ("conv", 128, 64), ("conv", 64, 64), ("up",), ("conv", 64, 32), ("conv", 32, 32),
          ("conv", 32, 2)]


class InitNet:
    """The encoder-decoder: parameters, forward, backward and Adam on the GPU."""

    def __init__(self, frame=64, seed=0, device=None, layers=LAYERS):
        self.dev = require_gpu(device)
        self.frame = frame
        self.layers = layers
        rng = np.random.default_rng(seed)
        self.params = []   # per conv: dict(w32, wb, b, gw, gb, mw, vw, mb, vb, cin, cout, relu)
        n_conv = sum(1 for l in layers if l[0] == "conv")
        k = 0
        for l in layers:
            if l[0] != "conv":
                continue
            k += 1
            cin, cout = l[1], l[2]
            std = np.sqrt(2.0 / (9 * cin))
            w = (rng.standard_normal((9 * cin, cout)) * std).astype(np.float32)
            d = dict(cin=cin, cout=cout, relu=(k < n_conv))
            d["w32"] = torch.from_numpy(w).to(self.dev)
            d["wb"] = d["w32"].to(torch.bfloat16)
            d["b"] = torch.zeros(cout, dtype=torch.float32, device=self.dev)
            for name, shape in (("gw", (9 * cin, cout)), ("gb", (cout,))):
                d[name] = torch.zeros(shape, dtype=torch.float32, device=self.dev)
            for name in ("mw", "vw"):
                d[name] = torch.zeros((9 * cin, cout), dtype=torch.float32, device=self.dev)
            for name in ("mb", "vb"):
                d[name] = torch.zeros(cout, dtype=torch.float32, device=self.dev)
            self.params.append(d)
        self.step_count = 0
        self._part = torch.zeros(1024, dtype=torch.float64, device=self.dev)
        self._ws = None

    # -- helpers --------------------------------------------------------------
    def _stream(self):
        return _vp(torch.cuda.current_stream(self.dev).cuda_stream)

    def _ws_for(self, n):
        if self._ws is None or self._ws.numel() < n:
            self._ws = torch.empty(n, dtype=torch.float32, device=self.dev)
        return self._ws

    def state_dict(self):
        return {f"w{i}": p["w32"].cpu() for i, p in enumerate(self.params)} | \
               {f"b{i}": p["b"].cpu() for i, p in enumerate(self.params)}

    def load_state_dict(self, sd):
        for i, p in enumerate(self.params):
            p["w32"].copy_(sd[f"w{i}"].to(self.dev))
            p["b"].copy_(sd[f"b{i}"].to(self.dev))
            p["wb"] = p["w32"].to(torch.bfloat16)

    # -- forward / backward ---------------------------------------------------
    def forward(self, x, keep=False):
        """x: (B, H, W, 1) bf16 -> (B, H, W, 2) fp32.  keep: save activations for backward."""
        s = self._stream()
        B, H, W, C = x.shape
        acts = [x]
        pi = 0
        out_f = None
        for l in self.layers:
            cur = acts[-1]
            if l[0] == "conv":
                p = self.params[pi]
                pi += 1
                last = pi == len(self.params)
                y = None if last else torch.empty((B, H, W, p["cout"]), dtype=torch.bfloat16,
                                                  device=self.dev)
                yf = torch.empty((B, H, W, p["cout"]), dtype=torch.float32, device=self.dev) \
                    if last else None
                _abi.call("pty_nn_conv3x3_fwd", _p(cur), _p(p["wb"]), _p(p["b"]), _p(y), _p(yf),
                          B, H, W, p["cin"], p["cout"], int(p["relu"]), s)
                if last:
                    out_f = yf
                    acts.append(None)
                else:
                    acts.append(y)
            elif l[0] == "pool":
                y = torch.empty((B, H // 2, W // 2, cur.shape[3]), dtype=torch.bfloat16, device=self.dev)
                _abi.call("pty_nn_pool2", _p(cur), _p(y), B, H, W, cur.shape[3], 0, s)
                H, W = H // 2, W // 2
                acts.append(y)
            elif l[0] == "up":
                y = torch.empty((B, 2 * H, 2 * W, cur.shape[3]), dtype=torch.bfloat16, device=self.dev)
                _abi.call("pty_nn_up2", _p(cur), _p(y), B, H, W, cur.shape[3], 0, s)
                H, W = 2 * H, 2 * W
                acts.append(y)
            if not keep and len(acts) > 2:
                acts[-2] = None
        self._acts = acts if keep else None
        return out_f

    def backward(self, dy):
        """dy: (B, H, W, 2) bf16 gradient of the loss w.r.t. the output; fills gw / gb."""
        s = self._stream()
        acts = self._acts
        B = dy.shape[0]
        pi = len(self.params)
        g = dy
        for li in range(len(self.layers) - 1, -1, -1):
            l = self.layers[li]
            x = acts[li]
            H, W = x.shape[1], x.shape[2]
            if l[0] == "conv":
                pi -= 1
                p = self.params[pi]
                if p["relu"]:
                    _abi.call("pty_nn_relu_bwd", _p(g), _p(acts[li + 1]), g.numel(), s)
                P = B * H * W
                splits = int(max(1, min(32, P // 8192)))
                ws = self._ws_for(splits * 9 * p["cin"] * p["cout"])
                _abi.call("pty_nn_conv3x3_wgrad", _p(x), _p(g), _p(p["gw"]), _p(p["gb"]), _p(ws),
                          splits, B, H, W, p["cin"], p["cout"], s)
                if li > 0:
                    dx = torch.empty_like(x)
                    _abi.call("pty_nn_conv3x3_dgrad", _p(g), _p(p["wb"]), _p(dx), B, H, W,
                              p["cin"], p["cout"], s)
                    g = dx
            elif l[0] == "pool":
                dx = torch.empty_like(x)
                _abi.call("pty_nn_pool2", _p(g), _p(dx), B, H, W, x.shape[3], 1, s)
                g = dx
            elif l[0] == "up":
                dx = torch.empty_like(x)
                _abi.call("pty_nn_up2", _p(g), _p(dx), B, H, W, x.shape[3], 1, s)
                g = dx
        self._acts = None

    def loss(self, pred, target, weight=None, grad=True, gscale=1.0):
        """Weighted MSE (mean over all outputs); returns (loss, dy bf16 or None)."""
        s = self._stream()
        n = pred.numel()
        dy = torch.empty(pred.shape, dtype=torch.bfloat16, device=self.dev) if grad else None
        period = weight.numel() if weight is not None else 1
        _abi.call("pty_nn_mse", _p(pred), _p(target), n, _p(dy), float(gscale), _p(self._part),
                  _p(weight), period, s)
        return float(self._part.sum().item()) / n, dy

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, gscale=1.0):
        s = self._stream()
        self.step_count += 1
        for p in self.params:
            _abi.call("pty_nn_adam", _p(p["w32"]), _p(p["wb"]), _p(p["gw"]), _p(p["mw"]), _p(p["vw"]),
                      p["w32"].numel(), lr, b1, b2, eps, self.step_count, float(gscale), s)
            _abi.call("pty_nn_adam", _p(p["b"]), None, _p(p["gb"]), _p(p["mb"]), _p(p["vb"]),
                      p["b"].numel(), lr, b1, b2, eps, self.step_count, float(gscale), s)


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------

def illumination_weight(P):
    """Per-pixel loss / stitching weight: |P|^2 normalised to max 1."""
    w = np.abs(P) ** 2
    return (w / w.max()).astype(np.float32)


def gauge_fix(patches, w):
    """Rotate each complex patch so that sum(w * patch) is real and positive."""
    c = (patches * w[None]).sum(axis=(1, 2))
    ph = np.where(np.abs(c) > 0, c / np.maximum(np.abs(c), 1e-30), 1.0)
    return patches * np.conj(ph)[:, None, None]


def input_scale(photons, frame):
    return float(np.sqrt(photons / (frame * frame)))


def training_frames(n_frames, seed=5, frame=64, pupil_radius=16, defocus=0.02, photons=1e6,
                    obj_size=384, per_object=1024, device=None, poisson=True):
    """Seeded frames from random smooth objects (the learned-init recipe).

    Returns x (n, frame, frame, 1) bf16 inputs, t (n, frame, frame, 2) fp32
    gauge-fixed (re, im) targets, the probe and the pixel weight, all on device.
    Intensities are formed by the engine's own frame kernels.
    """
    dev = require_gpu(device)
    rng = np.random.default_rng(seed)
    P = sim.make_probe(frame, pupil_radius, defocus, 0.0, rng)
    w = illumination_weight(P)
    xs, ts = [], []
    made = 0
    while made < n_frames:
        k = min(per_object, n_frames - made)
        O = sim.make_object((obj_size, obj_size), rng, (0.3, 1.0), 4.0)
        off = np.stack([rng.integers(0, obj_size - frame + 1, k),
                        rng.integers(0, obj_size - frame + 1, k)], 1)
        R = (-off).astype(np.int32)
        eng = Engine(k, frame, frame, obj_size, obj_size, R, device=dev)
        O_t = torch.from_numpy(O.astype(np.complex64)).to(dev)
        P_t = torch.from_numpy(P.astype(np.complex64)).to(dev)
        eng.set_state(O_t, P_t)
        I = eng.intensity()
        eng.close()
        I *= photons / float(I.sum(dtype=torch.float64).item() / k)
        if poisson:
            I = torch.poisson(I)
        xs.append((torch.sqrt(I) / input_scale(photons, frame)).to(torch.bfloat16)[..., None])
        patches = np.stack([O[y:y + frame, x:x + frame] for y, x in off])
        patches = gauge_fix(patches, w)
        ts.append(torch.from_numpy(np.stack([patches.real, patches.imag], -1).astype(np.float32)).to(dev))
        made += k
    return torch.cat(xs), torch.cat(ts), P, w


def train(net, x, t, w, steps=2000, batch=64, lr=2e-3, seed=0, log_every=0, time_budget=None):
    """Adam on mini-batches drawn with a seeded permutation; returns the loss history."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    wt = torch.from_numpy(np.repeat(w.ravel(), 2)).to(net.dev)   # weight per (pixel, channel)
    hist = []
    t0 = time.time()
    n = x.shape[0]
    perm = torch.randperm(n, generator=g)
    pos = 0
    for step in range(steps):
        if pos + batch > n:
            perm = torch.randperm(n, generator=g)
            pos = 0
        idx = perm[pos:pos + batch].to(net.dev)
        pos += batch
        xb = x.index_select(0, idx).contiguous()
        tb = t.index_select(0, idx).contiguous()
        lr_s = lr * min(1.0, (step + 1) / 100) * (0.5 * (1 + np.cos(np.pi * step / steps)) * 0.9 + 0.1)
        pred = net.forward(xb, keep=True)
        loss, dy = net.loss(pred, tb, wt, grad=True, gscale=1024.0)
        net.backward(dy)
        net.adam(lr=lr_s, gscale=1024.0)
        hist.append(loss)
        if log_every and step % log_every == 0:
            print(f"step {step} loss {loss:.4e} ({time.time() - t0:.1f}s)", flush=True)
        if time_budget is not None and time.time() - t0 > time_budget:
            break
    return hist


def predict_patches(net, I, photons, batch=256):
    """Complex patches (n, M, K) complex64 from centred intensities I (n, M, K) on device."""
    frame = I.shape[1]
    out = []
    for s in range(0, I.shape[0], batch):
        xb = (torch.sqrt(I[s:s + batch].to(torch.float32).clamp_min(0)) /
              input_scale(photons, frame)).to(torch.bfloat16)[..., None].contiguous()
        yb = net.forward(xb)
        out.append(torch.complex(yb[..., 0], yb[..., 1]))
    return torch.cat(out).contiguous()


def stitch(patches, R, U, V, weight=None, floor=1e-3, fill=1.0):
    """Weighted, phase-aligned overlap average of patches at R's windows (R <= 0)."""
    dev = patches.device
    N, M, K = patches.shape
    acc = torch.empty((U, V), dtype=torch.complex64, device=dev)
    wsum = torch.empty((U, V), dtype=torch.float32, device=dev)
    O = torch.empty((U, V), dtype=torch.complex64, device=dev)
    wt = torch.from_numpy(np.ascontiguousarray(weight, np.float32)).to(dev) if weight is not None else None
    Rc = np.ascontiguousarray(R, np.int32)
    torch.cuda.current_stream(dev).synchronize()
    _abi.call("pty_nn_stitch", _p(patches), _p(wt), Rc.ctypes.data_as(_vp), N, M, K, U, V,
              float(floor), float(fill), _p(acc), _p(wsum), _p(O),
              _vp(torch.cuda.current_stream(dev).cuda_stream))
    return O


def initial_object(net, I, R, P, U, V, photons, mask=None, alpha=1e-10, device=None):
    """Learned O0 if its eMod beats the flat start's, else the flat object.

    Returns (O0 complex64 numpy, info dict with both starting eMods and the
    choice).  The eMod of a start is the reference's first-iteration eMod
    (ptychography/era.py:561-563, 210) of exits = O0 * P.
    """
    from .era import ALGO_ERA
    dev = require_gpu(device)
    I_t = I if isinstance(I, torch.Tensor) else torch.from_numpy(np.asarray(I, np.float32))
    I_t = I_t.to(dev, torch.float32)
    Rs = np.asarray(R).copy()
    Rs -= Rs.max(axis=0)
    t0 = time.time()
    patches = predict_patches(net, I_t, photons)
    O_l = stitch(patches, Rs, U, V, illumination_weight(P))
    torch.cuda.synchronize()
    t_init = time.time() - t0

    def first_emod(O_start):
        N, M, K = I_t.shape
        eng = Engine(N, M, K, U, V, Rs, device=dev)
        mc = None
        ms = None
        if mask is not None and not (np.isscalar(mask) and mask == 1):
            mc = torch.from_numpy(np.asarray(mask, np.uint8)).to(dev)
            ms = torch.from_numpy(np.fft.ifftshift(np.asarray(mask, np.uint8))).to(dev)
        amp, I_norm = eng.prepare_amp(I_t, mc)
        eng.set_data(amp, ms, I_norm)
        O_t = O_start.clone()
        P_t = torch.from_numpy(np.asarray(P, np.complex64)).to(dev)
        eng.set_state(O_t, P_t)
        m = eng.run(ALGO_ERA, 1, 1, begin=True, alpha=alpha)
        eng.close()
        return float(m[0, 0])

    flat = torch.ones((U, V), dtype=torch.complex64, device=dev)
    e_learned = first_emod(O_l)
    e_flat = first_emod(flat)
    use = e_learned < e_flat
    info = dict(eMod_learned_start=e_learned, eMod_flat_start=e_flat, used_learned=bool(use),
                init_seconds=t_init)
    return (O_l if use else flat).cpu().numpy(), info
