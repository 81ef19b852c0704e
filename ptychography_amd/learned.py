"""Learned initial object: a 3x3-conv encoder-decoder run on the GPU's own
BF16 MFMA kernels (``csrc/nn.hip``), trained on seeded synthetic frames.

No counterpart in the reference: this is the initial-guess stage that feeds
``ERA`` / ``DM``.  The network maps a frame's centred sqrt-intensity
(64 x 64) to the complex object patch seen by that frame (real / imaginary
channels, gauge-fixed so the illumination-weighted patch mean is real and
positive -- a diffraction pattern cannot tell an object from the same object
times a global phase).  The patches' unknown phases are synchronised globally
from the pairwise overlap products of neighbouring windows (projected power
iterations, ``sync_phases``), and O0 is the engine's overlap projection of the
exits P * patch * phase (the |P|^2-weighted average).  The plain start stays
the fallback and the arbiter: ``learned_start`` keeps the learned O0 only if its
first-iteration eMod beats the plain start's.

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
            d["wT"] = torch.empty_like(d["wb"])
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
        self._refresh_transposed()

    # -- helpers --------------------------------------------------------------
    def _stream(self):
        return _vp(torch.cuda.current_stream(self.dev).cuda_stream)

    def _ws_for(self, n):
        if self._ws is None or self._ws.numel() < n:
            self._ws = torch.empty(n, dtype=torch.float32, device=self.dev)
        return self._ws

    def _refresh_transposed(self):
        """Per-tap transposed bf16 weights used by the forward direct convolution."""
        s = self._stream()
        for p in self.params:
            _abi.call("pty_nn_weight_t", _p(p["wb"]), _p(p["wT"]), p["cin"], p["cout"], s)

    def state_dict(self):
        return {f"w{i}": p["w32"].cpu() for i, p in enumerate(self.params)} | \
               {f"b{i}": p["b"].cpu() for i, p in enumerate(self.params)}

    def load_state_dict(self, sd):
        for i, p in enumerate(self.params):
            p["w32"].copy_(sd[f"w{i}"].to(self.dev))
            p["b"].copy_(sd[f"b{i}"].to(self.dev))
            p["wb"].copy_(p["w32"].to(torch.bfloat16))
        self._refresh_transposed()

    # -- forward / backward ---------------------------------------------------
    def forward(self, x, keep=False, intensity_scale=None):
        """x: (B, H, W, 1) bf16 -> (B, H, W, 2) fp32.  keep: save activations for backward.

        With ``intensity_scale`` x is instead (B, H, W) float32 centred
        intensities, and the first layer forms its input
        bf16(sqrt(max(x, 0)) / intensity_scale) as it reads them (inference)."""
        s = self._stream()
        from_I = intensity_scale is not None
        if from_I:
            assert not keep and x.dtype == torch.float32 and x.is_contiguous()
            B, H, W = x.shape
        else:
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
                if from_I and pi == 1:
                    assert p["cin"] == 1
                    _abi.call("pty_nn_conv3x3_fwd_intensity", _p(cur), float(intensity_scale),
                              _p(p["wb"]), _p(p["b"]), _p(y), _p(yf), B, H, W, p["cout"],
                              int(p["relu"]), s)
                else:
                    _abi.call("pty_nn_conv3x3_fwd", _p(cur), _p(p["wb"]), _p(p["wT"]), _p(p["b"]),
                              _p(y), _p(yf),
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
        """dy: (B, H, W, 2) bf16 gradient of the loss w.r.t. the output; fills gw / gb.

        The ReLU backward of a conv layer is fused into the dgrad of the conv
        that consumes its output when the two are adjacent."""
        s = self._stream()
        acts = self._acts
        B = dy.shape[0]
        pi = len(self.params)
        g = dy
        masked = False   # g already carries the ReLU mask of the layer being processed
        for li in range(len(self.layers) - 1, -1, -1):
            l = self.layers[li]
            x = acts[li]
            H, W = x.shape[1], x.shape[2]
            if l[0] == "conv":
                pi -= 1
                p = self.params[pi]
                if p["relu"] and not masked:
                    _abi.call("pty_nn_relu_bwd", _p(g), _p(acts[li + 1]), g.numel(), s)
                splits = _abi.load().pty_nn_wgrad_splits(B, H, W, p["cin"], p["cout"])
                ws = self._ws_for(splits * (9 * p["cin"] + 1) * p["cout"])
                _abi.call("pty_nn_conv3x3_wgrad", _p(x), _p(g), _p(p["gw"]), _p(p["gb"]), _p(ws),
                          splits, B, H, W, p["cin"], p["cout"], s)
                masked = False
                if li > 0:
                    prev_relu = self.layers[li - 1][0] == "conv" and self.params[pi - 1]["relu"]
                    dx = torch.empty_like(x)
                    _abi.call("pty_nn_conv3x3_dgrad", _p(g), _p(p["wb"]), _p(dx),
                              _p(x) if prev_relu else None, B, H, W, p["cin"], p["cout"], s)
                    g = dx
                    masked = prev_relu
            elif l[0] == "pool":
                dx = torch.empty_like(x)
                _abi.call("pty_nn_pool2", _p(g), _p(dx), B, H, W, x.shape[3], 1, s)
                g = dx
                masked = False
            elif l[0] == "up":
                dx = torch.empty_like(x)
                _abi.call("pty_nn_up2", _p(g), _p(dx), B, H, W, x.shape[3], 1, s)
                g = dx
                masked = False
        self._acts = None

    def loss(self, pred, target, weight=None, grad=True, gscale=1.0, sync=True):
        """Weighted MSE (mean over all outputs); returns (loss, dy bf16 or None).
        ``sync=False`` returns the loss as a 0-d device tensor (no host wait)."""
        s = self._stream()
        n = pred.numel()
        dy = torch.empty(pred.shape, dtype=torch.bfloat16, device=self.dev) if grad else None
        period = weight.numel() if weight is not None else 1
        _abi.call("pty_nn_mse", _p(pred), _p(target), n, _p(dy), float(gscale), _p(self._part),
                  _p(weight), period, s)
        total = self._part.sum() / n
        return (float(total.item()) if sync else total), dy

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, gscale=1.0):
        s = self._stream()
        self.step_count += 1
        for p in self.params:
            _abi.call("pty_nn_adam", _p(p["w32"]), _p(p["wb"]), _p(p["gw"]), _p(p["mw"]), _p(p["vw"]),
                      p["w32"].numel(), lr, b1, b2, eps, self.step_count, float(gscale), s)
            _abi.call("pty_nn_adam", _p(p["b"]), None, _p(p["gb"]), _p(p["mb"]), _p(p["vb"]),
                      p["b"].numel(), lr, b1, b2, eps, self.step_count, float(gscale), s)
        self._refresh_transposed()


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


def train(net, x, t, w, steps=2000, batch=64, lr=2e-3, seed=0, log_every=0, time_budget=None,
          augment=True):
    """Adam on mini-batches drawn with a seeded permutation; returns the loss history.

    Learning rate: linear warm-up over 100 steps, then cosine from ``lr`` to
    ``lr / 10`` over ``steps`` -- or over ``time_budget`` seconds when that is
    given (training then stops at the budget or at ``steps``).  ``augment``
    applies a random one of the 8 exact reflections / transposes to every
    sample (the probe of ``training_frames`` is symmetric under them)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    wt = torch.from_numpy(np.repeat(w.ravel(), 2)).to(net.dev)   # weight per (pixel, channel)
    losses = []
    torch.cuda.synchronize(net.dev)
    t0 = time.time()
    n = x.shape[0]
    perm = torch.randperm(n, generator=g)
    pos = 0
    for step in range(steps):
        elapsed = time.time() - t0
        if time_budget is not None and elapsed > time_budget:
            break
        frac = elapsed / time_budget if time_budget is not None else step / steps
        if pos + batch > n:
            perm = torch.randperm(n, generator=g)
            pos = 0
        idx = perm[pos:pos + batch].to(net.dev)
        pos += batch
        xb = torch.empty((batch,) + tuple(x.shape[1:]), dtype=x.dtype, device=net.dev)
        tb = torch.empty((batch,) + tuple(t.shape[1:]), dtype=t.dtype, device=net.dev)
        _abi.call("pty_nn_batch_gather", _p(x), _p(t), _p(idx), batch, x.shape[1], int(augment),
                  (seed * 1000003 + step) & 0xffffffff, _p(xb), _p(tb), net._stream())
        lr_s = lr * min(1.0, (step + 1) / 100) * (0.5 * (1 + np.cos(np.pi * min(frac, 1.0))) * 0.9 + 0.1)
        pred = net.forward(xb, keep=True)
        loss, dy = net.loss(pred, tb, wt, grad=True, gscale=1024.0, sync=False)
        net.backward(dy)
        net.adam(lr=lr_s, gscale=1024.0)
        losses.append(loss)
        if log_every and step % log_every == 0:
            print(f"step {step} loss {float(loss.item()):.4e} lr {lr_s:.2e} ({time.time() - t0:.1f}s)",
                  flush=True)
    return [float(v) for v in torch.stack(losses).cpu()] if losses else []


def save_weights(net, path):
    """Weights as bf16 bit patterns (what the forward pass uses) and fp32 biases."""
    out = {"frame": np.int32(net.frame), "layers": np.array([",".join(map(str, l)) for l in net.layers])}
    for i, p in enumerate(net.params):
        out[f"w{i}"] = p["wb"].view(torch.int16).cpu().numpy().view(np.uint16)
        out[f"b{i}"] = p["b"].cpu().numpy()
    np.savez_compressed(path, **out)


def load_weights(path, device=None):
    """An InitNet with the weights written by ``save_weights``."""
    d = np.load(path, allow_pickle=False)
    layers = []
    for l in d["layers"]:
        f = str(l).split(",")
        layers.append((f[0],) + tuple(int(v) for v in f[1:]))
    net = InitNet(frame=int(d["frame"]), device=device, layers=layers)
    for i, p in enumerate(net.params):
        wb = torch.from_numpy(d[f"w{i}"].view(np.int16)).view(torch.bfloat16).to(net.dev)
        p["wb"].copy_(wb)
        p["w32"].copy_(wb.float())
        p["b"].copy_(torch.from_numpy(d[f"b{i}"]).to(net.dev))
    net._refresh_transposed()
    return net


def predict_patches(net, I, photons, batch=256):
    """Complex patches (n, M, K) complex64 from centred intensities I (n, M, K) on device."""
    frame = I.shape[1]
    I = I.to(torch.float32).contiguous()
    scale = input_scale(photons, frame)
    out = []
    for s in range(0, I.shape[0], batch):
        # the first layer forms bf16(sqrt(I) / scale) itself (k_conv_c1)
        yb = net.forward(I[s:s + batch], intensity_scale=scale)
        out.append(torch.complex(yb[..., 0], yb[..., 1]))
    return torch.cat(out).contiguous()


def overlap_pairs(off, M, K, frac=0.75):
    """Pairs (i < j) whose windows are closer than ``frac`` of the frame in both
    axes, as an (npairs, 2) int32 array in (i, j) lexicographic order.
    Sorted by row, the candidates of each frame form one contiguous range."""
    off = np.asarray(off, np.int64)
    N = len(off)
    ty, tx = int(frac * M), int(frac * K)
    order = np.argsort(off[:, 0], kind="stable")
    ys, xs = off[order, 0], off[order, 1]
    lo = np.arange(1, N + 1)                             # later in row order only
    hi = np.searchsorted(ys, ys + ty, side="left")      # first y_j >= y_i + ty
    counts = np.maximum(hi - lo, 0)
    a = np.repeat(np.arange(N), counts)                 # candidate (a, b), both in row order
    b = np.arange(len(a)) - np.repeat(np.cumsum(counts) - counts - lo, counts)
    keep = np.abs(xs[b] - xs[a]) < tx
    i, j = order[a[keep]], order[b[keep]]
    lo_, hi_ = np.minimum(i, j), np.maximum(i, j)
    srt = np.argsort(lo_ * N + hi_)                     # unique keys: (i, j) lexicographic
    return np.stack([lo_[srt], hi_[srt]], 1).astype(np.int32)


def pair_csr(pairs, N):
    """CSR rows of the Hermitian pair matrix: row i lists (j, qs), qs = q for
    pair q = (i, j) and -q-1 (conjugate) for pair q = (j, i); columns ascending."""
    q = np.arange(len(pairs), dtype=np.int64)
    rows = np.concatenate([pairs[:, 0], pairs[:, 1]]).astype(np.int64)
    cols = np.concatenate([pairs[:, 1], pairs[:, 0]]).astype(np.int64)
    qs = np.concatenate([q, -q - 1])
    order = np.argsort(rows * N + cols)                 # unique keys: row-major, columns ascending
    rowptr = np.zeros(N + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=N), out=rowptr[1:])
    rowptr = rowptr.astype(np.int32)
    ent = np.stack([cols[order], qs[order]], 1).astype(np.int32)
    return rowptr, ent


def sync_phases(patches, off, weight=None, iters=300, frac=0.75):
    """Global phases z (N complex64, |z| = 1) that align the patches: projected
    power iterations on the pairwise overlap products (csrc/nn.hip)."""
    dev = patches.device
    N, M, K = patches.shape
    s = _vp(torch.cuda.current_stream(dev).cuda_stream)
    pairs = overlap_pairs(off, M, K, frac)
    rowptr, ent = pair_csr(pairs, N)
    c = torch.zeros((max(len(pairs), 1), 2), dtype=torch.float64, device=dev)
    off_t = torch.from_numpy(np.ascontiguousarray(off, np.int32)).to(dev)
    pairs_t = torch.from_numpy(np.ascontiguousarray(pairs)).to(dev)
    w_t = torch.from_numpy(np.ascontiguousarray(weight, np.float32)).to(dev) if weight is not None else None
    _abi.call("pty_nn_pair_overlap", _p(patches), _p(w_t), _p(off_t), _p(pairs_t), len(pairs), M, K,
              _p(c), s)
    rowptr_t = torch.from_numpy(rowptr).to(dev)
    ent_t = torch.from_numpy(ent).to(dev) if len(ent) else torch.zeros((1, 2), dtype=torch.int32, device=dev)
    z = torch.empty(N, dtype=torch.complex64, device=dev)
    work = torch.empty(max(3 * len(ent), 1), dtype=torch.float32, device=dev)
    _abi.call("pty_nn_phase_sync", _p(rowptr_t), _p(ent_t), _p(c), N, int(iters), _p(z), _p(work),
              len(ent), s)
    return z


def learned_start(net, eng, I_t, P_t, O_plain_t, photons=None, alpha=1e-10, sync_iters=150):
    """Learned O0 on an engine whose data are set, gated by the first-iteration eMod.

    Patches are predicted from the frames, phase-aligned by ``sync_phases`` and
    stitched by the overlap projection psup_O of the exits P * patch * z (the
    |P|^2-weighted average).  The starting eMod of a guess is the reference's
    first-iteration eMod (ptychography/era.py:561-563, 210) of exits O0 * P;
    the learned O0 is kept only if it beats ``O_plain_t``'s.  Returns (O0
    tensor, info).  Leaves the engine's state to be set by the caller.
    """
    from .era import ALGO_ERA
    dev = eng.device
    N, M, K = eng.N, eng.M, eng.K
    if photons is None:
        photons = float(I_t.sum(dtype=torch.float64).item()) / N
    torch.cuda.synchronize(dev)
    t0 = time.time()
    patches = predict_patches(net, I_t, photons)
    if (M, K) != (net.frame, net.frame):
        raise ValueError(f"the network takes {net.frame}x{net.frame} frames, the scan has {M}x{K}")
    off = -np.asarray(eng.R, np.int64)
    w = illumination_weight(P_t.cpu().numpy())
    z = sync_phases(patches, off, w, iters=sync_iters)
    exits = torch.empty((N, M, K), dtype=torch.complex64, device=dev)
    _abi.call("pty_nn_phase_exits", _p(patches), _p(z), _p(P_t), N, M, K, _p(exits),
              _vp(torch.cuda.current_stream(dev).cuda_stream))
    O_l, _ = eng.psup_O(exits, P_t, alpha=alpha)
    del exits, patches

    def first_emod(O_start):
        eng.set_state(O_start.clone(), P_t.clone())
        return float(eng.run(ALGO_ERA, 1, 1, begin=True, alpha=alpha)[0, 0])

    e_learned = first_emod(O_l)
    e_plain = first_emod(O_plain_t)
    use = bool(e_learned < e_plain)
    torch.cuda.synchronize(dev)
    info = dict(eMod_learned_start=e_learned, eMod_plain_start=e_plain, used_learned=use,
                init_seconds=time.time() - t0)
    return (O_l if use else O_plain_t), info


def initial_object(net, I, R, P, U, V, photons=None, mask=None, alpha=1e-10, device=None,
                   O_plain=None, sync_iters=150):
    """Stand-alone learned O0 (complex64 numpy) and the gate's info; the plain
    start is ``O_plain`` (default: a flat object of ones)."""
    dev = require_gpu(device)
    I_t = I if isinstance(I, torch.Tensor) else torch.from_numpy(np.asarray(I, np.float32))
    I_t = I_t.to(dev, torch.float32).contiguous()
    Rs = np.asarray(R).copy()
    Rs -= Rs.max(axis=0)
    N, M, K = I_t.shape
    eng = Engine(N, M, K, U, V, Rs, device=dev)
    try:
        mc = ms = None
        if mask is not None and not (np.isscalar(mask) and mask == 1):
            mc = torch.from_numpy(np.asarray(mask, np.uint8)).to(dev)
            ms = torch.from_numpy(np.fft.ifftshift(np.asarray(mask, np.uint8))).to(dev)
        amp, I_norm = eng.prepare_amp(I_t, mc)
        eng.set_data(amp, ms, I_norm)
        P_t = torch.from_numpy(np.asarray(P, np.complex64)).to(dev)
        plain = torch.ones((U, V), dtype=torch.complex64, device=dev) if O_plain is None else \
            torch.from_numpy(np.asarray(O_plain, np.complex64)).to(dev)
        O0, info = learned_start(net, eng, I_t, P_t, plain, photons, alpha, sync_iters)
        return O0.cpu().numpy(), info
    finally:
        eng.close()
