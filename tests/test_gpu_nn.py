"""BF16 MFMA kernels of the learned initialiser against PyTorch CPU float64.

Inputs and weights are bf16-representable, so the float64 references see the
exact operands; outputs compared in fp32 hold to 1e-5 relative (FP32
accumulation), bf16 outputs to 8e-3 (one bf16 rounding).
"""

import ctypes

import numpy as np
import pytest
import torch

from ptychography_amd import _abi
from ptychography_amd.engine import Engine
from ptychography_amd.learned import (InitNet, initial_object, overlap_pairs, sync_phases,
                                      training_frames, train)

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def P(t):
    return vp(t.data_ptr()) if t is not None else None


def bf(x, dev):
    return torch.from_numpy(np.asarray(x, np.float32)).to(torch.bfloat16).to(dev)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (96, 80, 72), (33, 17, 200)])
def test_mfma_gemm_probe(M, N, K, gpu_device):
    rng = np.random.default_rng(M + N + K)
    A = bf(rng.standard_normal((M, K)), gpu_device)
    B = bf(rng.standard_normal((K, N)), gpu_device)
    C = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
    _abi.call("pty_nn_gemm_probe", P(A), P(B), P(C), M, N, K, None)
    torch.cuda.synchronize()
    ref = A.double().cpu().numpy() @ B.double().cpu().numpy()
    assert rel(C.cpu().numpy(), ref) < 1e-6


def _conv_ref(x, w, b, relu):
    """x (B,H,W,Cin), w (9*Cin, Cout) -> (B,H,W,Cout), float64 CPU."""
    Bn, H, W, Cin = x.shape
    Cout = w.shape[1]
    wt = torch.from_numpy(w).reshape(3, 3, Cin, Cout).permute(3, 2, 0, 1).contiguous()
    xt = torch.from_numpy(x).permute(0, 3, 1, 2)
    y = torch.nn.functional.conv2d(xt, wt, torch.from_numpy(b) if b is not None else None, padding=1)
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 1).numpy()


def _wT(w, Cin, Cout, dev):
    wT = torch.empty_like(w)
    _abi.call("pty_nn_weight_t", P(w), P(wT), Cin, Cout, None)
    return wT


# shapes chosen to reach every kernel: first layer (1 input channel), few output
# channels, direct convolution with 32- and 64-channel column tiles at W = 8
# (tiles spanning two images, odd batch -> partial tile), 64 and 128, and the
# generic implicit GEMM (channel counts that are not multiples of 32); the
# weight gradients of the thin layers use the VALU kernel
CONV_SHAPES = [(2, 16, 16, 1, 32), (1, 32, 32, 64, 2), (2, 64, 64, 32, 2), (2, 8, 8, 32, 64), (3, 8, 8, 32, 32),
               (1, 64, 64, 32, 64), (1, 4, 128, 64, 32), (1, 16, 8, 64, 16), (2, 12, 12, 48, 40)]


@pytest.mark.parametrize("B,H,W,Cin,Cout", CONV_SHAPES)
def test_conv_forward_dgrad_wgrad(B, H, W, Cin, Cout, gpu_device):
    rng = np.random.default_rng(Cin * 7 + Cout + W)
    x = bf(rng.standard_normal((B, H, W, Cin)), gpu_device)
    w = bf(rng.standard_normal((9 * Cin, Cout)) * 0.2, gpu_device)
    wT = _wT(w, Cin, Cout, gpu_device)
    b = torch.from_numpy(rng.standard_normal(Cout).astype(np.float32)).to(gpu_device)
    xd, wd, bd = (t.double().cpu().numpy() for t in (x, w, b))
    ref = _conv_ref(xd, wd, bd, False)
    for wt_arg in (wT, None):   # direct kernel (where the shape allows) and generic GEMM
        y = torch.empty((B, H, W, Cout), dtype=torch.bfloat16, device=gpu_device)
        yf = torch.empty((B, H, W, Cout), dtype=torch.float32, device=gpu_device)
        _abi.call("pty_nn_conv3x3_fwd", P(x), P(w), P(wt_arg), P(b), P(y), P(yf), B, H, W, Cin, Cout,
                  0, None)
        torch.cuda.synchronize()
        assert rel(yf.cpu().numpy(), ref) < 1e-5
        assert rel(y.float().cpu().numpy(), ref) < 8e-3
    yr = torch.empty((B, H, W, Cout), dtype=torch.bfloat16, device=gpu_device)
    _abi.call("pty_nn_conv3x3_fwd", P(x), P(w), P(wT), P(b), P(yr), None, B, H, W, Cin, Cout, 1, None)
    torch.cuda.synchronize()
    assert rel(yr.float().cpu().numpy(), np.maximum(ref, 0)) < 8e-3
    # gradients of L = sum(g * conv(x, w))
    g = bf(rng.standard_normal((B, H, W, Cout)), gpu_device)
    xt = torch.from_numpy(xd).requires_grad_(True)
    wt = torch.from_numpy(wd).requires_grad_(True)
    bt = torch.from_numpy(bd).requires_grad_(True)
    wc = wt.reshape(3, 3, Cin, Cout).permute(3, 2, 0, 1)
    yt = torch.nn.functional.conv2d(xt.permute(0, 3, 1, 2), wc, bt, padding=1).permute(0, 2, 3, 1)
    (yt * torch.from_numpy(g.double().cpu().numpy())).sum().backward()
    if Cin > 1:
        dx = torch.empty_like(x)
        _abi.call("pty_nn_conv3x3_dgrad", P(g), P(w), P(dx), None, B, H, W, Cin, Cout, None)
        # with a ReLU mask (the producer's post-activation): dx = 0 where mask <= 0
        mask = bf(rng.standard_normal((B, H, W, Cin)), gpu_device)
        dxm = torch.empty_like(x)
        _abi.call("pty_nn_conv3x3_dgrad", P(g), P(w), P(dxm), P(mask), B, H, W, Cin, Cout, None)
        torch.cuda.synchronize()
        assert rel(dx.float().cpu().numpy(), xt.grad.numpy()) < 8e-3
        ref_m = xt.grad.numpy() * (mask.float().cpu().numpy() > 0)
        assert rel(dxm.float().cpu().numpy(), ref_m) < 8e-3
    # the library's split count and a single split must both match
    for splits in sorted({1, _abi.load().pty_nn_wgrad_splits(B, H, W, Cin, Cout)}):
        dw = torch.empty((9 * Cin, Cout), dtype=torch.float32, device=gpu_device)
        db = torch.empty(Cout, dtype=torch.float32, device=gpu_device)
        ws = torch.empty(splits * (9 * Cin + 1) * Cout, dtype=torch.float32, device=gpu_device)
        _abi.call("pty_nn_conv3x3_wgrad", P(x), P(g), P(dw), P(db), P(ws), splits, B, H, W, Cin, Cout,
                  None)
        torch.cuda.synchronize()
        assert rel(dw.cpu().numpy(), wt.grad.numpy()) < 1e-5, splits
        assert rel(db.cpu().numpy(), bt.grad.numpy()) < 1e-5, splits


def test_pool_up_and_relu_bwd(gpu_device):
    rng = np.random.default_rng(3)
    C = 16
    x = bf(rng.standard_normal((2, 8, 6, C)), gpu_device)
    y = torch.empty((2, 4, 3, C), dtype=torch.bfloat16, device=gpu_device)
    _abi.call("pty_nn_pool2", P(x), P(y), 2, 8, 6, C, 0, None)
    xd = x.double().cpu().numpy()
    ref = xd.reshape(2, 4, 2, 3, 2, C).mean(axis=(2, 4))
    torch.cuda.synchronize()
    assert rel(y.float().cpu().numpy(), ref) < 8e-3
    up = torch.empty((2, 16, 12, C), dtype=torch.bfloat16, device=gpu_device)
    _abi.call("pty_nn_up2", P(x), P(up), 2, 8, 6, C, 0, None)
    torch.cuda.synchronize()
    assert np.array_equal(up.float().cpu().numpy(), np.repeat(np.repeat(xd, 2, 1), 2, 2))
    # backward of up = 2x2 sum; of pool = 1/4 spread
    dxu = torch.empty_like(x)
    _abi.call("pty_nn_up2", P(up), P(dxu), 2, 8, 6, C, 1, None)
    torch.cuda.synchronize()
    assert rel(dxu.float().cpu().numpy(), 4 * xd) < 8e-3
    dxp = torch.empty_like(x)   # pool backward: each input pixel gets 1/4 of its output's gradient
    _abi.call("pty_nn_pool2", P(y), P(dxp), 2, 8, 6, C, 1, None)
    torch.cuda.synchronize()
    spread = np.repeat(np.repeat(y.double().cpu().numpy(), 2, 1), 2, 2) / 4
    assert rel(dxp.float().cpu().numpy(), spread) < 8e-3
    dy = bf(rng.standard_normal((3, 7, 5, 3)), gpu_device)   # 315 elements: vector body + tail
    yy = bf(rng.standard_normal(dy.shape), gpu_device)
    expect = dy.float().cpu().numpy() * (yy.float().cpu().numpy() > 0)
    _abi.call("pty_nn_relu_bwd", P(dy), P(yy), dy.numel(), None)
    torch.cuda.synchronize()
    assert np.array_equal(dy.float().cpu().numpy(), expect)


SMALL = [("conv", 1, 8), ("conv", 8, 8), ("pool",), ("conv", 8, 16), ("up",), ("conv", 16, 8),
         ("conv", 8, 2)]


def _net_ref(net, x, t, w):
    """float64 CPU emulation of the GPU network (bf16 rounding of every stored
    activation and of the output gradient), with autograd for the reference grads."""
    def r(a):  # round through bf16 like the GPU stores
        return a.to(torch.bfloat16).to(torch.float64)

    class RoundBf16(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a):
            return r(a)

        @staticmethod
        def backward(ctx, g):
            return r(g)

    params = []
    h = torch.from_numpy(x.float().cpu().numpy()).double().permute(0, 3, 1, 2)
    pi = 0
    for l in net.layers:
        if l[0] == "conv":
            p = net.params[pi]
            pi += 1
            wt = torch.from_numpy(p["wb"].float().cpu().numpy()).double().requires_grad_(True)
            bt = torch.from_numpy(p["b"].cpu().numpy()).double().requires_grad_(True)
            params.append((wt, bt))
            wc = wt.reshape(3, 3, p["cin"], p["cout"]).permute(3, 2, 0, 1)
            h = torch.nn.functional.conv2d(h, wc, bt, padding=1)
            if p["relu"]:
                h = RoundBf16.apply(torch.relu(h))
        elif l[0] == "pool":
            h = RoundBf16.apply(torch.nn.functional.avg_pool2d(h, 2))
        else:
            h = RoundBf16.apply(torch.nn.functional.interpolate(h, scale_factor=2, mode="nearest"))
    out = h.permute(0, 2, 3, 1)
    tt = torch.from_numpy(t.cpu().numpy()).double()
    wt_ = torch.from_numpy(np.repeat(w.ravel(), 2).reshape(out.shape[1], out.shape[2], 2)).double()
    loss = (wt_ * (out - tt) ** 2).mean()
    loss.backward()
    return out.detach().numpy(), float(loss), [(a.grad.numpy(), b.grad.numpy()) for a, b in params]


def test_network_gradients_match_float64_reference(gpu_device):
    rng = np.random.default_rng(11)
    net = InitNet(frame=16, seed=4, device=gpu_device, layers=SMALL)
    x = bf(np.abs(rng.standard_normal((4, 16, 16, 1))), gpu_device)
    t = torch.from_numpy(rng.standard_normal((4, 16, 16, 2)).astype(np.float32)).to(gpu_device)
    w = (rng.random((16, 16)) + 0.5).astype(np.float32)
    wt = torch.from_numpy(np.repeat(w.ravel(), 2)).to(gpu_device)
    pred = net.forward(x, keep=True)
    loss, dy = net.loss(pred, t, wt, grad=True, gscale=1.0)
    net.backward(dy)
    torch.cuda.synchronize()
    out_ref, loss_ref, grads_ref = _net_ref(net, x, t, w)
    assert rel(pred.cpu().numpy(), out_ref) < 1e-2
    assert abs(loss - loss_ref) < 2e-2 * loss_ref
    for p, (gw, gb) in zip(net.params, grads_ref):
        assert rel(p["gw"].cpu().numpy(), gw) < 5e-2
        assert rel(p["gb"].cpu().numpy(), gb) < 5e-2


def test_training_reduces_loss_and_gate(gpu_device):
    torch.manual_seed(0)
    x, t, Pr, w = training_frames(2048, seed=5, device=gpu_device, obj_size=192, per_object=512)
    net = InitNet(seed=1, device=gpu_device)
    hist = train(net, x, t, w, steps=150, batch=32, lr=2e-3)
    first, last = np.mean(hist[:10]), np.mean(hist[-10:])
    print(f"training loss {first:.4e} -> {last:.4e}")
    assert np.all(np.isfinite(hist)) and last < 0.7 * first
    # the gate keeps whichever start has the lower first-iteration eMod
    from ptychography_amd import sim
    s = sim.make_scan("c5", grid=(6, 6), noise=False)
    U, V = s.O.shape
    O0, info = initial_object(net, s.I, s.R, s.P, U, V, device=gpu_device)
    assert info["used_learned"] == (info["eMod_learned_start"] < info["eMod_plain_start"])
    assert O0.shape == (U, V)


def test_pair_overlap_sync_and_stitch(gpu_device):
    """Patches = windows of a known object times random phases: the pair
    products match numpy, the sync recovers the phases up to one global phase
    and the stitched object (overlap projection of P * patch * z) is the object."""
    from ptychography_amd import sim
    s = sim.make_scan("c5", grid=(8, 8), noise=False)
    off = -s.R
    N, M = len(off), s.P.shape[0]
    U, V = s.O.shape
    rng = np.random.default_rng(2)
    phi = rng.uniform(-np.pi, np.pi, N)
    true_p = np.stack([s.O[y:y + M, x:x + M] for y, x in off])
    patches = torch.from_numpy((true_p * np.exp(-1j * phi)[:, None, None]).astype(np.complex64)).to(gpu_device)
    w = (np.abs(s.P) ** 2 / (np.abs(s.P) ** 2).max()).astype(np.float32)
    pairs = overlap_pairs(off, M, M)
    c = torch.zeros((len(pairs), 2), dtype=torch.float64, device=gpu_device)
    # keep every argument tensor alive until the kernel has run
    w_t = torch.from_numpy(w).to(gpu_device)
    off_t = torch.from_numpy(off.astype(np.int32)).to(gpu_device)
    pairs_t = torch.from_numpy(pairs).to(gpu_device)
    _abi.call("pty_nn_pair_overlap", P(patches), P(w_t), P(off_t), P(pairs_t), len(pairs), M, M,
              P(c), None)
    torch.cuda.synchronize()
    pn = patches.cpu().numpy().astype(np.complex128)
    for q in (0, len(pairs) // 2, len(pairs) - 1):
        i, j = pairs[q]
        dy, dx = off[j] - off[i]
        a = (w * pn[i])[max(0, dy):M + min(0, dy), max(0, dx):M + min(0, dx)]
        b = (w * pn[j])[max(0, -dy):M + min(0, -dy), max(0, -dx):M + min(0, -dx)]
        ref = np.vdot(a, b)
        got = c[q, 0].item() + 1j * c[q, 1].item()
        assert abs(got - ref) <= 1e-6 * abs(ref) + 1e-9
    z = sync_phases(patches, off, w, iters=300).cpu().numpy()
    q = z * np.exp(-1j * phi)            # should be one constant phase
    q *= np.conj(q.mean() / abs(q.mean()))
    assert np.abs(np.angle(q)).max() < 1e-3
    eng = Engine(N, M, M, U, V, s.R - s.R.max(0), device=gpu_device)
    P_t = torch.from_numpy(s.P.astype(np.complex64)).to(gpu_device)
    exits = torch.empty_like(patches)
    zt = torch.from_numpy(z.astype(np.complex64)).to(gpu_device)
    _abi.call("pty_nn_phase_exits", P(patches), P(zt), P(P_t), N, M, M, P(exits), None)
    O, H = eng.psup_O(exits, P_t)
    eng.close()
    O = O.cpu().numpy()
    cov = H.cpu().numpy() > 1e-3 * H.max().item()
    g = (O[cov] * np.conj(s.O[cov])).sum() / (np.abs(s.O[cov]) ** 2).sum()
    assert abs(abs(g) - 1) < 1e-4
    assert rel(O[cov] / g, s.O[cov]) < 1e-4


def test_batch_gather_and_augmentation(gpu_device):
    from test_learned_host import _dihedral
    rng = np.random.default_rng(5)
    M, n, B = 16, 10, 6
    x = bf(rng.standard_normal((n, M, M)), gpu_device)
    t = torch.from_numpy(rng.standard_normal((n, M, M, 2)).astype(np.float32)).to(gpu_device)
    idx = torch.tensor([3, 0, 9, 3, 7, 1], dtype=torch.int64, device=gpu_device)
    xn, tn = x.float().cpu().numpy(), t.cpu().numpy()
    for aug in (0, 1):
        xb = torch.empty((B, M, M), dtype=torch.bfloat16, device=gpu_device)
        tb = torch.empty((B, M, M, 2), dtype=torch.float32, device=gpu_device)
        _abi.call("pty_nn_batch_gather", P(x), P(t), P(idx), B, M, aug, 1234, P(xb), P(tb), None)
        torch.cuda.synchronize()
        xo, to = xb.float().cpu().numpy(), tb.cpu().numpy()
        codes = []
        for b, k in enumerate(idx.cpu().numpy()):
            match = [c for c in range(8) if np.array_equal(xo[b], _dihedral(xn[k], c))
                     and np.array_equal(to[b], _dihedral(tn[k], c))]
            assert match, (aug, b)
            codes.append(match[0])
        if aug == 0:
            assert codes == [0] * B
        else:
            assert len(set(codes)) > 1


def test_first_layer_forms_its_input_from_intensities(gpu_device):
    # forward(I, intensity_scale) builds bf16(sqrt(max(I, 0)) / scale) inside
    # the first conv; the result must equal forward() on that input formed by torch
    from ptychography_amd.learned import InitNet
    net = InitNet(seed=3, device=gpu_device)
    g = torch.Generator(device=gpu_device).manual_seed(9)
    I = torch.rand((8, 64, 64), device=gpu_device, generator=g) * 400.0 - 5.0   # some negative
    scale = 3.7
    x = (torch.sqrt(I.clamp_min(0)) / scale).to(torch.bfloat16)[..., None].contiguous()
    a = net.forward(x)
    b = net.forward(I.contiguous(), intensity_scale=scale)
    assert torch.equal(a, b)
