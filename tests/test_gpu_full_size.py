"""ERA at the full headline size (4096 frames of 256^2, object 2782^2), checked
through properties that do not need the oracle (too slow at this size):

* fixed point: noise-free, background-free data formed from (O, P) leave the
  true object where it is (eMod ~ float32 round-off, object unchanged on the
  illuminated region);
* error reduction: with the probe known (method 1) an ERA iteration is a
  projection onto the modulus set followed by the least-squares object fit,
  i.e. the orthogonal projection onto the subspace of exits P * O_window, so
  the distance eMod between the two sets cannot grow from one iteration to
  the next (Fienup's error-reduction property), whatever the noise;
* determinism: two runs from the same start are bit-identical;
* the fused inverse-row / object-update kernel against the unfused kernels;
* DM: from the true object and probe on noise-free data the DM iterate f is
  consistent and satisfies the modulus constraint, so f, O and P stay put;
  the same with background refinement (method 4, C3's own algorithm) from the
  true background on data that include it;
  DM with background refinement (method 4) on the noisy data is
  bit-identical between two runs and its metrics stay finite.

Bars: eMod at the fixed point <= 2e-5 (ERA and DM) and object change <= 1e-5 (rel L2 on
pixels whose heatmap is >= 1e-3 of its maximum); eMod non-increasing within
a relative 1e-6 per step; fused vs unfused object rel L2 <= 1e-5 and eMod
within 1e-5 relative after 8 iterations.
"""

import numpy as np
import pytest
import torch

from ptychography_amd import sim
from ptychography_amd.engine import Engine
from tests.scans import rel_l2

pytestmark = pytest.mark.gpu


def _engine(scan, dev):
    I = scan["I"]
    N, M, K = I.shape
    U, V = scan["O"].shape
    eng = Engine(N, M, K, U, V, scan["R"].astype(np.int32), device=dev)
    mask = scan["mask"]
    mask_c = mask_s = None
    if not np.isscalar(mask):
        mask_c = torch.from_numpy(np.ascontiguousarray(mask).astype(np.uint8)).to(dev)
        mask_s = torch.from_numpy(np.ascontiguousarray(np.fft.ifftshift(mask)).astype(np.uint8)).to(dev)
    amp, I_norm = eng.prepare_amp(I, mask_c)
    eng.set_data(amp, mask_s, I_norm)
    return eng, (amp, mask_s)


def _state(eng, dev, O0, P0):
    O_t = torch.from_numpy(np.ascontiguousarray(O0).astype(np.complex64)).to(dev)
    P_t = torch.from_numpy(np.ascontiguousarray(P0).astype(np.complex64)).to(dev)
    eng.set_state(O_t, P_t)
    return O_t, P_t


@pytest.fixture(scope="module")
def c3_scan(gpu_device):
    print("full size: C3 scan (4096 x 256^2, Poisson, 1 % background, beamstop)", flush=True)
    scan = sim.make_scan_gpu(sim.RECIPES["c3"], gpu_device)
    assert tuple(scan["I"].shape) == (4096, 256, 256) and scan["O"].shape == (2782, 2782)
    yield scan
    del scan
    torch.cuda.empty_cache()


def test_c3_fixed_point(gpu_device):
    print("full size: fixed point", flush=True)
    spec = sim.RECIPES["c3"].with_(background_frac=0.0)
    scan = sim.make_scan_gpu(spec, gpu_device, poisson=False)
    eng, keep = _engine(scan, gpu_device)
    O_t, P_t = _state(eng, gpu_device, scan["O"], scan["P"])
    m = eng.run(0, 1, 3, begin=True)
    print("eMod at the true object:", m[:, 0])
    assert np.all(m[:, 0] <= 2e-5)
    H = eng.heatmap().cpu().numpy()
    lit = H >= 1e-3 * H.max()
    err = rel_l2(O_t.cpu().numpy()[lit], scan["O"][lit])
    print(f"object change on the illuminated region: {err:.2e}")
    assert err <= 1e-5
    eng.close()
    del keep, O_t, P_t, scan
    torch.cuda.empty_cache()


def test_c3_error_reduction_and_determinism(c3_scan, gpu_device):
    print("full size: eMod non-increasing, bit-identical reruns", flush=True)
    eng, keep = _engine(c3_scan, gpu_device)
    O0 = np.ones(c3_scan["O"].shape, np.complex64)
    runs = []
    for rep in range(2):
        O_t, P_t = _state(eng, gpu_device, O0, c3_scan["P"])
        m = eng.run(0, 1, 30, begin=True)
        runs.append((m, O_t.cpu().numpy()))
    e = runs[0][0][:, 0]
    print("eMod:", " ".join(f"{x:.5e}" for x in e[[0, 1, 2, 5, 10, 20, 29]]))
    steps = e[1:] / e[:-1] - 1.0
    assert np.all(steps <= 1e-6), f"eMod grew by {steps.max():.2e} at step {int(steps.argmax()) + 1}"
    assert e[-1] < 0.5 * e[0]
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    eng.close()
    del keep


def test_c3_fused_object_update_matches_unfused(c3_scan, gpu_device):
    print("full size: fused vs unfused object update", flush=True)
    eng, keep = _engine(c3_scan, gpu_device)
    sup, _ = eng.row_gather_status()
    assert sup
    O0 = np.ones(c3_scan["O"].shape, np.complex64)
    out = {}
    for fused in (True, False):
        eng.set_row_gather(fused)
        O_t, P_t = _state(eng, gpu_device, O0, c3_scan["P"])
        m = eng.run(0, 1, 8, begin=True)
        out[fused] = (m, O_t.cpu().numpy())
    err = rel_l2(out[True][1], out[False][1])
    print(f"object rel L2 fused vs unfused after 8 iterations: {err:.2e}")
    assert err <= 1e-5
    np.testing.assert_allclose(out[True][0][:, 0], out[False][0][:, 0], rtol=1e-5)
    eng.close()
    del keep


def test_c3_frame_fft_against_torch(gpu_device):
    # the engine's batched 2-D FFT over all 4096 frames against torch.fft in
    # complex128 (forward unnormalised, inverse with 1/(M K) like numpy)
    print("full size: 2-D FFT of 4096 x 256^2 against torch.fft (complex128)", flush=True)
    N, M = 4096, 256
    R = -sim.make_positions((64, 64), 40, 3, np.random.default_rng(2))
    R -= R.max(axis=0)
    U, V = int(-R[:, 0].min()) + M, int(-R[:, 1].min()) + M
    eng = Engine(N, M, M, U, V, R.astype(np.int32), device=gpu_device)
    g = torch.Generator(device=gpu_device).manual_seed(7)
    x = torch.randn((N, M, M), dtype=torch.complex64, device=gpu_device, generator=g)
    ref = torch.fft.fft2(x.to(torch.complex128))
    y = eng.fft2(x.clone())
    e_f = (torch.linalg.vector_norm(y.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)).item()
    z = eng.fft2(y, inverse=True)
    e_i = (torch.linalg.vector_norm(z.to(torch.complex128) - x.to(torch.complex128)) /
           torch.linalg.vector_norm(x.to(torch.complex128))).item()
    print(f"forward rel L2 {e_f:.2e}, round trip rel L2 {e_i:.2e}")
    assert e_f <= 1e-6 and e_i <= 1e-6
    eng.close()
    del x, y, z, ref
    torch.cuda.empty_cache()


def test_c3_dm_fixed_point(gpu_device):
    print("full size: DM fixed point", flush=True)
    spec = sim.RECIPES["c3"].with_(background_frac=0.0)
    scan = sim.make_scan_gpu(spec, gpu_device, poisson=False)
    eng, keep = _engine(scan, gpu_device)
    O_t, P_t = _state(eng, gpu_device, scan["O"], scan["P"])
    m = eng.run(1, 3, 3, begin=True)
    print("DM eMod at the true object and probe:", m[:, 0])
    assert np.all(m[:, 0] <= 2e-5)
    H = eng.heatmap()
    lit = (H >= 1e-3 * H.max()).cpu().numpy() if H is not None else np.ones(scan["O"].shape, bool)
    eo = rel_l2(O_t.cpu().numpy()[lit], scan["O"][lit])
    ep = rel_l2(P_t.cpu().numpy(), scan["P"])
    print(f"object change {eo:.2e}, probe change {ep:.2e}")
    assert eo <= 1e-5 and ep <= 1e-5
    eng.close()
    del keep, O_t, P_t, scan
    torch.cuda.empty_cache()


def test_c3_dm_background_fixed_point(gpu_device):
    # C3's own algorithm, DM with background refinement (method 4), on exact
    # data I = |FFT(O_n P)|^2 + B from the true object, probe and background B
    # (the recipe's smooth 1 % background), started from them: the modulus step
    # with background (Pmod_7) then leaves f and b unchanged, the object
    # projection returns O and the frame mean of b is b, so every iteration
    # is the identity.  No beamstop here: Pmod_7 zeroes the masked pixels of
    # f and b (era.py:589-592), which is no fixed point of exact data.
    print("full size: DM method 4 fixed point with the true background", flush=True)
    spec = sim.RECIPES["c3"].with_(beamstop_radius=0.0)
    scan = sim.make_scan_gpu(spec, gpu_device, poisson=False)
    assert scan["B"] is not None and np.isscalar(scan["mask"])
    eng, keep = _engine(scan, gpu_device)
    O_t, P_t = _state(eng, gpu_device, scan["O"], scan["P"])
    b_true = np.sqrt(np.fft.ifftshift(scan["B"])).astype(np.float32)
    eng.set_background(torch.from_numpy(b_true), per_frame=False)
    m = eng.run(1, 4, 3, begin=True)
    print("DM-bg eMod at the true object, probe and background:", m[:, 0])
    assert np.all(m[:, 0] <= 2e-5)
    H = eng.heatmap().cpu().numpy()
    lit = H >= 1e-3 * H.max()
    eo = rel_l2(O_t.cpu().numpy()[lit], scan["O"][lit])
    eb = rel_l2(eng.background().cpu().numpy(), b_true)
    print(f"object change {eo:.2e}, background change {eb:.2e}")
    assert eo <= 1e-5 and eb <= 1e-5
    eng.close()
    del keep, O_t, P_t, scan
    torch.cuda.empty_cache()


def test_c3_dm_background_deterministic(c3_scan, gpu_device):
    print("full size: DM method 4 bit-identical reruns", flush=True)
    eng, keep = _engine(c3_scan, gpu_device)
    b0 = torch.from_numpy(np.sqrt(np.fft.ifftshift(c3_scan["B"])).astype(np.float32))
    O0 = np.ones(c3_scan["O"].shape, np.complex64)
    runs = []
    for rep in range(2):
        O_t, P_t = _state(eng, gpu_device, O0, c3_scan["P"])
        eng.set_background(b0, per_frame=False)
        m = eng.run(1, 4, 6, begin=True)
        runs.append((m, O_t.cpu().numpy(), eng.background().cpu().numpy()))
    print("DM-bg eMod:", " ".join(f"{x:.4e}" for x in runs[0][0][:, 0]))
    assert np.all(np.isfinite(runs[0][0]))
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x, y)
    eng.close()
    del keep
