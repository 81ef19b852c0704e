"""ERA object-only steps with the inverse row FFT fused into the object update
(k_rows_gather) against the unfused kernels (k_rows_inv + k_gather_O); and DM
with background, whose object-only projection of the leading Pmod_7 output
runs in the same kernel, which then also writes f (instead of k_rows_inv
storing f and k_gather_O reading it back).

The fused kernel sums each object pixel's frames in the same ascending order
as the unfused gather, but it parks conj(P) psi' in LDS before adding it, while
the unfused gather lets the compiler contract the product into the running sum
(two FMAs): the two round differently.  Bar: object iterates within rel L2
1e-5 (the ERA parity tolerance against the oracle), eMod within 1e-5 relative,
eCon / eSup (FP64 partials per object row instead of per tile) within 1e-4.
Oracle parity of the fused path is covered by the 64^2 ERA tests in
test_gpu_solvers.py, which take it by default.
"""

import os

import numpy as np
import pytest
import torch

from ptychography_amd.engine import Engine
from tests.scans import broadband_start, rel_l2, small_scan

pytestmark = pytest.mark.gpu


def _run(s, dev, method, rowg, iters=(1, 6, 5), graphs=False, bg=None, obj_pad=0, before_run=None,
         algorithm=0, OP_iters=(1, 1), plan=0):
    N, M, K = s.I.shape
    U, V = s.O.shape
    # both paths on the same line-FFT plans (default: the default plans, also
    # for cutting the row segments at creation), so that only the fusion
    # differs; the plans themselves are compared in test_gpu_fft_plans.py
    saved = os.environ.get("PTY_FFT_PLAN")
    os.environ["PTY_FFT_PLAN"] = str(plan)
    try:
        eng = Engine(N, M, K, U + obj_pad, V + 2 * obj_pad, s.R.astype(np.int32), device=dev)
    finally:
        if saved is None:
            del os.environ["PTY_FFT_PLAN"]
        else:
            os.environ["PTY_FFT_PLAN"] = saved
    other = before_run() if before_run else None
    eng.set_path(False)
    eng.set_row_gather(rowg)
    sup, en = eng.row_gather_status()
    assert sup and en == rowg
    eng.set_graphs(graphs)
    amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), None)
    eng.set_data(amp, None, I_norm)
    O_t = torch.from_numpy(broadband_start((U + obj_pad, V + 2 * obj_pad)).astype(np.complex64)).to(dev)
    P_t = torch.from_numpy(s.P.astype(np.complex64)).to(dev)
    eng.set_state(O_t, P_t)
    if bg is not None:
        eng.set_background(torch.from_numpy(bg.astype(np.float32)), per_frame=False)
    rows, it, Os = [], 0, []
    for ci, n in enumerate(iters):
        rows.append(eng.run(algorithm, method, n, iter0=it, begin=(ci == 0), OP_iters=OP_iters))
        it += n
        Os.append(O_t.cpu().numpy())
    eng.close()
    if other is not None:
        other.close()
    return np.concatenate(rows), Os


def _compare(a, b):
    ma, Oa = a
    mb, Ob = b
    errs = [rel_l2(x, y) for x, y in zip(Oa, Ob)]
    print("object rel L2 fused vs unfused:", " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= 1e-5
    np.testing.assert_allclose(ma[:, 0], mb[:, 0], rtol=1e-5)   # eMod
    np.testing.assert_allclose(ma[:, 1:], mb[:, 1:], rtol=1e-4, atol=1e-300)


@pytest.mark.parametrize("frame,grid,step,jitter", [
    (64, (5, 5), 12, 1), (96, (4, 4), 20, 2), (128, (3, 4), 40, 3), (192, (2, 3), 48, 2),
    (256, (3, 3), 40, 3)])
def test_fused_matches_unfused_method1(frame, grid, step, jitter, gpu_device):
    print(f"row gather method 1 frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=jitter)
    _compare(_run(s, gpu_device, 1, True), _run(s, gpu_device, 1, False))


@pytest.mark.parametrize("frame,grid,step", [(256, (2, 16), 130), (64, (2, 80), 55)])
def test_fused_segmented_rows(frame, grid, step, gpu_device):
    # objects wide enough that each object row is split into two column
    # segments (frames straddling the boundary are transformed by both)
    print(f"row gather segmented rows frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=2)
    assert s.O.shape[1] > {256: 2200, 64: 4400}[frame]
    _compare(_run(s, gpu_device, 1, True), _run(s, gpu_device, 1, False))


def test_fused_segmented_rows_alternative_plan(gpu_device):
    # the row gather's alternative 256 plan parks 8 lines of 265 words, so a
    # work-group takes whole object rows up to ~4300 columns; wider objects
    # are still cut into two segments
    print("row gather segmented rows, alternative plan", flush=True)
    s = small_scan(frame=256, grid=(2, 36), step=130, jitter=2)
    assert s.O.shape[1] > 4400
    _compare(_run(s, gpu_device, 1, True, plan=4), _run(s, gpu_device, 1, False, plan=4))


def test_contexts_with_different_segment_widths(gpu_device):
    # a context for a narrower object of the same frame size, created after
    # this one, must not change what this one's fused kernel may use
    print("row gather: two contexts, different segment widths", flush=True)
    s = small_scan(frame=256, grid=(2, 16), step=130, jitter=2)      # two 1105-wide segments
    narrow = small_scan(frame=256, grid=(2, 2), step=40, jitter=1)   # one 296-wide segment

    def make_narrow():
        N, M, K = narrow.I.shape
        return Engine(N, M, K, *narrow.O.shape, narrow.R.astype(np.int32), device=gpu_device)

    _compare(_run(s, gpu_device, 1, True, before_run=make_narrow), _run(s, gpu_device, 1, False))


def test_fused_matches_unfused_background(gpu_device):
    print("row gather method 4 frame 64", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=1, background_frac=0.02)
    bg = np.random.default_rng(4).random((64, 64)) * 0.5
    _compare(_run(s, gpu_device, 4, True, bg=bg), _run(s, gpu_device, 4, False, bg=bg))


def test_fused_uncovered_rows_and_graphs(gpu_device):
    # object larger than the scanned area: rows / columns no frame covers
    # (the object there becomes 0 / (0 + alpha) = 0 in both paths); graphs on
    print("row gather padded object + graphs", flush=True)
    s = small_scan(frame=64, grid=(4, 4), step=16, jitter=2)
    a = _run(s, gpu_device, 1, True, graphs=True, obj_pad=40)
    b = _run(s, gpu_device, 1, False, graphs=False, obj_pad=40)
    _compare(a, b)
    assert np.all(a[1][-1][-40:, :] == 0) and np.all(a[1][-1][:, -80:] == 0)


@pytest.mark.parametrize("frame,grid,step,method,op", [
    (64, (5, 5), 12, 4, (1, 1)), (96, (4, 4), 20, 6, (1, 2)), (256, (3, 3), 40, 4, (1, 1)),
    (256, (2, 16), 130, 4, (1, 1))])
def test_dm_background_projection_fused(frame, grid, step, method, op, gpu_device):
    # f written by the fused kernel must equal what k_rows_inv stores: every
    # frame row exactly once, also where rows straddle two column segments
    # (the last case); method 6 with OP_iters (1, 2) fuses on its object-only
    # steps.  DM is not a contraction: at 256^2 (3 x 3 scan) the first fused
    # step differs from the unfused one by 1.8e-8 (rel L2: a few one-ulp
    # differences) and the gap then grows about 1.5x per iteration (1.6e-5
    # after 8), so the objects are compared after 1 and 5 iterations; at 64^2
    # the two paths stay bit-identical.
    print(f"row gather DM method {method} frame {frame} grid {grid}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=2, background_frac=0.02)
    bg = np.random.default_rng(4).random((frame, frame)) * 0.5
    kw = dict(bg=bg, algorithm=1, OP_iters=op, iters=(1, 4))
    _compare(_run(s, gpu_device, method, True, **kw), _run(s, gpu_device, method, False, **kw))
