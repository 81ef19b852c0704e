"""DM methods 1 and 4 with the iterate held as spectra.

With the probe fixed and no far-field probe constraint, the engine carries
DM's iterate f as F = FFT2(f): with background refinement (method 4) it folds
the next iteration's leading Pmod_7 into the column pass of the DM step
(``dm_bg_fourier_iter``); without (method 1) it forms each iteration's metric
in the next iteration's column pass (``dm1_fourier_iter``, the last one in
the run's tail).  The real-space path (PTY_DM_FOURIER=0, or any other DM
method) stores f itself.  The two
are the same algorithm and differ by rounding.  DM is not a contraction: a
one-ulp difference grows about 1.5x per iteration (the background state, of
small values, faster), so the one-step results must agree to 1e-6 (rel L2)
and after 3 iterations the object, f (through ``pty_get_iterate``, which
transforms it back) and the background to the row-gather tests' 1e-5; eMod
within 1e-5 relative.  Oracle parity of the spectral
path is covered by the DM method-4 cases of test_gpu_solvers.py and
test_gpu_step_parity.py, which take it by default.
"""

import os

import numpy as np
import pytest
import torch

from ptychography_amd.engine import Engine
from tests.scans import broadband_start, rel_l2, small_scan

pytestmark = pytest.mark.gpu


class _Env:
    def __init__(self, real):
        self.real = real

    def __enter__(self):
        self.old = os.environ.get("PTY_DM_FOURIER")
        if self.real:
            os.environ["PTY_DM_FOURIER"] = "0"
        else:
            os.environ.pop("PTY_DM_FOURIER", None)

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("PTY_DM_FOURIER", None)
        else:
            os.environ["PTY_DM_FOURIER"] = self.old


def _setup(s, dev, bg, graphs=False, P0=None):
    N, M, K = s.I.shape
    U, V = s.O.shape
    eng = Engine(N, M, K, U, V, s.R.astype(np.int32), device=dev)
    amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), None)
    eng.set_data(amp, None, I_norm)
    O_t = torch.from_numpy(broadband_start((U, V)).astype(np.complex64)).to(dev)
    P_t = torch.from_numpy((s.P if P0 is None else P0).astype(np.complex64)).to(dev)
    eng.set_state(O_t, P_t)
    if bg is not None:
        eng.set_background(torch.from_numpy(bg.astype(np.float32)), per_frame=False)
    eng.set_graphs(graphs)
    return eng, O_t, P_t, amp


def _run(s, dev, bg, plan, method=4, graphs=False, P0=None, Ps=None):
    """plan: list of (iterations, real-space path?) runs, the first with begin;
    the probe after each run goes to the list Ps when given."""
    eng, O_t, P_t, amp = _setup(s, dev, bg if method == 4 else None, graphs, P0)
    try:
        rows, Os, it = [], [], 0
        for k, (n, real) in enumerate(plan):
            with _Env(real):
                rows.append(eng.run(1, method, n, iter0=it, begin=(k == 0)))
            it += n
            Os.append(O_t.cpu().numpy())
            if Ps is not None:
                Ps.append(P_t.cpu().numpy())
        f, b = eng.get_iterate(f=True, background=method == 4)
        return (np.concatenate(rows), Os, f.cpu().numpy(),
                None if b is None else b.cpu().numpy())
    finally:
        eng.close()


CASES = [(64, (5, 5), 12), (256, (3, 3), 48)]


@pytest.mark.parametrize("frame,grid,step", CASES)
def test_spectral_and_real_space_paths_agree(frame, grid, step, gpu_device):
    print(f"DM method 4 spectra vs real space, frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=2, background_frac=0.02)
    bg = np.random.default_rng(4).random((frame, frame)) * 0.5
    ma, Oa, fa, ba = _run(s, gpu_device, bg, [(1, False), (2, False)])
    mb, Ob, fb, bb = _run(s, gpu_device, bg, [(1, True), (2, True)])
    errs = [rel_l2(x, y) for x, y in zip(Oa, Ob)]
    print("object rel L2 after 1 / 3:", " ".join(f"{e:.1e}" for e in errs),
          f"f {rel_l2(fa, fb):.1e} b {rel_l2(ba, bb):.1e}")
    assert errs[0] < 1e-6 and errs[1] < 1e-5
    assert rel_l2(fa, fb) < 1e-5 and rel_l2(ba, bb) < 1e-5
    assert np.allclose(ma[:, 0], mb[:, 0], rtol=1e-5)
    assert np.allclose(ma[:, 2], mb[:, 2], rtol=1e-4)


def test_split_runs_equal_one_run(gpu_device):
    """The spectra, the folded leading step and its background sums carry
    from one pty_run to the next: 2 + 4 iterations = 6, bit for bit."""
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=2, background_frac=0.02)
    bg = np.random.default_rng(5).random((64, 64)) * 0.5
    ma, Oa, fa, ba = _run(s, gpu_device, bg, [(6, False)])
    mb, Ob, fb, bb = _run(s, gpu_device, bg, [(2, False), (4, False)])
    assert np.array_equal(Oa[-1], Ob[-1])
    assert np.array_equal(fa, fb) and np.array_equal(ba, bb)
    assert np.array_equal(ma, mb)


def test_switching_paths_mid_run(gpu_device):
    """2 iterations on spectra, then 1 in real space (the engine transforms
    F back), against 3 in real space."""
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=2, background_frac=0.02)
    bg = np.random.default_rng(6).random((64, 64)) * 0.5
    ma, Oa, fa, ba = _run(s, gpu_device, bg, [(2, False), (1, True)])
    mb, Ob, fb, bb = _run(s, gpu_device, bg, [(2, True), (1, True)])
    print(f"object {rel_l2(Oa[-1], Ob[-1]):.1e} f {rel_l2(fa, fb):.1e}")
    assert rel_l2(Oa[-1], Ob[-1]) < 1e-5 and rel_l2(fa, fb) < 1e-5
    assert np.allclose(ma[:, 0], mb[:, 0], rtol=1e-5)


# ---- method 1 ---------------------------------------------------------------

@pytest.mark.parametrize("frame,grid,step", CASES)
def test_method1_spectral_and_real_space_paths_agree(frame, grid, step, gpu_device):
    """Every iteration's metrics (the model error formed one iteration later
    or in the tail) and the objects after 1 and 2 iterations on the
    illuminated region (sum |P|^2 >= 1e-6 of its maximum).  Bar per
    iteration: max(1e-5, 10 x the difference between the complex64 and the
    complex128 oracle runs from the same start) -- the rounding two complex64
    evaluations of the same step may differ by; without background the
    weakly lit border divides round-off by a small heatmap."""
    import oracle
    print(f"DM method 1 spectra vs real space, frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=2)
    H = oracle.make_P_heatmap(s.P, s.R - s.R.max(axis=0), s.O.shape)
    region = H >= 1e-6 * H.max()
    O0 = broadband_start(s.O.shape).astype(np.complex64)
    ref = {}
    for dt in ("double", "single"):
        tr = []
        oracle.DM(s.I, s.R.copy(), s.P.copy(), O0.astype(np.complex128), 2, method=1, dtype=dt,
                  callback=lambda i, O_, P_: tr.append(O_.copy()))
        ref[dt] = tr
    bound = [max(1e-5, 10 * rel_l2(a[region], b[region]))
             for a, b in zip(ref["single"], ref["double"])]
    ma, Oa, fa, _ = _run(s, gpu_device, None, [(1, False), (1, False)], method=1)
    mb, Ob, fb, _ = _run(s, gpu_device, None, [(1, True), (1, True)], method=1)
    errs = [rel_l2(x[region], y[region]) for x, y in zip(Oa, Ob)]
    print("object rel L2 after 1 / 2:", " ".join(f"{e:.1e}" for e in errs),
          "bound", " ".join(f"{e:.1e}" for e in bound), f"f {rel_l2(fa, fb):.1e}")
    assert errs[0] < 1e-6 and errs[1] < bound[1] and rel_l2(fa, fb) < 1e-5
    assert np.allclose(ma[:, 0], mb[:, 0], rtol=1e-5)
    assert np.allclose(ma[:, 1], mb[:, 1], rtol=1e-4)
    assert np.allclose(ma[:, 2], mb[:, 2], rtol=1e-4)


@pytest.mark.parametrize("graphs", [False, True])
def test_method1_split_runs_equal_one_run(graphs, gpu_device):
    """The run's tail writes the last metrics; the next run rebuilds O from
    the spectra through the same kernels: 2 + 4 iterations = 6 bit for bit,
    eager and with graph replay."""
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=2)
    ma, Oa, fa, _ = _run(s, gpu_device, None, [(6, False)], method=1, graphs=graphs)
    mb, Ob, fb, _ = _run(s, gpu_device, None, [(2, False), (4, False)], method=1, graphs=graphs)
    assert np.array_equal(Oa[-1], Ob[-1])
    assert np.array_equal(fa, fb)
    assert np.array_equal(ma, mb)


def test_method1_graphs_equal_eager(gpu_device):
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=2)
    ma, Oa, fa, _ = _run(s, gpu_device, None, [(5, False)], method=1, graphs=False)
    mb, Ob, fb, _ = _run(s, gpu_device, None, [(5, False)], method=1, graphs=True)
    assert np.array_equal(Oa[-1], Ob[-1]) and np.array_equal(fa, fb) and np.array_equal(ma, mb)


# ---- method 3 (object + probe every iteration) --------------------------------

def _probe_start(s):
    rng = np.random.default_rng(17)
    return s.P * (1.0 + 0.2 * rng.standard_normal(s.P.shape))


@pytest.mark.parametrize("frame,grid,step", CASES)
def test_method3_spectral_and_real_space_paths_agree(frame, grid, step, gpu_device):
    """As for method 1, with the probe (whole array) under the same bar."""
    import oracle
    print(f"DM method 3 spectra vs real space, frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=2)
    P0 = _probe_start(s)
    H = oracle.make_P_heatmap(s.P, s.R - s.R.max(axis=0), s.O.shape)
    region = H >= 1e-6 * H.max()
    O0 = broadband_start(s.O.shape).astype(np.complex64)
    ref = {}
    for dt in ("double", "single"):
        tr = []
        oracle.DM(s.I, s.R.copy(), P0.astype(np.complex64).astype(np.complex128),
                  O0.astype(np.complex128), 2, method=3, dtype=dt,
                  callback=lambda i, O_, P_: tr.append((O_.copy(), P_.copy())))
        ref[dt] = tr
    bO = [max(1e-5, 10 * rel_l2(a[0][region], b[0][region])) for a, b in zip(ref["single"], ref["double"])]
    bP = [max(1e-5, 10 * rel_l2(a[1], b[1])) for a, b in zip(ref["single"], ref["double"])]
    Pa, Pb = [], []
    ma, Oa, fa, _ = _run(s, gpu_device, None, [(1, False), (1, False)], method=3, P0=P0, Ps=Pa)
    mb, Ob, fb, _ = _run(s, gpu_device, None, [(1, True), (1, True)], method=3, P0=P0, Ps=Pb)
    eO = [rel_l2(x[region], y[region]) for x, y in zip(Oa, Ob)]
    eP = [rel_l2(x, y) for x, y in zip(Pa, Pb)]
    print("object", " ".join(f"{e:.1e}" for e in eO), "bound", " ".join(f"{e:.1e}" for e in bO))
    print("probe ", " ".join(f"{e:.1e}" for e in eP), "bound", " ".join(f"{e:.1e}" for e in bP),
          f"f {rel_l2(fa, fb):.1e}")
    assert eO[0] < 1e-6 and eP[0] < 1e-6
    assert eO[1] < bO[1] and eP[1] < bP[1] and rel_l2(fa, fb) < 1e-5
    assert np.allclose(ma[:, 0], mb[:, 0], rtol=1e-5)
    assert np.allclose(ma[:, 1:], mb[:, 1:], rtol=1e-4)


@pytest.mark.parametrize("graphs", [False, True])
def test_method3_split_runs_equal_one_run(graphs, gpu_device):
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=2)
    P0 = _probe_start(s)
    Pa, Pb = [], []
    ma, Oa, fa, _ = _run(s, gpu_device, None, [(6, False)], method=3, graphs=graphs, P0=P0, Ps=Pa)
    mb, Ob, fb, _ = _run(s, gpu_device, None, [(2, False), (4, False)], method=3, graphs=graphs,
                         P0=P0, Ps=Pb)
    assert np.array_equal(Oa[-1], Ob[-1]) and np.array_equal(Pa[-1], Pb[-1])
    assert np.array_equal(fa, fb)
    assert np.array_equal(ma, mb)
