"""GPU ERA / DM against the NumPy oracle on identical seeded inputs.

Bar (stated tolerances): iterates agree to relative L2 <= 1e-5 over the first
20 iterations on the illuminated region (sum |P|^2 >= 1e-6 max); per-iteration
eMod / eCon / eSup agree to 1e-4 relative; final eMod within 1 %.
"""

import numpy as np
import pytest

import oracle
import ptychography_amd as pty
from tests.scans import broadband_start, flat_start, rel_l2, small_scan

pytestmark = pytest.mark.gpu

ITER_TOL = 1e-5
METRIC_TOL = 1e-4


def _run_both(fn_gpu, fn_ref, args, kwargs, iters):
    """Run both solvers, recording O and P after every iteration."""
    traj_g, traj_r = [], []
    I, R, P, O = args
    out_g = fn_gpu(I, R.copy(), None if P is None else P.copy(), None if O is None else O.copy(),
                   iters, callback=lambda i, O_, P_: traj_g.append((O_.copy(), P_.copy())), **kwargs)
    out_r = fn_ref(I, R.copy(), None if P is None else P.copy(), None if O is None else O.copy(),
                   iters, callback=lambda i, O_, P_: traj_r.append((O_.copy(), P_.copy())), **kwargs)
    return out_g, out_r, traj_g, traj_r


def _O0(s):
    return broadband_start(s.O.shape)


def _illuminated(P, R, O_shape):
    H = oracle.make_P_heatmap(P, R - R.max(axis=0), O_shape)
    return H >= 1e-6 * H.max()


def _check_traj(traj_g, traj_r, region, upd, tol=ITER_TOL):
    worst = 0.0
    for k, ((Og, Pg), (Or, Pr)) in enumerate(zip(traj_g, traj_r)):
        if upd in ("O", "OP"):
            e = rel_l2(Og[region], Or[region])
            worst = max(worst, e)
            assert e <= tol, f"iteration {k}: object rel L2 {e:.2e}"
        if upd in ("P", "OP"):
            e = rel_l2(Pg, Pr)
            worst = max(worst, e)
            assert e <= tol, f"iteration {k}: probe rel L2 {e:.2e}"
    return worst


def _check_metrics(ig, ir, keys=("eMod", "eCon", "eSup"), tol=METRIC_TOL):
    for key in keys:
        g = np.array(ig[key])
        r = np.array(ir[key])
        assert g.shape == r.shape
        err = np.abs(g - r) / np.maximum(np.abs(r), 1e-12)
        assert err.max() <= tol, f"{key}: worst relative error {err.max():.2e} at {err.argmax()}"


def test_era_method1_twenty_iterations(gpu_device):
    print("ERA method 1, 20 iterations", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=1)
    (Og, ig), (Or, ir), tg, tr = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, s.P, _O0(s)),
                                           dict(method=1), 20)
    region = _illuminated(s.P, s.R, Or.shape)
    worst = _check_traj(tg, tr, region, "O")
    print("worst iterate rel L2", worst)
    _check_metrics(ig, ir)
    assert Og.dtype == np.complex128
    assert rel_l2(ig["I"], ir["I"]) < 1e-5
    assert rel_l2(ig["heatmap"], ir["heatmap"]) < 1e-6


def test_era_masked_flat_probe_start(gpu_device):
    print("ERA method 3, masked, random probe start", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1, beamstop_radius=2.0, dead_frac=0.05)
    P0 = flat_start(s)
    (Og, Pg, ig), (Or, Pr, ir), tg, tr = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, P0, _O0(s)),
                                                   dict(method=3, mask=s.mask), 20)
    region = _illuminated(s.P, s.R, Or.shape)
    _check_traj(tg, tr, region, "OP", tol=1e-4)
    _check_metrics(ig, ir, tol=1e-3)


@pytest.mark.parametrize("method", [2, 3])
def test_era_probe_methods(method, gpu_device):
    print(f"ERA method {method}", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = flat_start(s)
    O0 = s.O if method == 2 else _O0(s)
    out_g, out_r, tg, tr = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, P0, O0),
                                     dict(method=method), 15)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "P" if method == 2 else "OP", tol=1e-4)
    _check_metrics(out_g[-1], out_r[-1], tol=1e-3)


def test_era_op_schedule_centering_pmodp(gpu_device):
    print("ERA method 3, OP_iters=(2,3), centring, Pmod_probe=4", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = np.roll(flat_start(s), (3, -2), axis=(0, 1))
    out_g, out_r, tg, tr = _run_both(
        pty.ERA, oracle.ERA, (s.I, s.R, P0, _O0(s)),
        dict(method=3, OP_iters=(2, 3), probe_centering=True, Pmod_probe=4), 12)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "OP", tol=1e-4)
    _check_metrics(out_g[-1], out_r[-1], tol=1e-3)


@pytest.mark.parametrize("method", [4, 6])
def test_era_background_methods(method, gpu_device):
    print(f"ERA method {method} (background)", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1, background_frac=0.02)
    P0 = s.P if method == 4 else flat_start(s)
    rng = np.random.default_rng(21)
    B0 = rng.random((32, 32)) * 0.5
    out_g, out_r, tg, tr = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, P0, _O0(s)),
                                     dict(method=method, background=B0), 12)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "O" if method == 4 else "OP", tol=1e-4)
    _check_metrics(out_g[-1], out_r[-1], tol=1e-3)
    assert rel_l2(out_g[-1]["background"], out_r[-1]["background"]) < 1e-4


def test_dm_method1(gpu_device):
    print("DM method 1", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=1)
    (Og, ig), (Or, ir), tg, tr = _run_both(pty.DM, oracle.DM, (s.I, s.R, s.P, _O0(s)),
                                           dict(method=1), 20)
    region = _illuminated(s.P, s.R, Or.shape)
    _check_traj(tg, tr, region, "O")
    _check_metrics(ig, ir)
    assert rel_l2(ig["I"], ir["I"]) < 1e-5


@pytest.mark.parametrize("method", [2, 3])
def test_dm_probe_methods(method, gpu_device):
    print(f"DM method {method}", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = flat_start(s)
    O0 = s.O if method == 2 else _O0(s)
    out_g, out_r, tg, tr = _run_both(pty.DM, oracle.DM, (s.I, s.R, P0, O0),
                                     dict(method=method), 12)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "P" if method == 2 else "OP", tol=1e-4)
    _check_metrics(out_g[-1], out_r[-1], tol=1e-3)


def test_dm_centering_schedule(gpu_device):
    print("DM method 3, OP_iters=(1,2), centring", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = np.roll(flat_start(s), (2, 3), axis=(0, 1))
    out_g, out_r, tg, tr = _run_both(pty.DM, oracle.DM, (s.I, s.R, P0, _O0(s)),
                                     dict(method=3, OP_iters=(1, 2), probe_centering=True), 10)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "OP", tol=1e-4)
    _check_metrics(out_g[-1], out_r[-1], tol=1e-3)


@pytest.mark.parametrize("method", [4, 6])
def test_dm_background_methods(method, gpu_device):
    print(f"DM method {method} (background)", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1, background_frac=0.02,
                   beamstop_radius=1.5)
    P0 = s.P if method == 4 else flat_start(s)
    B0 = np.random.default_rng(4).random((32, 32)) * 0.5
    out_g, out_r, tg, tr = _run_both(pty.DM, oracle.DM, (s.I, s.R, P0, _O0(s)),
                                     dict(method=method, background=B0, mask=s.mask), 10)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "O" if method == 4 else "OP", tol=1e-4)
    _check_metrics(out_g[-1], out_r[-1], tol=1e-3)
    assert rel_l2(out_g[-1]["background"], out_r[-1]["background"]) < 1e-4


def test_r_shifted_in_place_and_errors(gpu_device):
    s = small_scan(frame=32, grid=(3, 3), step=8)
    R = s.R.copy() - 5
    pty.ERA(s.I, R, s.P, None, 1, method=1)
    assert R.max(axis=0).tolist() == [0, 0]
    with pytest.raises(ValueError):
        pty.ERA(s.I, s.R.copy(), s.P, s.O, 1)          # method undefined in the reference
    from ptychography_amd._abi import PtyError
    with pytest.raises(PtyError):
        pty.ERA(s.I, s.R.copy(), s.P, np.ones((20, 20)), 1, method=1)  # windows outside O
