"""GPU ERA / DM against the NumPy oracle on identical seeded inputs.

Bar (stated tolerances), per iteration over the first 12-20 iterations:

* iterate error  |x_gpu - x_ref|_2 / |x_ref|_2  (object on the illuminated
  region, sum |P|^2 >= 1e-6 max; probe everywhere) must not exceed
  max(1e-5, 10 * max_{j<=k} d_j), where d_j is the same error of the oracle
  run in complex64 (the reference's own dtype='single' path) against the
  complex128 oracle at iteration j -- GPU complex64 must stay within the same
  order of magnitude as the reference's own single-precision mode.  ERA method 1 stays under 1e-5 outright; DM and
  probe-updating runs amplify round-off (their complex64 NumPy runs drift to
  1e-4..1e-3 within 10-20 iterations), so they are held to the drift bound;
* eMod / eCon / eSup per iteration: relative error <= max(1e-4, 10 * the
  running max of the complex64 oracle's relative error); for DM, whose
  iterates drift furthest, also <= 10 * the complex64 oracle's iterate drift
  (the metrics are functions of the iterates);
* final eMod within 1 %.
"""

import numpy as np
import pytest

import oracle
import ptychography_amd as pty
from tests.scans import broadband_start, flat_start, rel_l2, small_scan

pytestmark = pytest.mark.gpu

ITER_TOL = 1e-5
METRIC_TOL = 1e-4
DRIFT_FACTOR = 10.0


def _traj(fn, args, kwargs, iters, **extra):
    I, R, P, O = args
    tr = []
    out = fn(I, R.copy(), None if P is None else P.copy(), None if O is None else O.copy(),
             iters, callback=lambda i, O_, P_: tr.append((O_.copy(), P_.copy())), **kwargs, **extra)
    return out, tr


def _run_both(fn_gpu, fn_ref, args, kwargs, iters):
    """GPU run, complex128 oracle run and complex64 oracle run, with iterates."""
    out_g, traj_g = _traj(fn_gpu, args, kwargs, iters)
    out_r, traj_r = _traj(fn_ref, args, kwargs, iters)
    out_s, traj_s = _traj(fn_ref, args, kwargs, iters, dtype='single')
    return out_g, out_r, traj_g, traj_r, (out_s, traj_s)


def _O0(s):
    return broadband_start(s.O.shape)


def _illuminated(P, R, O_shape):
    H = oracle.make_P_heatmap(P, R - R.max(axis=0), O_shape)
    return H >= 1e-6 * H.max()


def _errs(traj_a, traj_b, region, upd):
    out = []
    for (Oa, Pa), (Ob, Pb) in zip(traj_a, traj_b):
        e = 0.0
        if upd in ("O", "OP"):
            e = max(e, rel_l2(Oa[region], Ob[region]))
        if upd in ("P", "OP"):
            e = max(e, rel_l2(Pa, Pb))
        out.append(e)
    return np.array(out)


def _check_traj(traj_g, traj_r, region, upd, single=None, tol=ITER_TOL):
    eg = _errs(traj_g, traj_r, region, upd)
    bound = np.full_like(eg, tol)
    if single is not None:
        es = np.maximum.accumulate(_errs(single[1], traj_r, region, upd))
        bound = np.maximum(bound, DRIFT_FACTOR * es)
    print("iterate error  gpu:", " ".join(f"{x:.1e}" for x in eg))
    print("bound             :", " ".join(f"{x:.1e}" for x in bound))
    bad = np.nonzero(eg > bound)[0]
    assert bad.size == 0, f"iteration {bad[0]}: rel L2 {eg[bad[0]]:.2e} > bound {bound[bad[0]]:.2e}"
    return None if single is None else es


def _check_metrics(ig, ir, single_info=None, keys=("eMod", "eCon", "eSup"), tol=METRIC_TOL,
                   iter_drift=None):
    # eCon is a difference of successive iterates: its relative error is the
    # iterate error divided by eCon itself (eCon ~ 1e-2 turns a 1e-4 iterate
    # drift into 1e-2), so it gets a 10x looser floor and 3x the drift factor.
    # The kernels' own eCon error is pinned separately, step by step, by
    # test_gpu_step_parity.py: within 1e-5 relative or 6e-6 absolute per step
    # for every method, from the oracle's own state.
    for key in keys:
        tol = METRIC_TOL * (10 if key == "eCon" else 1)
        factor = DRIFT_FACTOR * (3 if key == "eCon" else 1)
        g = np.array(ig[key])
        r = np.array(ir[key])
        assert g.shape == r.shape
        err = np.abs(g - r) / np.maximum(np.abs(r), 1e-12)
        bound = np.full_like(err, tol)
        if single_info is not None:
            s = np.array(single_info[key])
            ds = np.maximum.accumulate(np.abs(s - r) / np.maximum(np.abs(r), 1e-12))
            bound = np.maximum(bound, factor * ds)
        if iter_drift is not None:
            # the metrics are functions of the iterates: where the complex64
            # oracle's iterates have drifted by d, the metrics may differ by
            # up to DRIFT_FACTOR * d as well
            bound = np.maximum(bound, DRIFT_FACTOR * np.asarray(iter_drift)[:len(bound)])
        bad = np.nonzero(err > bound)[0]
        assert bad.size == 0, f"{key}: iteration {bad[0]} rel err {err[bad[0]]:.2e} > {bound[bad[0]]:.2e}"
    assert abs(ig["eMod"][-1] - ir["eMod"][-1]) <= 0.01 * ir["eMod"][-1]


def test_era_method1_twenty_iterations(gpu_device):
    print("ERA method 1, 20 iterations", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=1)
    (Og, ig), (Or, ir), tg, tr, single = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, s.P, _O0(s)),
                                                   dict(method=1), 20)
    region = _illuminated(s.P, s.R, Or.shape)
    _check_traj(tg, tr, region, "O", single)
    _check_metrics(ig, ir, single[0][1])
    assert Og.dtype == np.complex128
    assert rel_l2(ig["I"], ir["I"]) < 1e-5
    assert rel_l2(ig["heatmap"], ir["heatmap"]) < 1e-6


def test_era_masked_flat_probe_start(gpu_device):
    print("ERA method 3, masked, random probe start", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1, beamstop_radius=2.0, dead_frac=0.05)
    P0 = flat_start(s)
    (Og, Pg, ig), (Or, Pr, ir), tg, tr, single = _run_both(
        pty.ERA, oracle.ERA, (s.I, s.R, P0, _O0(s)), dict(method=3, mask=s.mask), 20)
    region = _illuminated(s.P, s.R, Or.shape)
    _check_traj(tg, tr, region, "OP", single)
    _check_metrics(ig, ir, single[0][-1])


@pytest.mark.parametrize("method", [2, 3])
def test_era_probe_methods(method, gpu_device):
    print(f"ERA method {method}", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = flat_start(s)
    O0 = s.O if method == 2 else _O0(s)
    out_g, out_r, tg, tr, single = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, P0, O0),
                                     dict(method=method), 15)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "P" if method == 2 else "OP", single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1])


def test_era_op_schedule_centering_pmodp(gpu_device):
    print("ERA method 3, OP_iters=(2,3), centring, Pmod_probe=4", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = np.roll(flat_start(s), (3, -2), axis=(0, 1))
    out_g, out_r, tg, tr, single = _run_both(
        pty.ERA, oracle.ERA, (s.I, s.R, P0, _O0(s)),
        dict(method=3, OP_iters=(2, 3), probe_centering=True, Pmod_probe=4), 12)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "OP", single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1])


@pytest.mark.parametrize("method,frame", [(4, 32), (6, 32), (4, 64)])
def test_era_background_methods(method, frame, gpu_device):
    # 64^2 frames: object-only steps run the fused inverse-row / object-update kernel
    print(f"ERA method {method} (background), frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=(5, 5), step=frame * 3 // 16, jitter=1, background_frac=0.02)
    P0 = s.P if method == 4 else flat_start(s)
    rng = np.random.default_rng(21)
    B0 = rng.random((frame, frame)) * 0.5
    out_g, out_r, tg, tr, single = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, P0, _O0(s)),
                                     dict(method=method, background=B0), 12)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, "O" if method == 4 else "OP", single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1])
    eb = rel_l2(single[0][-1]["background"], out_r[-1]["background"])
    assert rel_l2(out_g[-1]["background"], out_r[-1]["background"]) < max(1e-4, DRIFT_FACTOR * eb)


def test_dm_method1(gpu_device):
    print("DM method 1", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=1)
    (Og, ig), (Or, ir), tg, tr, single = _run_both(pty.DM, oracle.DM, (s.I, s.R, s.P, _O0(s)),
                                                   dict(method=1), 20)
    region = _illuminated(s.P, s.R, Or.shape)
    _check_traj(tg, tr, region, "O", single)
    _check_metrics(ig, ir, single[0][1])
    eI = rel_l2(single[0][1]["I"], ir["I"])
    assert rel_l2(ig["I"], ir["I"]) < max(1e-5, DRIFT_FACTOR * eI)


@pytest.mark.parametrize("method", [2, 3])
def test_dm_probe_methods(method, gpu_device):
    print(f"DM method {method}", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = flat_start(s)
    O0 = s.O if method == 2 else _O0(s)
    out_g, out_r, tg, tr, single = _run_both(pty.DM, oracle.DM, (s.I, s.R, P0, O0),
                                     dict(method=method), 20)
    region = _illuminated(s.P, s.R, s.O.shape)
    drift = _check_traj(tg, tr, region, "P" if method == 2 else "OP", single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1], iter_drift=drift)


def test_dm_centering_schedule(gpu_device):
    print("DM method 3, OP_iters=(1,2), centring", flush=True)
    s = small_scan(frame=32, grid=(5, 5), step=6, jitter=1)
    P0 = np.roll(flat_start(s), (2, 3), axis=(0, 1))
    out_g, out_r, tg, tr, single = _run_both(pty.DM, oracle.DM, (s.I, s.R, P0, _O0(s)),
                                     dict(method=3, OP_iters=(1, 2), probe_centering=True), 20)
    region = _illuminated(s.P, s.R, s.O.shape)
    drift = _check_traj(tg, tr, region, "OP", single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1], iter_drift=drift)


@pytest.mark.parametrize("method,frame,op", [(4, 32, 1), (6, 32, 1), (4, 64, 1), (6, 64, (1, 2))])
def test_dm_background_methods(method, frame, op, gpu_device):
    # at 64^2 the object-only projections after the leading Pmod_7 run fused
    # into its inverse row FFT (method 6 on its non-OP iterations)
    print(f"DM method {method} (background), frame {frame}, OP_iters {op}", flush=True)
    s = small_scan(frame=frame, grid=(5, 5), step=frame * 3 // 16, jitter=1, background_frac=0.02,
                   beamstop_radius=1.5)
    P0 = s.P if method == 4 else flat_start(s)
    B0 = np.random.default_rng(4).random((frame, frame)) * 0.5
    out_g, out_r, tg, tr, single = _run_both(pty.DM, oracle.DM, (s.I, s.R, P0, _O0(s)),
                                     dict(method=method, background=B0, mask=s.mask, OP_iters=op), 20)
    region = _illuminated(s.P, s.R, s.O.shape)
    drift = _check_traj(tg, tr, region, "O" if method == 4 else "OP", single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1], iter_drift=drift)
    eb = rel_l2(single[0][-1]["background"], out_r[-1]["background"])
    assert rel_l2(out_g[-1]["background"], out_r[-1]["background"]) < max(1e-4, DRIFT_FACTOR * eb)


@pytest.mark.parametrize("method,op", [(1, (1, 1)), (3, (1, 3))])
def test_dm_default_path_without_callback(method, op, gpu_device):
    # without a callback one pty_run covers all iterations, so an object-only
    # projection starts from the previous iteration's metric projection (the
    # per-iteration calls of the tests above never take that path)
    # With OP_iters (1, 3) the reference's DM itself diverges on this scan
    # after about 7 iterations (eMod 37 at the 9th, 1.8e7 at the 18th, in both
    # precisions of the oracle), so method 3 runs 6 iterations (the projection
    # is reused at iterations 2 and 5), from a perturbed true probe.
    iters = 20 if method == 1 else 6
    print(f"DM method {method}, OP_iters {op}, no callback, {iters} iterations", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, jitter=1)
    P0 = s.P if method == 1 else s.P * (1.0 + 0.2 * np.random.default_rng(8).standard_normal(s.P.shape))
    args = (s.I, s.R.copy(), P0.copy(), _O0(s))
    kw = dict(method=method, OP_iters=op)
    out_g = pty.DM(*args, iters, device=gpu_device, **kw)
    out_r = oracle.DM(s.I, s.R.copy(), P0.copy(), _O0(s), iters, **kw)
    out_s = oracle.DM(s.I, s.R.copy(), P0.copy(), _O0(s), iters, dtype='single', **kw)
    _check_metrics(out_g[-1], out_r[-1], out_s[-1])
    region = _illuminated(s.P, s.R, s.O.shape)
    eO = rel_l2(out_g[0][region], out_r[0][region])
    bO = max(ITER_TOL, DRIFT_FACTOR * rel_l2(out_s[0][region], out_r[0][region]))
    print(f"final object {eO:.2e} (bound {bO:.2e})")
    assert eO <= bO
    if method == 3:
        eP = rel_l2(out_g[1], out_r[1])
        assert eP <= max(ITER_TOL, DRIFT_FACTOR * rel_l2(out_s[1], out_r[1]))


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


@pytest.mark.parametrize("fused", [True, False])
def test_era_and_dm_128_frames_both_paths(fused, gpu_device):
    """128^2 frames through the fused per-frame kernel (opt-in) and through the
    staged kernels (the default): both hold the oracle parity bar."""
    print(f"ERA/DM at 128^2, fused={fused}", flush=True)
    from ptychography_amd import engine
    s = small_scan(frame=128, grid=(4, 4), step=24, jitter=2)
    engine.FUSED = fused
    try:
        (Og, ig), (Or, ir), tg, tr, single = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, s.P, _O0(s)),
                                                       dict(method=1), 10)
        region = _illuminated(s.P, s.R, Or.shape)
        _check_traj(tg, tr, region, "O", single)
        _check_metrics(ig, ir, single[0][1])
        assert rel_l2(ig["I"], ir["I"]) < 1e-5
        out_g, out_r, tg, tr, single = _run_both(pty.DM, oracle.DM, (s.I, s.R, flat_start(s), _O0(s)),
                                                 dict(method=3), 8)
        _check_traj(tg, tr, region, "OP", single)
        _check_metrics(out_g[-1], out_r[-1], single[0][-1])
        eI = rel_l2(single[0][-1]["I"], out_r[-1]["I"])
        assert rel_l2(out_g[-1]["I"], out_r[-1]["I"]) < max(1e-4, DRIFT_FACTOR * eI)
    finally:
        engine.FUSED = None


def test_fused_and_staged_paths_agree(gpu_device):
    import torch
    from ptychography_amd.engine import Engine
    s = small_scan(frame=256, grid=(3, 3), step=64, jitter=2)
    N, M, K = s.I.shape
    outs = []
    for fused in (True, False):
        eng = Engine(N, M, K, *s.O.shape, s.R.astype(np.int32), device=gpu_device)
        eng.set_path(fused)
        amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), None)
        eng.set_data(amp, None, I_norm)
        O_t = torch.from_numpy(_O0(s).astype(np.complex64)).to(gpu_device)
        P_t = torch.from_numpy(s.P.astype(np.complex64)).to(gpu_device)
        eng.set_state(O_t, P_t)
        m = eng.run(0, 1, 3, begin=True)
        outs.append((O_t.cpu().numpy(), m, eng.intensity().cpu().numpy()))
        eng.close()
    (Of, mf, If), (Os, ms, Is) = outs
    assert rel_l2(Of, Os) < 1e-5
    assert np.allclose(mf[:, 0], ms[:, 0], rtol=1e-5)
    assert rel_l2(If, Is) < 1e-5
