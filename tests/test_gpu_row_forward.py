"""The forward row stage in object-row order (k_rows_fwd_orow) against the
frame-order kernel (k_rows_fwd, selected with PTY_ROW_FWD_OROW=0 at context
creation), for the three sources that read the object: O_win P (ERA, DM on
spectra, the metric passes), 2 O_win P - f (DM in real space) and
f + w O_win P (HIO).

Both form every frame row with the same complex arithmetic and transform it
with the same line FFT; only where the object values come from differs (LDS
staging of the object-row segment vs global loads of each window).  So every
result is bit-identical: ERA, DM and HIO iterates, metrics and info, for
objects of one and of several column segments, on the alternative row plan,
with packed scans and with sub-pixel (per-frame) probes."""

import numpy as np
import pytest

from ptychography_amd import DM, ERA, ERA_many, HIO
from tests.scans import broadband_start, small_scan

pytestmark = pytest.mark.gpu


def _both(fn, monkeypatch):
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PTY_ROW_FWD_OROW", flag)
        out.append(fn())
    monkeypatch.delenv("PTY_ROW_FWD_OROW")
    return out


def _same(a, b):
    # (O, info) or (O, P, info): every array and every metric bit for bit
    assert len(a) == len(b)
    for x, y in zip(a[:-1], b[:-1]):
        assert np.array_equal(x, y)
    for k in ("eMod", "eCon", "eSup", "exits", "I", "heatmap"):
        if k in a[-1]:
            assert np.array_equal(np.asarray(a[-1][k]), np.asarray(b[-1][k])), k


@pytest.mark.parametrize("frame,grid,step,method", [
    (32, (5, 5), 8, 1), (64, (5, 5), 12, 1), (96, (4, 4), 20, 3), (128, (3, 4), 40, 1),
    (192, (2, 3), 48, 3), (256, (3, 3), 40, 1), (256, (3, 3), 40, 3)])
def test_era_object_row_forward_is_bit_identical(frame, grid, step, method, gpu_device, monkeypatch):
    print(f"row forward ERA method {method} frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=grid, step=step, jitter=2, seed=60 + frame,
                   beamstop_radius=frame / 32, dead_frac=0.03)
    O0 = broadband_start(s.O.shape, seed=frame)
    run = lambda: ERA(s.I, s.R.copy(), s.P.copy(), O0.copy(), 5, method=method, mask=s.mask,
                      device=gpu_device)
    _same(*_both(run, monkeypatch))


@pytest.mark.parametrize("plan", ["0", "1"])
def test_wide_object_several_segments(plan, gpu_device, monkeypatch):
    # an object row of ~4300 columns is cut into several segments; frames are
    # listed by the segment their window starts in and read columns past it
    print(f"row forward wide object, plan mask {plan}", flush=True)
    monkeypatch.setenv("PTY_FFT_PLAN", plan)
    s = small_scan(frame=256, grid=(2, 32), step=130, jitter=2, seed=71)
    assert s.O.shape[1] > 4000
    O0 = broadband_start(s.O.shape, seed=71)
    run = lambda: ERA(s.I, s.R.copy(), s.P.copy(), O0.copy(), 3, method=1, device=gpu_device)
    _same(*_both(run, monkeypatch))


@pytest.mark.parametrize("method", [1, 3, 4])
def test_dm_object_row_forward_is_bit_identical(method, gpu_device, monkeypatch):
    print(f"row forward DM method {method}", flush=True)
    bg = method == 4
    s = small_scan(frame=64, grid=(4, 4), step=16, seed=72, beamstop_radius=3.0, dead_frac=0.05,
                   background_frac=0.01 if bg else 0.0)
    O0 = broadband_start(s.O.shape, seed=72)
    kw = {"background": np.full(s.I.shape[1:], 0.01 * s.I.mean())} if bg else {}
    run = lambda: DM(s.I, s.R.copy(), s.P.copy(), O0.copy(), 6, method=method, mask=s.mask,
                     device=gpu_device, **kw)
    _same(*_both(run, monkeypatch))


@pytest.mark.parametrize("method", [1, 3, 4])
def test_dm_real_space_is_bit_identical(method, gpu_device, monkeypatch):
    # DM on f in real space: the row pass forms 2 O_win P - f
    print(f"row forward DM method {method} in real space", flush=True)
    monkeypatch.setenv("PTY_DM_FOURIER", "0")
    bg = method == 4
    s = small_scan(frame=64, grid=(4, 4), step=16, seed=76, beamstop_radius=3.0, dead_frac=0.05,
                   background_frac=0.01 if bg else 0.0)
    O0 = broadband_start(s.O.shape, seed=76)
    kw = {"background": np.full(s.I.shape[1:], 0.01 * s.I.mean())} if bg else {}
    run = lambda: DM(s.I, s.R.copy(), s.P.copy(), O0.copy(), 6, method=method, mask=s.mask,
                     device=gpu_device, **kw)
    _same(*_both(run, monkeypatch))


@pytest.mark.parametrize("update", ["sample", "both"])
def test_hio_is_bit_identical(update, gpu_device, monkeypatch):
    # HIO's row pass forms f + w O_win P
    print(f"row forward HIO update={update}", flush=True)
    s = small_scan(frame=64, grid=(5, 5), step=12, seed=77)
    O0 = broadband_start(s.O.shape, seed=77)
    run = lambda: HIO(s.I, s.R.copy(), s.P.copy(), O0.copy(), 6, beta=0.9, update=update,
                      device=gpu_device)
    _same(*_both(run, monkeypatch))


def test_packed_scans(gpu_device, monkeypatch):
    print("row forward packed scans", flush=True)
    specs = [dict(frame=64, grid=(4, 5), step=16, seed=73), dict(frame=64, grid=(3, 3), step=20, seed=74)]
    jobs = []
    for sp in specs:
        s = small_scan(**sp)
        jobs.append(dict(I=s.I, R=s.R, P=s.P, O=broadband_start(s.O.shape, seed=sp["seed"]), method=3))
    run = lambda: ERA_many([dict(j, R=j["R"].copy()) for j in jobs], 4, device=gpu_device)
    a, b = _both(run, monkeypatch)
    for ra, rb in zip(a, b):
        _same(ra, rb)


def test_subpixel_positions(gpu_device, monkeypatch):
    print("row forward sub-pixel positions", flush=True)
    rng = np.random.default_rng(75)
    M, G, step = 64, 4, 14.3
    yy, xx = np.mgrid[:G, :G]
    R = -np.stack([yy.ravel() * step, xx.ravel() * step], 1) - rng.uniform(0, 1, (G * G, 2))
    R -= R.max(axis=0)
    U = M + int(np.ceil(np.ptp(R[:, 0])))
    V = M + int(np.ceil(np.ptp(R[:, 1])))
    I = rng.uniform(0.0, 1e4, (G * G, M, M)).astype(np.float32)
    y, x = np.mgrid[:M, :M]
    P = np.exp(-((y - M / 2) ** 2 + (x - M / 2) ** 2) / (2 * 9.0 ** 2)).astype(np.complex128)
    O0 = broadband_start((U, V), seed=75)
    run = lambda: ERA(I, R.copy(), P.copy(), O0.copy(), 4, method=1, device=gpu_device)
    _same(*_both(run, monkeypatch))


def test_object_wider_than_the_row_lists_take(gpu_device, monkeypatch):
    # V > 8192: no object-row lists (the packed column offset holds 13 bits)
    # and no fused row gather; the frame-order kernels run and match the oracle
    import oracle
    from tests.scans import rel_l2
    print("object wider than 8192 columns", flush=True)
    s = small_scan(frame=128, grid=(2, 66), step=126, jitter=1, seed=78)
    assert s.O.shape[1] > 8192
    O0 = broadband_start(s.O.shape, seed=78)
    O_g, info = ERA(s.I, s.R.copy(), s.P.copy(), O0.copy(), 3, method=1, device=gpu_device)
    O_r, info_r = oracle.ERA(s.I, s.R.copy(), s.P.copy(), O0.copy(), 3, method=1)
    H = info_r["heatmap"]
    cov = H > 1e-3 * H.max()
    print(f"  object rel L2 {rel_l2(O_g[cov], O_r[cov]):.1e}", flush=True)
    assert rel_l2(O_g[cov], O_r[cov]) < 1e-5
    e, er = np.asarray(info["eMod"]), np.asarray(info_r["eMod"])
    assert np.max(np.abs(e - er) / er) < 1e-5
