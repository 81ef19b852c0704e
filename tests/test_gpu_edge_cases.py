"""Edge cases of the ERA / DM call surface against the oracle.

* iters = 0: the start is returned and the metric lists are empty;
* one frame whose window is the whole object (N = 1): fitted exactly, so
  eMod is held to max(1.01 x the oracle's, 1e-6) -- complex64's round-off floor;
* O = None (the reference's default shape) and P = None drawn from the same
  seeded generator on both sides (the reference draws np.random.random);
* complex128 inputs give complex128 outputs;
* a mask of ones gives bit-identical results to mask = 1.

Bars as in test_gpu_solvers.py: iterates within max(1e-5, 10 x the complex64
oracle's drift), final eMod within 1 %.
"""

import numpy as np
import pytest

import oracle
import ptychography_amd as pty
from tests.scans import rel_l2, small_scan
from tests.test_gpu_solvers import _O0, _check_traj, _illuminated, _run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["ERA", "DM"])
def test_zero_iterations_return_the_start(solver, gpu_device):
    print(f"{solver} with iters = 0", flush=True)
    s = small_scan(frame=32, grid=(3, 3), step=8)
    O0 = _O0(s)
    fn_g, fn_r = getattr(pty, solver), getattr(oracle, solver)
    Og, Pg, ig = fn_g(s.I, s.R.copy(), s.P, O0, 0, method=3)
    Or, Pr, ir = fn_r(s.I, s.R.copy(), s.P, O0, 0, method=3)
    assert len(ig["eMod"]) == len(ir["eMod"]) == 0
    assert rel_l2(Og, Or) <= 1e-6 and rel_l2(Pg, Pr) <= 1e-6


def test_single_frame_covering_the_object(gpu_device):
    print("ERA method 3 with one frame", flush=True)
    s = small_scan(frame=32, grid=(1, 1), step=8, jitter=0)
    assert s.I.shape[0] == 1 and s.O.shape == (32, 32)
    out_g, out_r, tg, tr, single = _run_both(pty.ERA, oracle.ERA, (s.I, s.R, s.P, _O0(s)),
                                             dict(method=3), 6)
    _check_traj(tg, tr, np.ones(s.O.shape, bool), "OP", single)
    # one frame is fitted exactly: the complex128 oracle's eMod falls to ~1e-10,
    # complex64 stops at its round-off (~1e-7), so the bar is that floor
    eg, er = np.array(out_g[-1]["eMod"]), np.array(out_r[-1]["eMod"])
    print("eMod gpu   :", " ".join(f"{x:.1e}" for x in eg))
    print("eMod oracle:", " ".join(f"{x:.1e}" for x in er))
    assert np.all(eg <= np.maximum(1.01 * er, 1e-6))


def test_default_object_and_random_probe(gpu_device):
    print("ERA method 1 with O = None, P = None (seeded)", flush=True)
    s = small_scan(frame=32, grid=(4, 4), step=8)
    # method 1 returns (O, info); the probe both sides drew is rebuilt from the seed
    Og, ig = pty.ERA(s.I, s.R.copy(), None, None, 5, method=1, rng=np.random.default_rng(5))
    Or, ir = oracle.ERA(s.I, s.R.copy(), None, None, 5, method=1, rng=np.random.default_rng(5))
    assert Og.shape == Or.shape
    rng = np.random.default_rng(5)
    P = rng.random((32, 32)) + 1j * rng.random((32, 32))
    region = _illuminated(P, s.R, Or.shape)
    err = rel_l2(Og[region], Or[region])
    print(f"object rel L2 {err:.2e}")
    assert err <= 1e-5
    assert abs(ig["eMod"][-1] - ir["eMod"][-1]) <= 0.01 * ir["eMod"][-1]


def test_double_precision_inputs_give_double_outputs(gpu_device):
    print("ERA method 1, complex128 in / out", flush=True)
    s = small_scan(frame=32, grid=(3, 3), step=8)
    I = s.I.astype(np.float64)
    P = s.P.astype(np.complex128)
    O0 = _O0(s).astype(np.complex128)
    Og, ig = pty.ERA(I, s.R.copy(), P, O0, 4, method=1)
    Or, ir = oracle.ERA(I, s.R.copy(), P, O0, 4, method=1)
    assert Og.dtype == np.complex128
    region = _illuminated(s.P, s.R, s.O.shape)
    assert rel_l2(Og[region], Or[region]) <= 1e-5


@pytest.mark.parametrize("solver", ["ERA", "DM"])
def test_mask_of_ones_equals_no_mask(solver, gpu_device):
    print(f"{solver}: mask of ones vs mask = 1", flush=True)
    s = small_scan(frame=64, grid=(4, 4), step=12)
    fn = getattr(pty, solver)
    a = fn(s.I, s.R.copy(), s.P, _O0(s), 5, method=3)
    b = fn(s.I, s.R.copy(), s.P, _O0(s), 5, method=3, mask=np.ones((64, 64), bool))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(np.asarray(a[2]["eMod"]), np.asarray(b[2]["eMod"]))


def test_progress_bar_prints_and_leaves_results_alone(gpu_device, capsys):
    import ptychography_amd as pty
    from tests.scans import small_scan
    s = small_scan(frame=32, grid=(3, 3), step=8)
    O1, i1 = pty.ERA(s.I, s.R.copy(), s.P, None, 4, method=1, device=gpu_device)
    capsys.readouterr()
    O2, i2 = pty.ERA(s.I, s.R.copy(), s.P, None, 4, method=1, device=gpu_device, progress=True)
    out = capsys.readouterr().out
    assert out.startswith("algrithm progress iteration convergence modulus error\n")
    assert out.count("\rERA: [") == 4 and "Done...\r\n" in out
    assert np.array_equal(O1, O2) and i1["eMod"] == i2["eMod"]
    assert f" 3 {i2['eCon'][-1]:.12g} {i2['eMod'][-1]:.12g} " in out
