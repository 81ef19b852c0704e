"""Pin the oracle's sub-pixel positions (oracle/subpixel.py) on the CPU.

The reference shifts by float amounts with the Fourier shift theorem
(bagOfns.roll, ptychography/bagOfns.py:729-764) and forms float-position exit
waves by shifting the object (utility_Ptych.makeExits, :42-48).  The oracle
keeps integer windows and shifts the probe of each frame by the remainder.
What fixes that model:

* the frequency numbering of the phase ramp, against the numbers the
  reference's make_xy(origin=(0, 0)) gives (bagOfns.py:471-492), which differ
  from numpy's fftfreq at the Nyquist bin;
* the shift theorem: integer shifts are np.roll, fractional shifts of a
  band-limited signal are its exact translate;
* the diffraction intensities of the probe-shift exits agree with the
  reference's object-shift exits for band-limited inputs, and a flipped sign
  or rounded positions do not;
* ERA / DM with float positions: integer-valued floats are the integer run,
  a tiny fraction changes the result by a tiny amount, data made by the
  reference's object-shift exits is a fixed point that rounded positions miss,
  and ERA's alternating-projection sandwich holds.
"""

import numpy as np
import pytest

import oracle
from oracle import subpixel as sp
from oracle.era import make_exits
from tests.scans import band_limited_scan, raster_subpixel_scan


def test_freq_index_numbers_bins_as_make_xy():
    # make_xy(8, origin=(0, 0)) along one axis, read from bagOfns.py:489-490
    assert sp.freq_index(8).tolist() == [0, 1, 2, 3, 4, -3, -2, -1]
    assert sp.freq_index(4).tolist() == [0, 1, 2, -1]
    # numpy puts the Nyquist bin at -N/2; the reference at +N/2
    assert (np.fft.fftfreq(8) * 8).tolist()[4] == -4.0


@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (16, 12)])
@pytest.mark.parametrize("shift", [(1, 0), (0, -3), (2, 5), (-7, 11), (16, -1)])
def test_integer_shift_is_np_roll(shape, shift):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    y = sp.roll(x, (float(shift[0]), float(shift[1])))
    ref = np.roll(np.roll(x, shift[0], 0), shift[1], 1)
    assert np.allclose(y, ref, atol=1e-12)


@pytest.mark.parametrize("s", [(0.5, 0.0), (-0.25, 0.375), (1.3, -2.7)])
def test_fractional_shift_translates_a_band_limited_signal(s):
    """roll(x, s)[r] = x(r - s) for x = sum c_k exp(2 pi i k r / N) with every
    |k| < N / 2 (below the Nyquist bin, where the sign convention would matter)."""
    N0, N1 = 12, 16
    rng = np.random.default_rng(2)
    ks = [(a, b) for a in range(-5, 6) for b in range(-7, 8)]
    c = rng.standard_normal(len(ks)) + 1j * rng.standard_normal(len(ks))

    def x(r0, r1):
        return sum(ck * np.exp(2j * np.pi * (a * r0 / N0 + b * r1 / N1)) for ck, (a, b) in zip(c, ks))

    r0, r1 = np.mgrid[:N0, :N1].astype(float)
    y = sp.roll(x(r0, r1), s)
    assert np.allclose(y, x(r0 - s[0], r1 - s[1]), atol=1e-10)


def test_split_positions():
    R = np.array([[0.0, -0.4], [-1.5, -2.6], [-3.2, 0.0], [-2.5, -0.5]])
    Rw, d = sp.split_positions(R)
    assert Rw.dtype == np.int64
    assert np.array_equal(-Rw + d, -R)
    assert np.all(np.abs(d) <= 0.5)
    assert Rw.tolist() == [[0, 0], [-2, -3], [-3, 0], [-2, 0]]   # rint: half to even
    assert not sp.is_fractional(np.array([[0.0, -2.0]]))
    assert not sp.is_fractional(np.array([[0, -2]]))
    assert sp.is_fractional(np.array([[0.0, -2.5]]))


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def test_probe_shift_exits_match_the_object_shift_intensities():
    O, P, R, I = band_limited_scan()
    Ir = np.fft.ifftshift(I, axes=(-2, -1))
    Rw, d = sp.split_positions(R)
    Im = np.abs(np.fft.fft2(make_exits(O, sp.frame_probes(P, d), Rw))) ** 2
    assert _rel(Im, Ir) < 1e-6
    # PLANTED: the probe shifted the other way, or the positions rounded
    Ib = np.abs(np.fft.fft2(make_exits(O, sp.frame_probes(P, -d), Rw))) ** 2
    Ii = np.abs(np.fft.fft2(make_exits(O, P, Rw))) ** 2
    assert _rel(Ib, Ir) > 1e-3 and _rel(Ii, Ir) > 1e-3


@pytest.mark.parametrize("alg,method", [("ERA", 1), ("ERA", 4), ("DM", 1), ("DM", 4)])
def test_integer_valued_float_positions_are_the_integer_run(alg, method):
    O, P, R, I = band_limited_scan(seed=3, M=16, N=6, U=28)
    Ri = np.rint(R).astype(np.int64)
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    B = np.full(P.shape, 0.01) if method == 4 else None
    a = fn(I, Ri.copy(), P.copy(), np.ones_like(O), 4, method=method, background=B)
    b = fn(I, Ri.astype(np.float64), P.copy(), np.ones_like(O), 4, method=method, background=B)
    assert np.array_equal(a[0], b[0])
    assert a[-1]["eMod"] == b[-1]["eMod"]


@pytest.mark.parametrize("alg", ["ERA", "DM"])
def test_a_tiny_fraction_changes_the_run_by_a_tiny_amount(alg):
    O, P, R, I = band_limited_scan(seed=4, M=16, N=6, U=28)
    Ri = np.rint(R)
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    a = fn(I, Ri.astype(np.int64), P.copy(), np.ones_like(O), 4, method=1)
    b = fn(I, Ri - 1e-7, P.copy(), np.ones_like(O), 4, method=1)
    assert a[0].shape == b[0].shape
    assert 0 < _rel(b[0], a[0]) < 1e-5


@pytest.mark.parametrize("alg", ["ERA", "DM"])
def test_object_shift_data_is_a_fixed_point(alg):
    O, P, R, I = band_limited_scan()
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    _, info = fn(I, R.copy(), P.copy(), O.copy(), 5, method=1)
    assert max(info["eMod"]) < 1e-5
    _, info_r = fn(I, np.rint(R).astype(np.int64), P.copy(), O.copy(), 5, method=1)
    assert min(info_r["eMod"]) > 1e-3


def test_planted_flipped_shift_breaks_the_fixed_point(monkeypatch):
    O, P, R, I = band_limited_scan()
    right = sp.frame_probes
    monkeypatch.setattr(sp, "frame_probes", lambda P, d: right(P, -d))
    _, info = oracle.ERA(I, R.copy(), P.copy(), O.copy(), 2, method=1)
    assert info["eMod"][0] > 1e-3


def test_fractional_positions_beat_rounded_ones_from_a_flat_start():
    O, P, R, I = band_limited_scan()
    O0 = np.ones_like(O)
    _, f = oracle.ERA(I, R.copy(), P.copy(), O0.copy(), 60, method=1)
    _, r = oracle.ERA(I, np.rint(R).astype(np.int64), P.copy(), O0.copy(), 60, method=1)
    assert f["eMod"][-1] < 0.5 * r["eMod"][-1]


def test_era_sandwich_with_fractional_positions():
    """Object-only ERA alternates two projections, also with per-frame probes:
    eMod_{i+1} <= eSup_i <= eMod_i."""
    O, P, R, I = band_limited_scan(seed=6)
    _, info = oracle.ERA(I, R.copy(), P.copy(), np.ones_like(O), 8, method=1)
    e, s = np.array(info["eMod"]), np.array(info["eSup"])
    assert np.all(s <= e * (1 + 1e-9))
    assert np.all(e[1:] <= s[:-1] * (1 + 1e-9))


@pytest.mark.parametrize("kw", [dict(method=2), dict(method=3), dict(method=1, Pmod_probe=2)])
def test_probe_updates_are_rejected(kw):
    O, P, R, I = band_limited_scan(seed=7, M=16, N=4, U=28)
    with pytest.raises(ValueError):
        oracle.ERA(I, R.copy(), P.copy(), np.ones_like(O), 1, **kw)


# ---- position refinement ----------------------------------------------------

def _perturbed_state(seed=12, M=32, N=10, U=64):
    O, P, R, I = band_limited_scan(seed=seed, M=M, N=N, U=U)
    Rw, d = sp.split_positions(R)
    rng = np.random.default_rng(seed)
    d0 = d + rng.uniform(-0.2, 0.2, d.shape)
    amp = np.sqrt(np.fft.ifftshift(I, axes=(-2, -1)))
    return O, P, Rw, d, d0, amp


def test_position_terms_match_finite_differences():
    """The gradient against central differences of f_n, the Gauss-Newton
    matrix against the Jacobian of |psi_hat| by central differences."""
    O, P, Rw, d, d0, amp = _perturbed_state()
    T = sp.position_terms(O, P, Rw, d0, amp, True)
    h = 1e-5

    def mag(dd):
        ex = make_exits(O, sp.frame_probes(P, dd), Rw)
        return np.abs(np.fft.fftn(ex, axes=(-2, -1)))

    fd_g, J = [], []
    for a in range(2):
        e = np.zeros_like(d0)
        e[:, a] = h
        fp = sp.position_terms(O, P, Rw, d0 + e, amp, True)[:, 0]
        fm = sp.position_terms(O, P, Rw, d0 - e, amp, True)[:, 0]
        fd_g.append((fp - fm) / (2 * h))
        J.append((mag(d0 + e) - mag(d0 - e)) / (2 * h))
    fd_g = np.stack(fd_g, 1)
    assert np.abs(T[:, 1:3] - fd_g).max() < 1e-6 * np.abs(fd_g).max()
    H = np.stack([2 * np.sum(J[0] * J[0], axis=(1, 2)), 2 * np.sum(J[0] * J[1], axis=(1, 2)),
                  2 * np.sum(J[1] * J[1], axis=(1, 2))], 1)
    assert np.abs(T[:, 3:] - H).max() < 1e-6 * np.abs(H).max()
    # f_n is the frame's share of the unnormalised modulus error
    ex = make_exits(O, sp.frame_probes(P, d0), Rw)
    f = np.sum((np.abs(np.fft.fftn(ex, axes=(-2, -1))) - amp) ** 2, axis=(1, 2))
    assert np.allclose(T[:, 0], f, rtol=1e-12)


def test_position_terms_planted_sign_mistake():
    """PLANTED: the derivative probes of the opposite shift give a gradient
    that disagrees with the finite differences."""
    O, P, Rw, d, d0, amp = _perturbed_state(seed=13)
    right = sp.probe_derivatives
    T_ok = sp.position_terms(O, P, Rw, d0, amp, True)
    try:
        sp.probe_derivatives = lambda P, d: -right(P, d)
        T_bad = sp.position_terms(O, P, Rw, d0, amp, True)
    finally:
        sp.probe_derivatives = right
    assert np.all(np.sign(T_bad[:, 1:3]) == -np.sign(T_ok[:, 1:3]))


def test_position_step_moves_towards_the_true_shifts():
    from ptychography_amd.era import position_step
    O, P, Rw, d, d0, amp = _perturbed_state(seed=14)
    T = sp.position_terms(O, P, Rw, d0, amp, True)
    d1 = position_step(T, d0)
    f1 = sp.position_terms(O, P, Rw, d1, amp, True)[:, 0]
    assert np.all(f1 < T[:, 0])
    assert np.abs(d1 - d).max() < 0.3 * np.abs(d0 - d).max()


def refine_with_oracle(I, R, P, iters, every, method=1, background=None, O=None):
    """ERA with position refinement, on the oracle: runs of `every` iterations,
    then one position step (the facade's ``position_step``) on the oracle's
    position terms.  Returns the object, the positions and the eMod list."""
    from ptychography_amd.era import position_step
    Rs = np.asarray(R, np.float64).copy()
    Rs -= Rs.max(axis=0)
    Rw, d = sp.split_positions(Rs)
    U = P.shape[0] + int(np.ceil(np.ptp(Rs[:, 0])))
    V = P.shape[1] + int(np.ceil(np.ptp(Rs[:, 1])))
    O = np.ones((U, V), complex) if O is None else np.array(O, complex)
    amp = np.sqrt(np.fft.ifftshift(I, axes=(-2, -1)))
    e, i = [], 0
    while i < iters:
        n = min(every - i % every, iters - i)
        out = oracle.ERA(I, Rs.copy(), P.copy(), O, n, method=method, background=background,
                         windows=(Rw, d))
        O = out[0]
        e += out[-1]["eMod"]
        i += n
        if i % every == 0 and i < iters:
            d = position_step(sp.position_terms(O, P, Rw, d, amp, True), d)
    return O, Rw - d, e


def test_refinement_recovers_perturbed_positions():
    O, P, R, I = raster_subpixel_scan(seed=15)
    rng = np.random.default_rng(15)
    R0 = R + rng.uniform(-0.4, 0.4, R.shape)

    def rms(Rest):
        s, t = -Rest, -R
        err = (s - s.mean(0)) - (t - t.mean(0))
        return float(np.sqrt((err ** 2).mean()))

    _, Rr, e_r = refine_with_oracle(I, R0, P, 100, 5)
    Rn = R0 - R0.max(axis=0)
    _, e_n = oracle.ERA(I, Rn.copy(), P.copy(), np.ones_like(O), 100, method=1)
    e_n = e_n["eMod"]
    assert rms(Rr) < 0.5 * rms(Rn)
    assert e_r[-1] < 0.5 * e_n[-1]
