"""Pins the back-projection oracle with closed forms (no reference vectors exist:
the reference's back_projection.py does not run as written)."""

import numpy as np

from oracle.back_projection import Back_projection
from scans import small_scan


def test_flat_object_exact_data_gives_the_probe_back():
    """I_n = |F(P * 1)|^2 centred: modulus and phase are both P's, so every exit
    wave is P and the object is sum|P|^2 / (sum|P|^2 + alpha) (1 where covered)."""
    s = small_scan(frame=32, grid=(4, 4), step=8)
    P = s.P
    I1 = np.fft.ifftshift(np.abs(np.fft.fftn(P)) ** 2)
    I = np.broadcast_to(I1, (len(s.R),) + I1.shape).copy()
    R = s.R.copy()
    O, info = Back_projection(I, R, P, None, alpha=1e-10)
    assert np.allclose(info["exits"], P[None], atol=1e-12 * np.abs(P).max())
    H = info["heatmap"]
    cov = H > 1e-6 * H.max()
    assert np.allclose(O[cov], (H / (H + 1e-10))[cov], rtol=1e-10)
    assert R.max(axis=0).tolist() == [0, 0]


def test_exit_spectra_carry_the_masked_modulus_and_the_probe_phase():
    s = small_scan(frame=32, grid=(3, 3), step=8, beamstop_radius=3.0, dead_frac=0.1, seed=7)
    O, info = Back_projection(s.I, s.R.copy(), s.P, None, mask=s.mask)
    F = np.fft.fftn(info["exits"], axes=(-2, -1))
    m = np.fft.ifftshift(s.mask)
    amp = np.fft.ifftshift(np.sqrt(s.I), axes=(-2, -1))
    assert np.allclose(np.abs(F), m * amp, atol=1e-9 * amp.max())
    ph = np.exp(1j * np.angle(np.fft.fftn(s.P)))
    on = (m * amp) > 1e-6 * amp.max()
    unit = F / np.where(on, np.abs(F), 1)
    assert np.allclose(unit[on], np.broadcast_to(ph, F.shape)[on], atol=1e-8)
