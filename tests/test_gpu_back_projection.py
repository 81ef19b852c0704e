"""Back-projection initial object on the GPU against the oracle.

The back projection takes the phase of the probe's far field, fftn(P).  Where
that far field vanishes (outside a hard pupil) its phase is round-off and
differs between any two FFTs (float32 here, float64 in the oracle), while the
data there are not zero (the object scatters outside the pupil).  The parity
scans therefore use a Gaussian-apodised pupil whose far field stays well above
round-off; the hard-pupil case is checked inside the pupil only.
"""

import numpy as np
import pytest

import oracle.back_projection as obp
import ptychography_amd as pty
from ptychography_amd import sim
from scans import rel_l2, small_scan

pytestmark = pytest.mark.gpu


def _soft_probe(frame, sigma_frac=0.3, defocus=0.02):
    i = np.fft.fftfreq(frame, 1.0 / frame)
    q2 = np.fft.fftshift(i[:, None] ** 2 + i[None, :] ** 2)
    pupil = np.exp(-q2 / (2 * (sigma_frac * frame) ** 2) - 1j * np.pi * defocus * (64.0 / frame) ** 2 * q2)
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(pupil)))


@pytest.mark.parametrize("frame,masked", [(32, False), (64, True), (128, True)])
def test_back_projection_matches_the_oracle(frame, masked, gpu_device):
    s = small_scan(frame=frame, grid=(4, 4), step=frame // 4, seed=frame,
                   beamstop_radius=(frame / 16 if masked else 0.0), dead_frac=(0.05 if masked else 0.0))
    P = _soft_probe(frame)
    I = sim.frames_intensity(s.O, P, -s.R)
    I *= 1e6 / I.sum(axis=(1, 2)).mean()
    Rg, Rr = s.R.copy(), s.R.copy()
    Og, ig = pty.Back_projection(I, Rg, P, None, mask=s.mask, device=gpu_device)
    Or, ir = obp.Back_projection(I, Rr, P, None, mask=s.mask)
    assert np.array_equal(Rg, Rr)
    H = ir["heatmap"]
    cov = H > 1e-3 * H.max()
    assert rel_l2(ig["heatmap"], H) < 1e-6
    assert rel_l2(ig["exits"], ir["exits"]) < 1e-5
    assert rel_l2(Og[cov], Or[cov]) < 1e-5


def test_hard_pupil_agrees_inside_the_pupil(gpu_device):
    s = small_scan(frame=64, grid=(4, 4), step=16, seed=9)
    Og, ig = pty.Back_projection(s.I, s.R.copy(), s.P, None, device=gpu_device)
    Or, ir = obp.Back_projection(s.I, s.R.copy(), s.P, None)
    Fg = np.fft.fftn(ig["exits"], axes=(-2, -1))
    Fr = np.fft.fftn(ir["exits"], axes=(-2, -1))
    PF = np.abs(np.fft.fftn(s.P))
    inside = PF > 1e-3 * PF.max()
    assert rel_l2(Fg[:, inside], Fr[:, inside]) < 1e-5
