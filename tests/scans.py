"""Small seeded scans shared by the tests (sizes the NumPy oracle finishes in seconds)."""

import numpy as np

from ptychography_amd import sim


def small_spec(frame=32, grid=(4, 4), step=8, jitter=1, seed=3, **kw):
    radius = frame / 4.0
    defocus = 0.02 * (64.0 / frame) ** 2
    base = dict(frame=frame, grid=grid, step=step, jitter=jitter, pupil_radius=radius,
                defocus=defocus, photons=1e6, seed=seed)
    base.update(kw)
    return sim.ScanSpec(**base)


def small_scan(**kw):
    return sim.make_scan(small_spec(**kw))


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def flat_start(scan, rng_seed=11):
    """A fixed random probe start and the reference's default flat object."""
    rng = np.random.default_rng(rng_seed)
    M = scan.I.shape[1]
    P0 = np.fft.fftshift(np.fft.ifftn(np.abs(np.fft.fftn(scan.P)) *
                                      np.exp(2j * np.pi * rng.random((M, M)))))
    return P0


def broadband_start(shape, seed=13, eps=0.1):
    """1 + eps * white complex noise: an initial object whose exit-wave spectra
    have no numerically-zero pixels.

    From a flat object and a band-limited probe the model spectrum outside the
    pupil is pure round-off, and the modulus step amp / (|psi_hat| + 1e-10)
    turns that round-off into O(1) values with arbitrary phase -- two correct
    FFTs (even two float64 ones) then disagree there.  Parity runs start from
    this object instead.
    """
    rng = np.random.default_rng(seed)
    return 1.0 + eps * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
