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


def _freq_index(n):
    """FFT bin frequencies in [-(n/2 - 1), n/2] (bagOfns.make_xy(origin=(0, 0)))."""
    j = np.arange(n)
    return ((j + n // 2 - 1) % n) - (n // 2 - 1)


def _smooth(shape, band, rng):
    F = np.zeros(shape, complex)
    ky = _freq_index(shape[0])[:, None]
    kx = _freq_index(shape[1])[None, :]
    sel = (np.abs(ky) <= band) & (np.abs(kx) <= band)
    F[sel] = rng.standard_normal(sel.sum()) + 1j * rng.standard_normal(sel.sum())
    return np.fft.ifft2(F) * np.sqrt(F.size)


def band_limited_scan(seed=0, M=32, N=12, U=56):
    """Object band-limited to |k| <= 4 of its grid, a Gaussian probe with a
    weak quadratic phase, fractional positions already shifted so their
    maxima are 0 (the preamble leaves them unchanged), and the centred
    intensities of the reference's object-shift exits (the oracle's
    restatement of utility_Ptych.makeExits, the data-making checker)."""
    from oracle import subpixel
    rng = np.random.default_rng(seed)
    O = 1 + 0.3 * _smooth((U, U), 4, rng)
    y, x = np.mgrid[:M, :M]
    P = np.exp(-((y - M / 2 + 0.3) ** 2 + (x - M / 2 - 0.7) ** 2) / (2 * 3.0 ** 2)) * \
        np.exp(1j * 0.05 * (x - M / 2) ** 2)
    R = -rng.uniform(0, U - M, size=(N, 2))
    R -= R.max(axis=0)
    ex = subpixel.make_exits_object_shift(O, P, R)
    I = np.fft.fftshift(np.abs(np.fft.fft2(ex)) ** 2, axes=(-2, -1))
    return O, P, R, I


def raster_subpixel_scan(seed=0, M=32, G=6, step=9.0, band=6, sigma=5.0):
    """G x G raster of fractional positions (step plus a uniform [0, 1)
    jitter), an object band-limited to |k| <= band, a Gaussian probe of width
    sigma with a weak quadratic phase, and the centred intensities of the
    probe-shift model (the oracle's make_exits with per-frame probes).
    Returns (O, P, R, I) with R already shifted so its maxima are 0."""
    from oracle import make_exits, subpixel
    rng = np.random.default_rng(seed)
    N = G * G
    yy, xx = np.mgrid[:G, :G]
    R = -np.stack([yy.ravel() * step, xx.ravel() * step], 1) - rng.uniform(0, 1, (N, 2))
    R -= R.max(axis=0)
    U = M + int(np.ceil(np.ptp(R[:, 0])))
    V = M + int(np.ceil(np.ptp(R[:, 1])))
    O = 1 + 0.4 * _smooth((U, V), band, rng)
    y, x = np.mgrid[:M, :M]
    P = (np.exp(-((y - M / 2) ** 2 + (x - M / 2) ** 2) / (2 * sigma ** 2)) *
         np.exp(1j * 0.03 * ((x - M / 2) ** 2 + (y - M / 2) ** 2)))
    Rw, d = subpixel.split_positions(R)
    ex = make_exits(O, subpixel.frame_probes(P, d), Rw)
    I = np.fft.fftshift(np.abs(np.fft.fft2(ex)) ** 2, axes=(-2, -1))
    return O, P, R, I
