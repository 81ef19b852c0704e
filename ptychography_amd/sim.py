"""Seeded synthetic ptychography scans.

Plays the role of the reference's ``forward_sim``
(``ptychography/forward_models.py:12-203``) and keeps its conventions:

* the probe is built as a far-field pupil (disc of radius ``A`` pixels) times a
  phase, then ``P = fftshift(ifftn(ifftshift(pupil)))`` (forward_models.py:125-138);
* ``R`` holds the NEGATED window offsets, so frame ``n`` sees
  ``O[-R[n,0] : M - R[n,0], -R[n,1] : K - R[n,1]]`` and ``R.max(axis=0) == 0``
  (forward_models.py:149-163);
* ``I_n = ifftshift(|fftn(P * O_n)|^2)`` -- DC at the frame centre
  (forward_models.py:167-171);
* ``mask`` is ``True`` for good pixels (forward_models.py:182-188).

Unlike the reference (unseeded ``np.random`` and the brog/twain images) every
draw here comes from ``np.random.default_rng(seed)`` and objects are smoothed
random amplitude/phase maps, following the named workload recipes below.
"""

from dataclasses import dataclass, field, replace
from math import factorial

import numpy as np


@dataclass(frozen=True)
class ScanSpec:
    frame: int = 64                 # detector M = K
    grid: tuple = (12, 12)          # raster positions (rows, cols)
    step: int = 16                  # raster step in object pixels
    jitter: int = 1                 # integer jitter, uniform in [-jitter, jitter]
    pupil_radius: float = 16.0      # far-field pupil radius (pixels)
    defocus: float = 0.0            # quadratic phase exp(-i pi defocus q^2)
    zernike_rms: float = 0.0        # random Zernike phase (orders <= 4), rad per term
    photons: float = 1e6            # mean photons per frame before noise
    poisson: bool = True
    amp_range: tuple = (0.5, 1.0)   # object amplitude range
    sigma: float = 4.0              # smoothing length of the object maps (pixels)
    beamstop_radius: float = 0.0    # masked disc at the detector centre (pixels)
    dead_frac: float = 0.0          # fraction of random dead pixels (masked)
    background_frac: float = 0.0    # smooth additive background, fraction of mean I
    obj_size: int = 0               # object side; 0 = smallest that holds every window
    seed: int = 0

    def with_(self, **kw):
        return replace(self, **kw)


# Named recipes.  Sizes follow the workload descriptions; defocus values are
# chosen so the probe footprint overlaps neighbouring positions by ~50-70 %.
RECIPES = {
    # 144 x 64^2 frames, object 256^2 (ERA method 1 on the CPU path)
    "c1": ScanSpec(frame=64, grid=(12, 12), step=16, jitter=1, pupil_radius=16,
                   defocus=0.02, photons=1e6, obj_size=256, seed=0),
    # 1024 x 128^2 frames, object ~1124^2, Zernike probe, 10 % masked
    "c2": ScanSpec(frame=128, grid=(32, 32), step=32, jitter=2, pupil_radius=32,
                   defocus=0.012, zernike_rms=0.6, photons=1e7, beamstop_radius=8.0,
                   dead_frac=0.086, seed=1),
    # 4096 x 256^2 frames, object ~2782^2, background 1 % of mean I, 8 px beamstop
    "c3": ScanSpec(frame=256, grid=(64, 64), step=40, jitter=3, pupil_radius=64,
                   defocus=0.0035, photons=1e8, beamstop_radius=4.0,
                   background_frac=0.01, seed=2),
    # held-out learned-init scan: 32 x 32 positions of 64^2 frames
    "c5": ScanSpec(frame=64, grid=(32, 32), step=16, jitter=1, pupil_radius=16,
                   defocus=0.02, photons=1e6, amp_range=(0.3, 1.0), seed=6),
}


def _gauss_smooth(a, sigma):
    """Periodic Gaussian smoothing through the FFT (float64)."""
    ny, nx = a.shape
    qy = np.fft.fftfreq(ny)[:, None]
    qx = np.fft.fftfreq(nx)[None, :]
    g = np.exp(-2.0 * (np.pi * sigma) ** 2 * (qy ** 2 + qx ** 2))
    return np.fft.ifft2(np.fft.fft2(a) * g).real


def _rescale(a, lo, hi):
    amin, amax = a.min(), a.max()
    if amax == amin:
        return np.full_like(a, lo)
    return (a - amin) * (hi - lo) / (amax - amin) + lo


def make_object(shape, rng, amp_range=(0.5, 1.0), sigma=4.0):
    """Smoothed random amplitude in ``amp_range`` times a smoothed random phase in [-pi, pi]."""
    amp = _rescale(_gauss_smooth(rng.random(shape), sigma), 0.0, 1.0)
    amp = amp_range[0] + (amp_range[1] - amp_range[0]) * amp
    phase = _rescale(_gauss_smooth(rng.random(shape), sigma), -np.pi, np.pi)
    return (amp * np.exp(1j * phase)).astype(np.complex128)


def _zernike_radial(n, m, rho):
    out = np.zeros_like(rho)
    for k in range((n - m) // 2 + 1):
        c = ((-1) ** k * factorial(n - k)
             / (factorial(k) * factorial((n + m) // 2 - k) * factorial((n - m) // 2 - k)))
        out += c * rho ** (n - 2 * k)
    return out


def zernike_phase(q_rho, q_theta, rng, rms, max_order=4):
    """Random sum of Zernike polynomials Z_n^m, 1 <= n <= max_order, on the unit disc."""
    ph = np.zeros_like(q_rho)
    for n in range(1, max_order + 1):
        for m in range(-n, n + 1, 2):
            rad = _zernike_radial(n, abs(m), q_rho)
            ang = np.cos(m * q_theta) if m >= 0 else np.sin(-m * q_theta)
            ph += rng.normal(0.0, rms) * rad * ang
    return ph


def make_probe(frame, radius, defocus=0.0, zernike_rms=0.0, rng=None):
    """Real-space probe from a far-field disc pupil (forward_models.py:125-138)."""
    i = np.fft.fftfreq(frame, 1.0 / frame)
    qi, qj = np.meshgrid(i, i, indexing="ij")
    q2 = np.fft.fftshift(qi ** 2 + qj ** 2)
    pupil = (q2 < radius ** 2).astype(np.complex128)
    phase = -np.pi * defocus * q2
    if zernike_rms > 0.0:
        qs_i = np.fft.fftshift(qi)
        qs_j = np.fft.fftshift(qj)
        rho = np.sqrt(q2) / radius
        theta = np.arctan2(qs_i, qs_j)
        phase = phase + np.where(rho < 1.0, zernike_phase(rho, theta, rng, zernike_rms), 0.0)
    pupil *= np.exp(1j * phase)
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(pupil)))


def make_positions(grid, step, jitter, rng):
    """Jittered raster window offsets (>= 0) and the object size that holds them."""
    ny, nx = grid
    oy, ox = np.meshgrid(np.arange(ny) * step, np.arange(nx) * step, indexing="ij")
    off = np.stack([oy.ravel(), ox.ravel()], axis=1).astype(np.int64) + jitter
    if jitter > 0:
        off += rng.integers(-jitter, jitter + 1, size=off.shape)
    return off - off.min(axis=0)


def make_mask(frame, beamstop_radius, dead_frac, rng):
    """Good-pixel mask, centred (DC at frame//2): beamstop disc plus random dead pixels."""
    c = frame // 2
    yy, xx = np.meshgrid(np.arange(frame) - c, np.arange(frame) - c, indexing="ij")
    mask = np.ones((frame, frame), dtype=bool)
    if beamstop_radius > 0:
        mask &= (yy ** 2 + xx ** 2) >= beamstop_radius ** 2
    if dead_frac > 0:
        mask &= rng.random((frame, frame)) >= dead_frac
    return mask


def frames_intensity(O, P, offsets, chunk=256):
    """Noise-free centred intensities ``ifftshift(|fftn(P * O_n)|^2)``, float64."""
    M, K = P.shape
    out = np.empty((len(offsets), M, K), dtype=np.float64)
    for s in range(0, len(offsets), chunk):
        part = offsets[s:s + chunk]
        ex = np.empty((len(part), M, K), dtype=np.complex128)
        for k, (y, x) in enumerate(part):
            ex[k] = O[y:y + M, x:x + K] * P
        f = np.fft.fftn(ex, axes=(-2, -1))
        out[s:s + len(part)] = np.fft.ifftshift((f.conj() * f).real, axes=(-2, -1))
    return out


@dataclass
class Scan:
    I: np.ndarray          # (N, M, K) centred intensities
    R: np.ndarray          # (N, 2) int64, negated window offsets, max 0
    mask: object           # (M, K) bool, True = good, or 1
    P: np.ndarray          # (M, K) complex128 true probe
    O: np.ndarray          # (U, V) complex128 true object
    B: object              # (M, K) centred background or None
    spec: ScanSpec = field(default=None)

    @property
    def offsets(self):
        return -self.R


def make_scan(spec="c1", noise=None, **overrides):
    """Build a scan from a named recipe (or a ``ScanSpec``), with overrides."""
    if isinstance(spec, str):
        spec = RECIPES[spec]
    if overrides:
        spec = spec.with_(**overrides)
    if noise is not None:
        spec = spec.with_(poisson=bool(noise))
    rng = np.random.default_rng(spec.seed)
    M = spec.frame
    offsets = make_positions(spec.grid, spec.step, spec.jitter, rng)
    U = max(int(offsets[:, 0].max()) + M, spec.obj_size)
    V = max(int(offsets[:, 1].max()) + M, spec.obj_size)
    O = make_object((U, V), rng, spec.amp_range, spec.sigma)
    P = make_probe(M, spec.pupil_radius, spec.defocus, spec.zernike_rms, rng)
    I = frames_intensity(O, P, offsets)
    # scale so the mean frame holds `photons` counts
    scale = spec.photons / max(I.sum() / len(I), 1e-300)
    I *= scale
    P = P * np.sqrt(scale)
    B = None
    if spec.background_frac > 0:
        B = _rescale(_gauss_smooth(rng.random((M, M)), 2.0 * spec.sigma), 0.5, 1.5)
        B *= spec.background_frac * I.mean() / B.mean()
        I += B[None]
    if spec.poisson:
        I = rng.poisson(I).astype(np.float64)
    mask = make_mask(M, spec.beamstop_radius, spec.dead_frac, rng)
    if mask.all():
        mask = 1
    R = -offsets
    R -= R.max(axis=0)
    return Scan(I=I, R=R, mask=mask, P=P, O=O, B=B, spec=spec)


def forward_sim(shape_P=(32, 64), shape_O=(128, 128), A=14, defocus=0., photons_pupil=1, ny=10,
                nx=10, random_offset=None, background=None, mask=None, counting_statistics=False,
                rng=None):
    """The reference's ``forward_sim`` (ptychography/forward_models.py:12-203):
    same arguments and returns ``(I, R, M, P, O, B)``, the same steps, with
    every random draw taken from ``rng`` (a ``numpy.random.Generator``; default
    seed 0) and a smoothed random object in place of the reference's two
    bundled images.

    * P: far-field disc of radius A times sqrt(photons_pupil) and the defocus
      phase exp(-i pi defocus q^2), brought to the sample plane
      (forward_models.py:125-138);
    * R: an ny x nx grid from rint(linspace(0, M - U, n)), plus integer
      offsets floor(random * random_offset), clipped to the object, shifted
      so each column's maximum is 0 (:149-163);
    * I = ifftshift(|fftn(O_n P)|^2) (:165-171); B = rint(random * background)
      added to every frame (:173-178); M: `mask` random dead pixels
      (False = bad) (:180-186); Poisson counts if counting_statistics (:189-190).
    """
    rng = np.random.default_rng(0) if rng is None else rng
    Mf, Kf = shape_P
    i = np.fft.fftfreq(Mf, 1.0 / Mf)
    j = np.fft.fftfreq(Kf, 1.0 / Kf)
    qi, qj = np.meshgrid(i, j, indexing="ij")
    q2 = np.fft.fftshift(qi ** 2 + qj ** 2)
    P = (q2 < A ** 2).astype(np.complex128) * np.sqrt(photons_pupil)
    P *= np.exp(-1.0j * np.pi * defocus * q2)
    P = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(P)))

    O = make_object(shape_O, rng, amp_range=(0.0, 1.0))

    ys = np.rint(np.linspace(0, Mf - shape_O[0], ny)).astype(np.int64)
    xs = np.rint(np.linspace(0, Kf - shape_O[1], nx)).astype(np.int64)
    ys, xs = np.meshgrid(ys, xs, indexing="ij")
    R = np.stack([ys.ravel(), xs.ravel()], axis=1)
    if random_offset is not None:
        R += (rng.random(R.shape) * int(random_offset)).astype(np.int64)
        R[R > 0] = 0
        R[:, 0] = np.maximum(R[:, 0], Mf - shape_O[0])
        R[:, 1] = np.maximum(R[:, 1], Kf - shape_O[1])
    R[:, 0] -= R[:, 0].max()
    R[:, 1] -= R[:, 1].max()

    I = frames_intensity(O, P, -R)
    if background is not None:
        B = np.rint(rng.random(I[0].shape) * background)
        I += B
    else:
        B = 0
    if mask is not None:
        ss = (rng.random(mask) * I.shape[1]).astype(np.int64)
        fs = (rng.random(mask) * I.shape[2]).astype(np.int64)
        Mask = np.ones(I[0].shape, dtype=bool)
        Mask[ss, fs] = False
    else:
        Mask = 1
    if counting_statistics:
        I = rng.poisson(I).astype(np.float64)
    return I, R, Mask, P, O, B


def c4_specs(count=16, seed0=100, frames=(64, 96, 128, 192, 256), positions=(100, 2000)):
    """The mixed multi-scan workload: scan s (seed seed0 + s) draws its frame
    side from ``frames`` and its position count from ``positions``, a raster
    step giving 30-50 % overlap of the probe footprint (half the frame), and
    follows the 128^2 Zernike-probe recipe scaled to its frame size (pupil
    radius frame/4, same maximum defocus phase, 10^7 photons, beamstop and
    8.6 % dead pixels)."""
    out = []
    for s in range(count):
        rng = np.random.default_rng(seed0 + s)
        frame = int(rng.choice(frames))
        n = int(rng.integers(positions[0], positions[1] + 1))
        gy = max(1, int(round(np.sqrt(n))))
        gx = max(1, int(round(n / gy)))
        overlap = rng.uniform(0.3, 0.5)
        f = frame / 128.0
        out.append(ScanSpec(frame=frame, grid=(gy, gx), step=max(1, int(round(frame / 2 * (1 - overlap)))),
                            jitter=max(1, int(round(2 * f))), pupil_radius=32.0 * f,
                            defocus=0.012 / f ** 2, zernike_rms=0.6, photons=1e7,
                            beamstop_radius=8.0 * f, dead_frac=0.086, seed=seed0 + s))
    return out


def make_scan_gpu(spec, device, poisson=True):
    """``make_scan`` for large scans: the same draws, with the intensities
    formed by the engine's own frame kernels on ``device`` (complex64) and
    returned as a float32 device tensor.  Returns a dict with I, R (int64,
    max 0), P (scaled like ``make_scan``), O, B, mask, spec."""
    import torch
    from .engine import Engine

    rng = np.random.default_rng(spec.seed)
    M = spec.frame
    off = make_positions(spec.grid, spec.step, spec.jitter, rng)
    U = max(int(off[:, 0].max()) + M, spec.obj_size)
    V = max(int(off[:, 1].max()) + M, spec.obj_size)
    O = make_object((U, V), rng, spec.amp_range, spec.sigma)
    P = make_probe(M, spec.pupil_radius, spec.defocus, spec.zernike_rms, rng)
    R = -off
    R -= R.max(axis=0)
    N = len(R)
    gen = Engine(N, M, M, U, V, R, device=device)
    try:
        O_t = torch.from_numpy(O.astype(np.complex64)).to(device)
        P_t = torch.from_numpy(P.astype(np.complex64)).to(device)
        gen.set_state(O_t, P_t)
        I = gen.intensity()
    finally:
        gen.close()
    scale = spec.photons / max(float(I.sum(dtype=torch.float64).item()) / N, 1e-300)
    I *= scale
    P = P * np.sqrt(scale)
    B = None
    if spec.background_frac > 0:
        B = _rescale(_gauss_smooth(rng.random((M, M)), 2.0 * spec.sigma), 0.5, 1.5)
        B *= spec.background_frac * float(I.mean().item()) / B.mean()
        I += torch.from_numpy(B.astype(np.float32)).to(device)[None]
    if poisson:
        g = torch.Generator(device=device).manual_seed(spec.seed)
        I = torch.poisson(I, generator=g)
    mask = make_mask(M, spec.beamstop_radius, spec.dead_frac, rng)
    if mask.all():
        mask = 1
    return dict(I=I, R=R, P=P, O=O, B=B, mask=mask, spec=spec)
