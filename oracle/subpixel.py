"""Sub-pixel positions for the oracle.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference shifts arrays by non-integer amounts with the Fourier shift
theorem (``bagOfns.roll``, ptychography/bagOfns.py:729-764, its ramp from
``phase_ramp`` / ``make_xy(origin=(0, 0))``, :966-979 and :471-492): the
spectrum is multiplied by exp(-2 pi i (s0 ky / N0 + s1 kx / N1)) with the
frequency index k of bin j taken as ((j + N//2 - 1) mod N) - (N//2 - 1), i.e.
in [-(N/2 - 1), N/2].  Its ERA / DM take integer positions only; the model
used for fractional ones is a probe that moves by sub-pixel amounts over an
object sampled on its own grid:

* a window start s_n = -R_n (after the preamble's shift, R <= 0) splits into
  the integer window w_n = rint(s_n) and the fraction d_n = s_n - w_n in
  [-1/2, 1/2];
* frame n sees P_n = roll(P, d_n) times the object window at w_n, i.e. the
  probe value at frame pixel r is P(r - d_n);
* the overlap projection onto the object uses P_n in place of P (numerator
  conj(P_n) psi_n, heatmap sum |P_n|^2).

For integer positions d_n = 0 and everything is the integer model bit for bit.

The reference's own float-position exits (``utility_Ptych.makeExits``,
ptychography/utility_Ptych.py:42-48) shift the whole object instead:
exit_n = roll(O, R_n)[:M, :K] * P.  For an object and a probe band-limited
well inside the frame's Nyquist band the two give the same diffraction
intensities (the exit waves differ by a sub-pixel translation of the frame
grid); ``make_exits_object_shift`` restates that form so tests can compare.
"""

import numpy as np


def freq_index(n):
    """Frequency index of each FFT bin, as the reference's make_xy(origin=(0, 0)) numbers them."""
    j = np.arange(n)
    return ((j + n // 2 - 1) % n) - (n // 2 - 1)


def roll(x, shift):
    """bagOfns.roll for a 2-D array and a (possibly fractional) shift (s0, s1):
    y[i, j] = x[i - s0, j - s1], through the Fourier shift theorem."""
    x = np.asarray(x)
    N0, N1 = x.shape[-2:]
    ky = freq_index(N0)[:, None]
    kx = freq_index(N1)[None, :]
    ramp = np.exp(-2.0j * np.pi * (shift[0] * ky / N0 + shift[1] * kx / N1))
    xh = np.fft.fftn(x, axes=(-2, -1))
    # complex64 input: single-precision transforms, the ramp formed in double
    return np.fft.ifftn(xh * ramp.astype(xh.dtype), axes=(-2, -1))


def is_fractional(R):
    R = np.asarray(R)
    return np.issubdtype(R.dtype, np.floating) and np.any(R != np.rint(R))


def split_positions(R):
    """(integer R for the windows, fractions d) of positions R <= 0 (after the
    preamble shift): w = rint(-R), d = -R - w, returned as R_int = -w."""
    s = -np.asarray(R, dtype=np.float64)
    w = np.rint(s)
    return (-w).astype(np.int64), s - w


def frame_probes(P, d):
    """(N, M, K) stack of P_n = roll(P, d_n)."""
    return np.stack([roll(P, dn) for dn in d]).astype(P.dtype)


def make_exits_object_shift(O, P, R):
    """utility_Ptych.makeExits (utility_Ptych.py:42-48): exit_n =
    bagOfns.roll(O, R_n)[:M, :K] * P, R_n float (after the preamble's shift)."""
    M, K = P.shape
    return np.stack([roll(O, (float(r[0]), float(r[1])))[:M, :K] * P for r in R])


def probe_derivatives(P, d):
    """(N, 2, M, K): d P_n / d d_n along y and x, P_n = roll(P, d_n), i.e.
    IFFT(FFT(P) ramp_n (-2 pi i f / N)) on each axis (the derivative of
    bagOfns.roll's phase ramp, bagOfns.py:729-764)."""
    M, K = P.shape
    ky = freq_index(M)[:, None]
    kx = freq_index(K)[None, :]
    Ph = np.fft.fftn(P)
    out = np.empty((len(d), 2) + P.shape, Ph.dtype)
    for n, (dy, dx) in enumerate(d):
        # the ramp and its derivative factors in double, rounded to P's precision
        ramp = Ph * np.exp(-2.0j * np.pi * (dy * ky / M + dx * kx / K))
        out[n, 0] = np.fft.ifftn((ramp * (-2.0j * np.pi * ky / M)).astype(Ph.dtype))
        out[n, 1] = np.fft.ifftn((ramp * (-2.0j * np.pi * kx / K)).astype(Ph.dtype))
    return out


def position_terms(O, P, Rw, d, amp, mask):
    """Per frame n, with psi_hat = FFT(O_win(Rw_n) P_n), r = |psi_hat| - amp
    over the good pixels (amp and mask ifftshifted, as the preamble holds
    them) and J_a = Re(conj(psi_hat) dpsi_hat / d d_a) / |psi_hat|:
    (N, 6) = f_n = sum r^2, the gradient 2 sum r J_y, 2 sum r J_x and the
    Gauss-Newton matrix 2 sum J_y J_y, 2 sum J_y J_x, 2 sum J_x J_x.  The
    modulus error the reference minimises (era.py:560-572) as a function of
    the frame positions, the coordinate-gradient of its notebooks
    (ipython_notbooks/1dcoordsUpdate_1dsample_2dprobe.ipynb) for the
    probe-shift model."""
    M, K = P.shape
    # complex64 inputs keep the transforms in single precision (the probe
    # stacks are formed in double and rounded, as on the GPU)
    cdt = np.result_type(O.dtype, P.dtype)
    Pn = frame_probes(P, d).astype(cdt)
    dP = probe_derivatives(P, d).astype(cdt)
    mask = np.broadcast_to(np.asarray(mask, bool), (M, K))
    out = np.zeros((len(d), 6))
    for n in range(len(d)):
        r0, r1 = int(Rw[n][0]), int(Rw[n][1])
        Ow = O[-r0:M - r0, -r1:K - r1]
        ps = np.fft.fftn(Ow * Pn[n]).astype(np.complex128)
        a = np.abs(ps)
        r = np.where(mask, a - amp[n], 0.0)
        J = [np.where(mask, (ps.conj() * np.fft.fftn(Ow * dP[n, k])).real / np.maximum(a, 1e-300), 0.0)
             for k in range(2)]
        out[n] = [np.sum(r * r), 2 * np.sum(r * J[0]), 2 * np.sum(r * J[1]),
                  2 * np.sum(J[0] * J[0]), 2 * np.sum(J[0] * J[1]), 2 * np.sum(J[1] * J[1])]
    return out
