"""NumPy restatement of the reference ERA solver and its projection helpers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Follows
``ptychography/era.py`` of andyofmelbourne/Ptychography for one MPI rank.
Arrays stay in the caller's precision (complex128 for float64 intensities,
complex64 for float32, exactly as the reference's ``preamble`` decides).

Build additions (not in the reference, marked where they occur):

* ``info['eSup']`` -- the support/consistency error per iteration,
  ``sqrt(M*K * sum_n |psi'_n - P*O_{i+1,n}|^2 / I_norm)``: the legacy class's
  ``calc_eSup`` (``ptychography/Ptychography_2dsample_2dprobe_farfield.py:282-286``,
  unitary FFT) rescaled by Parseval to era.py's unnormalised FFT convention.
* ``callback(i, O, P)`` -- called at the end of every iteration so tests can
  compare iterates; ``None`` by default.
* ``rng`` -- where the reference draws unseeded ``np.random.random`` numbers
  (default probe, default background) this uses ``np.random`` too, unless a
  ``numpy.random.Generator``-like object with ``.random(shape)`` is given.
* ``resume`` / ``state_callback`` -- single-step checks start an iteration
  from a recorded state: ``resume`` may hold ``i0`` (the global number of the
  first iteration, which drives the OP_iters schedule and the Pmod_probe
  window), the cached ``P_heatmap`` / ``O_heatmap``, the per-frame
  ``background`` state, ``Pamp`` (the far-field probe modulus of the run's
  first probe) and, for DM, the iterate ``exits`` and the eCon reference
  ``O_ref`` / ``P_ref``.  ``state_callback(i, state)`` receives the
  same quantities (copies) after iteration i.  Without them the loop is the
  reference's.
"""

import numpy as np


# ---------------------------------------------------------------------------
# index helpers
# ---------------------------------------------------------------------------

def multiroll(x, shift):
    """Periodic roll of every axis of ``x`` by ``shift`` (era.py:631-740).

    ``y[i, j, ...] = x[(i - s0) mod n0, (j - s1) mod n1, ...]``; identical to
    nested ``np.roll`` calls, one per axis.
    """
    x = np.asarray(x)
    shift = tuple(int(s) for s in shift)
    if len(shift) != x.ndim:
        raise ValueError("multiroll needs one shift per axis")
    return np.roll(x, shift, axis=tuple(range(x.ndim)))


# How frame windows of the object are read.  "roll" rolls the whole object per
# frame, as the reference's make_exits / make_O_heatmap do: its cost model,
# which the timed CPU baseline keeps.  "slice" reads the same window values
# directly; when every window lies inside the object (R <= 0 after the
# preamble, window inside U x V) the results are bit-identical (the same
# values, multiplied and summed in the same order), which
# tests/test_oracle_pinning.py checks.  Used for the larger fixture runs.
WINDOWS = "roll"


class use_windows:
    """Context manager selecting the window mode ("roll" or "slice")."""

    def __init__(self, mode):
        if mode not in ("roll", "slice"):
            raise ValueError("mode must be 'roll' or 'slice'")
        self.mode = mode

    def __enter__(self):
        global WINDOWS
        self.saved, WINDOWS = WINDOWS, self.mode
        return self

    def __exit__(self, *exc):
        global WINDOWS
        WINDOWS = self.saved


def _inside(R, M, K, shape):
    R = np.asarray(R)
    return (R.max() <= 0 and (M - R[:, 0]).max() <= shape[0] and (K - R[:, 1]).max() <= shape[1])


def make_exits(O, P, R, exits=None):
    """Exit waves ``psi_n = roll(O, r_n)[:M, :K] * P`` (era.py:426-433).

    In the default "roll" mode this keeps the reference's cost model: every
    frame rolls the whole object.
    """
    # P may be a (N, M, K) stack of per-frame probes (sub-pixel positions,
    # oracle/subpixel.py): frame n then uses P[n]
    per_frame = P.ndim == 3
    M, K = P.shape[-2:]
    if exits is None:
        exits = np.empty((len(R), M, K), dtype=P.dtype)
    if WINDOWS == "slice" and _inside(R, M, K, O.shape):
        for n in range(len(R)):
            r0, r1 = int(R[n][0]), int(R[n][1])
            exits[n] = O[-r0:M - r0, -r1:K - r1] * (P[n] if per_frame else P)
        return exits
    for n in range(len(R)):
        exits[n] = multiroll(O, (R[n][0], R[n][1]))[:M, :K] * (P[n] if per_frame else P)
    return exits


def make_P_heatmap(P, R, shape):
    """Sum of |P|^2 placed at every frame window of the object (era.py:402-410)
    (per-frame |P_n|^2 for a stack of per-frame probes)."""
    M, K = P.shape[-2:]
    heat = np.zeros(shape, dtype=P.real.dtype)
    p2 = (P.conj() * P).real
    for n, r in enumerate(R):
        heat[-r[0]:M - r[0], -r[1]:K - r[1]] += p2[n] if P.ndim == 3 else p2
    return heat


def make_O_heatmap(O, R, shape):
    """Sum over frames of the |O|^2 window, cropped to the probe shape (era.py:412-417)."""
    o2 = (O * O.conj()).real
    M, K = shape
    if WINDOWS == "slice" and _inside(R, M, K, O.shape):
        heat = np.zeros((M, K), dtype=O.real.dtype)
        for r in R:
            r0, r1 = int(r[0]), int(r[1])
            heat += o2[-r0:M - r0, -r1:K - r1]
        return heat
    heat = np.zeros(O.shape, dtype=O.real.dtype)
    for r in R:
        heat += multiroll(o2, (r[0], r[1]))
    return heat[:shape[0], :shape[1]]


# ---------------------------------------------------------------------------
# consistency (overlap) projections
# ---------------------------------------------------------------------------

def psup_O(exits, P, R, O_shape, P_heatmap=None, alpha=1.0e-10):
    """Object update ``O = sum_n window_n^T(conj(P) psi_n) / (H + alpha)`` (era.py:336-369).

    ``H`` (the P heatmap) is recomputed only when ``P_heatmap is None``.
    """
    M, K = P.shape[-2:]
    numer = np.zeros(O_shape, P.dtype)
    if P_heatmap is None:
        P_heatmap = make_P_heatmap(P, R, O_shape)
    Pc = P.conj()
    for n, (r, ex) in enumerate(zip(R, exits)):
        numer[-r[0]:M - r[0], -r[1]:K - r[1]] += ex * (Pc[n] if P.ndim == 3 else Pc)
    return numer / (P_heatmap + alpha), P_heatmap


def psup_P(exits, O, R, O_heatmap=None, alpha=1.0e-10):
    """Probe update ``P = sum_n conj(O_n) psi_n / (sum_n |O_n|^2 + alpha)`` (era.py:371-400)."""
    M, K = exits[0].shape
    numer = np.zeros((M, K), exits.dtype)
    if O_heatmap is None:
        O_heatmap = np.ascontiguousarray(make_O_heatmap(O, R, (M, K)))
    Oc = O.conj()
    for r, ex in zip(R, exits):
        numer += ex * Oc[-r[0]:M - r[0], -r[1]:K - r[1]]
    return numer / (O_heatmap + alpha), O_heatmap


# ---------------------------------------------------------------------------
# modulus projections
# ---------------------------------------------------------------------------

def _mask_is_array(mask):
    return isinstance(mask, np.ndarray) and mask.ndim >= 1


def Pmod_1(amp, exits, mask=1, alpha=1.0e-10, eMod_calc=False):
    """Fourier-space modulus replacement, masked pixels kept (era.py:560-572).

    ``Mg = |psi_hat| + alpha``; ``eMod = sum mask (Mg - amp)^2``;
    ``psi_hat *= amp / Mg`` except where ``mask == 0`` (factor 1).
    """
    Mg = np.sqrt((exits.conj() * exits).real) + alpha
    eMod = np.sum((Mg - amp) ** 2 * mask) if eMod_calc else None
    fac = amp / Mg
    if _mask_is_array(mask):
        bad = np.where(mask == 0)
        if len(bad[0]) > 0:
            fac[:, bad[0], bad[1]] = 1.
    exits *= fac
    return exits, eMod


def pmod_1(amp, exits, mask=1, alpha=1.0e-10, eMod_calc=False):
    """FFT -> Pmod_1 -> inverse FFT over the last two axes (era.py:550-558)."""
    ex = np.fft.fftn(exits, axes=(-2, -1))
    ex, eMod = Pmod_1(amp, ex, mask, alpha, eMod_calc)
    ex = np.fft.ifftn(ex, axes=(-2, -1))
    if eMod_calc:
        return ex, eMod
    return ex


def Pmod_7(amp, background, exits, mask=1, alpha=1.0e-10, eMod_calc=False):
    """Modulus replacement with an incoherent background (era.py:583-595).

    ``Mg = sqrt(|psi_hat|^2 + b^2 + alpha)``; ``fac = mask * amp / Mg``.  The
    reference's ``if mask is 1`` test never fires once the preamble has turned
    the mask into an array, so masked pixels get factor 0 (psi_hat and b are
    zeroed there).  ``background`` is scaled in place.
    """
    Mg = np.sqrt((exits.conj() * exits).real + background ** 2 + alpha)
    eMod = np.sum((Mg - amp) ** 2 * mask) if eMod_calc else None
    fac = mask * amp / Mg
    exits *= fac
    background *= fac
    return exits, background, eMod


def pmod_7(amp, background, exits, mask=1, alpha=1.0e-10, eMod_calc=False):
    """FFT -> Pmod_7 -> inverse FFT (era.py:574-581)."""
    ex = np.fft.fftn(exits, axes=(-2, -1))
    ex, background, eMod = Pmod_7(amp, background, ex, mask, alpha, eMod_calc)
    ex = np.fft.ifftn(ex, axes=(-2, -1))
    if eMod_calc:
        return ex, background, eMod
    return ex, background


def Pmod_P(amp, P, mask=1, alpha=1.0e-10):
    """Far-field modulus constraint on the probe itself (era.py:597-610)."""
    Pf = np.fft.fftn(P, axes=(-2, -1))
    fac = amp / (np.sqrt((Pf.conj() * Pf).real) + alpha)
    if _mask_is_array(mask):
        bad = np.where(mask == 0)
        if len(bad[0]) > 0:
            fac[bad[0], bad[1]] = 1.
    Pf *= fac
    return np.fft.ifftn(Pf, axes=(-2, -1))


# ---------------------------------------------------------------------------
# probe centring (era.py:172-186)
# ---------------------------------------------------------------------------

def center_of_mass(a):
    """Intensity-weighted centroid per axis (what ``scipy.ndimage.center_of_mass`` returns)."""
    a = np.asarray(a, dtype=np.float64)
    tot = a.sum()
    rows = np.arange(a.shape[0], dtype=np.float64)
    cols = np.arange(a.shape[1], dtype=np.float64)
    return (float((a.sum(axis=1) * rows).sum() / tot),
            float((a.sum(axis=0) * cols).sum() / tot))


def probe_centre_shift(P):
    """``cm = rint(centre_of_mass(|P|^2)) - shape // 2`` (era.py:175-176; Python 2 ``/``)."""
    a = (P.conj() * P).real
    cm = np.rint(np.array(center_of_mass(a))).astype(np.int64)
    return cm - np.array(a.shape, dtype=np.int64) // 2


# ---------------------------------------------------------------------------
# preamble (era.py:435-548)
# ---------------------------------------------------------------------------

def _random(rng, shape):
    if rng is None:
        return np.random.random(shape)
    return rng.random(shape)


def preamble(I, R, P, O, iters, OP_iters, mask, background, method, alpha, dtype,
             rng=None, windows=None):
    """Resolve method/update/dtype, default O and P, amplitudes, mask and R.

    ``R`` is shifted IN PLACE so that each column's maximum is 0 (era.py:504-505).
    Returns a dict with the same quantities the reference's tuple carries.
    ``windows=(Rw, d)`` (test hook, for runs that continue a sub-pixel run
    with moved positions) gives the integer windows and the per-frame shifts
    directly, bypassing the shift and split of R.
    """
    if method is None:
        if O is None and P is None:
            method = 3
        elif O is None:
            method = 1
        elif P is None:
            method = 2
        else:
            raise ValueError("method=None needs O or P to be None (the reference "
                             "leaves 'update' undefined here)")
        if background is not None:
            method += 3
    if method in (1, 4):
        update = 'O'
    elif method in (2, 5):
        update = 'P'
    elif method in (3, 6):
        update = 'OP'
    else:
        raise ValueError("method must be one of 1..6")

    if isinstance(OP_iters, (int, np.integer)):
        OP_iters = (int(OP_iters), 1)

    if dtype is None:
        dtype = 'single' if I.dtype == np.float32 else 'double'
    if dtype == 'single':
        r_dtype, c_dtype = np.float32, np.complex64
    elif dtype == 'double':
        r_dtype, c_dtype = np.float64, np.complex128
    else:
        raise ValueError("dtype must be None, 'single' or 'double'")

    from . import subpixel
    frac = subpixel.is_fractional(R) or windows is not None
    if frac and method not in (1, 4):
        raise ValueError("sub-pixel positions are modelled for the object-only methods 1 and 4")
    if O is None:
        shape = (I.shape[1] + int(np.ceil(R[:, 0].max() - R[:, 0].min())),
                 I.shape[2] + int(np.ceil(R[:, 1].max() - R[:, 1].min())))
        O = np.ones(shape, dtype=c_dtype)
    if P is None:
        P = _random(rng, I[0].shape) + 1J * _random(rng, I[0].shape)
    P = P.astype(c_dtype)
    O = O.astype(c_dtype)

    I_norm = np.sum(mask * I)
    amp = np.fft.ifftshift(np.sqrt(I).astype(r_dtype), axes=(-2, -1))
    mask = np.asarray(mask)
    if mask.ndim > 0:
        mask = np.fft.ifftshift(mask)

    if windows is None:
        R[:, 0] -= R[:, 0].max()
        R[:, 1] -= R[:, 1].max()

    # sub-pixel positions: integer windows and per-frame probes (oracle/subpixel.py)
    Pn = None
    if windows is not None:
        R = np.asarray(windows[0], dtype=np.int64)
        Pn = subpixel.frame_probes(P, np.asarray(windows[1], dtype=np.float64))
    elif frac:
        R, d = subpixel.split_positions(R)
        Pn = subpixel.frame_probes(P, d)
    elif np.issubdtype(R.dtype, np.floating):
        R = R.astype(np.int64)      # integer-valued floats: the integer model

    exits = make_exits(O, P if Pn is None else Pn, R)

    if method in (4, 5, 6):
        if background is None:
            background = (_random(rng, exits.shape).astype(r_dtype) + 0.1).astype(r_dtype)
        else:
            tmp = np.empty(exits.shape, dtype=r_dtype)
            tmp[:] = np.sqrt(np.fft.ifftshift(background))
            background = tmp
    # outside methods 4-6 a passed background is left untouched; the reference
    # only uses it for info['background'] (era.py:311-314)

    Pamp = np.fft.fftn(P, axes=(-2, -1))
    Pamp = np.sqrt((Pamp.conj() * Pamp).real)
    return dict(method=method, update=update, dtype=r_dtype, c_dtype=c_dtype,
                OP_iters=OP_iters, O=O, P=P, amp=amp, Pamp=Pamp,
                background=background, R=R, mask=mask, I_norm=I_norm,
                N=len(I), exits=exits, Pn=Pn)


def _state(update, O, P):
    if update == 'O':
        return O
    if update == 'P':
        return P
    return np.hstack((O.ravel(), P.ravel()))


def _econ(bak, cur):
    d = bak - cur
    return np.sqrt(np.sum((d * d.conj()).real) / np.sum((cur * cur.conj()).real))


def _centre(P, O):
    cm = probe_centre_shift(P)
    return multiroll(P, -cm), multiroll(O, -cm), cm


# ---------------------------------------------------------------------------
# ERA (era.py:13-333)
# ---------------------------------------------------------------------------

def _copy(x):
    return None if x is None else np.array(x, copy=True)


def _resume(st, resume, bg_mode):
    """(i0, P_heatmap, O_heatmap, background) of a resumed run."""
    background = st['background']
    if resume is None:
        return 0, None, None, background
    if resume.get('Pamp') is not None:
        st['Pamp'] = np.asarray(resume['Pamp'], dtype=st['dtype'])
    if bg_mode and resume.get('background') is not None:
        background = np.array(resume['background'], dtype=st['dtype'])
    return (int(resume.get('i0', 0)), _copy(resume.get('P_heatmap')),
            _copy(resume.get('O_heatmap')), background)


def ERA(I, R, P, O, iters, OP_iters=1, mask=1, background=None, method=None,
        Pmod_probe=False, probe_centering=False, hardware='cpu', alpha=1.0e-10,
        dtype=None, full_output=True, callback=None, rng=None, resume=None,
        state_callback=None, windows=None):
    """Error-reduction ptychography, one-rank restatement of era.py:13-333."""
    st = preamble(I, R, P, O, iters, OP_iters, mask, background, method, alpha,
                  dtype, rng, windows)
    method, update, OP_iters = st['method'], st['update'], st['OP_iters']
    bg_mode = method in (4, 5, 6)
    i0, P_heatmap, O_heatmap, background = _resume(st, resume, bg_mode)
    O, P, amp, Pamp = st['O'], st['P'], st['amp'], st['Pamp']
    R, mask = st['R'], st['mask']
    I_norm, exits = st['I_norm'], st['exits']
    MK = P.shape[0] * P.shape[1]

    eMods, eCons, eSups = [], [], []
    # the probes the frames see (per-frame for sub-pixel positions; P never
    # changes in the object-only methods those are modelled for)
    Pn = st['Pn']
    if Pn is not None and Pmod_probe is not None and Pmod_probe is not False:
        raise ValueError("sub-pixel positions are modelled without the far-field probe constraint")

    for i in range(i0, i0 + iters):
        bak = _state(update, O, P).copy()

        # modulus projection (era.py:160 / :230)
        if bg_mode:
            exits, background, eMod = pmod_7(amp, background, exits, mask,
                                             alpha=alpha, eMod_calc=True)
        else:
            exits, eMod = pmod_1(amp, exits, mask, alpha=alpha, eMod_calc=True)

        # consistency projection (era.py:163-189 / :233-259)
        if update == 'O':
            O, P_heatmap = psup_O(exits, P if Pn is None else Pn, R, O.shape, P_heatmap, alpha)
        elif update == 'P':
            P, O_heatmap = psup_P(exits, O, R, O_heatmap, alpha)
        else:
            if i % OP_iters[1] == 0:
                for _ in range(OP_iters[0]):
                    O, P_heatmap = psup_O(exits, P, R, O.shape, None, alpha)
                    P, O_heatmap = psup_P(exits, O, R, None, alpha)
                if probe_centering:
                    P, O, _ = _centre(P, O)
                    P_heatmap = O_heatmap = None
            else:
                O, P_heatmap = psup_O(exits, P, R, O.shape, P_heatmap, alpha)

        # far-field probe modulus (era.py:193-194)
        if Pmod_probe is not None and i < Pmod_probe:
            P = Pmod_P(Pamp, P, mask, alpha)

        # frame-averaged background (era.py:265-270)
        if bg_mode:
            background[:] = np.mean(background, axis=0)

        psi_mod = exits
        exits = make_exits(O, P if Pn is None else Pn, R)

        # build addition: support error on the frames just projected
        d = psi_mod - exits
        eSups.append(float(np.sqrt(MK * np.sum((d * d.conj()).real) / I_norm)))

        eCons.append(_econ(bak, _state(update, O, P)))
        eMods.append(np.sqrt(eMod / I_norm))
        if callback is not None:
            callback(i, O, P)
        if state_callback is not None:
            state_callback(i, dict(O=O.copy(), P=P.copy(), exits=exits.copy(),
                                   background=_copy(background) if bg_mode else None,
                                   P_heatmap=_copy(P_heatmap), O_heatmap=_copy(O_heatmap)))

    if not full_output:
        if update == 'O':
            return O
        if update == 'P':
            return P
        return O, P

    info = {}
    info['I'] = np.fft.fftshift(np.abs(np.fft.fftn(exits, axes=(-2, -1))) ** 2,
                                axes=(-2, -1))
    info['eMod'] = eMods
    info['eCon'] = eCons
    info['eSup'] = eSups
    info['heatmap'] = P_heatmap
    if background is not None:
        b = background[0] if background.ndim == 3 else background
        info['background'] = np.fft.fftshift(b) ** 2
    if update == 'O':
        return O, info
    if update == 'P':
        return P, info
    return O, P, info
