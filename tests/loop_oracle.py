"""An independent pure-loop restatement of the reference ERA / DM iteration.

TEST INFRASTRUCTURE ONLY: the second checker that pins ``oracle.ERA`` /
``oracle.DM`` (the NumPy restatement that every GPU parity test compares
against).  It is written again from the reference text, not from ``oracle/``,
and shares no code with it:

* every array access is an explicit per-pixel loop with modular index
  arithmetic: the periodic roll of ``multiroll`` (ptychography/era.py:707-740)
  is ``O[(i - r0) % U][(j - r1) % V]``, a frame window is the index map
  ``(i, j) -> (i - r0, j - r1)``;
* the 2-D transforms are direct O(n^2) sums over a twiddle table
  (``numpy.fft``'s sign and scaling: forward unnormalised, inverse 1/(M K)),
  separable over rows and columns -- no FFT;
* ``fftshift`` / ``ifftshift`` are index maps, the centre of mass is an
  explicit weighted sum, ``rint`` is Python's half-to-even ``round``.

It follows ptychography/era.py:13-333 (loop :152-291, helpers :336-433,
preamble :435-548, modulus steps :550-610) and ptychography/dm.py:13-333,
:337-341, for one MPI rank (every collective is the identity).  It also
computes the two build additions the oracle reports, with the same
definitions: ERA's eSup_i = sqrt(M K sum |psi'_n - P O_{i+1,n}|^2 / I_norm),
DM's eSup_i = sqrt(M K sum |f_{i+1} - f_in|^2 / I_norm), f_in being the
iterate the DM update starts from.

It is slow (pure Python) and meant for frames of at most ~8 x 8 pixels and a
handful of positions.
"""

import cmath
import math

import numpy as np


# ---------------------------------------------------------------------------
# index maps and transforms
# ---------------------------------------------------------------------------

def _twiddles(n, sign):
    return [cmath.exp(sign * 2j * math.pi * k / n) for k in range(n)]


def _dft_lines(rows, n, sign):
    """Direct DFT of every line in ``rows`` (a list of length-n lists)."""
    w = _twiddles(n, sign)
    out = []
    for line in rows:
        res = []
        for k in range(n):
            acc = 0j
            for t in range(n):
                acc += line[t] * w[(k * t) % n]
            res.append(acc)
        out.append(res)
    return out


def dft2(frame, inverse=False):
    """2-D DFT of an M x K frame (list of lists or array), direct sums.

    Forward: X[k, l] = sum_{i, j} x[i, j] exp(-2 pi i (k i / M + l j / K)).
    Inverse: the conjugate kernel scaled by 1 / (M K) (numpy's ifftn).
    """
    M = len(frame)
    K = len(frame[0])
    sign = 1 if inverse else -1
    rows = _dft_lines([[complex(frame[i][j]) for j in range(K)] for i in range(M)], K, sign)
    cols = _dft_lines([[rows[i][j] for i in range(M)] for j in range(K)], M, sign)
    out = [[cols[j][i] for j in range(K)] for i in range(M)]
    if inverse:
        s = 1.0 / (M * K)
        out = [[v * s for v in row] for row in out]
    return out


def ifftshift2(a):
    """out[i, j] = a[(i + M//2) % M, (j + K//2) % K] over the last two axes."""
    a = np.asarray(a)
    M, K = a.shape[-2:]
    out = np.empty_like(a)
    for i in range(M):
        for j in range(K):
            out[..., i, j] = a[..., (i + M // 2) % M, (j + K // 2) % K]
    return out


def fftshift2(a):
    """out[i, j] = a[(i - M//2) % M, (j - K//2) % K] over the last two axes."""
    a = np.asarray(a)
    M, K = a.shape[-2:]
    out = np.empty_like(a)
    for i in range(M):
        for j in range(K):
            out[..., i, j] = a[..., (i - M // 2) % M, (j - K // 2) % K]
    return out


def roll2(a, s0, s1):
    """Periodic roll of a 2-D array: out[i, j] = a[(i - s0) % U, (j - s1) % V]."""
    U, V = a.shape
    out = np.empty_like(a)
    for i in range(U):
        for j in range(V):
            out[i, j] = a[(i - s0) % U, (j - s1) % V]
    return out


def _abs2(z):
    return z.real * z.real + z.imag * z.imag


# ---------------------------------------------------------------------------
# projections (era.py:336-433, :550-610; dm.py:337-341)
# ---------------------------------------------------------------------------

def make_exits(O, P, R):
    """psi_n[i, j] = O[(i - r0) % U, (j - r1) % V] * P[i, j]   (era.py:426-433)."""
    N = len(R)
    M, K = P.shape
    U, V = O.shape
    ex = np.zeros((N, M, K), dtype=complex)
    for n in range(N):
        r0, r1 = int(R[n][0]), int(R[n][1])
        for i in range(M):
            for j in range(K):
                ex[n, i, j] = O[(i - r0) % U, (j - r1) % V] * P[i, j]
    return ex


def p_heatmap(P, R, shape):
    """sum_n |P|^2 at window n: H[i - r0, j - r1] += |P[i, j]|^2   (era.py:402-410)."""
    M, K = P.shape
    H = np.zeros(shape)
    for r0, r1 in R:
        for i in range(M):
            for j in range(K):
                H[i - int(r0), j - int(r1)] += _abs2(P[i, j])
    return H


def o_heatmap(O, R, shape):
    """sum_n |roll(O, r_n)|^2 cropped to the probe   (era.py:412-417)."""
    U, V = O.shape
    G = np.zeros(shape)
    for r0, r1 in R:
        for i in range(shape[0]):
            for j in range(shape[1]):
                G[i, j] += _abs2(O[(i - int(r0)) % U, (j - int(r1)) % V])
    return G


def psup_O(exits, P, R, O_shape, H=None, alpha=1e-10):
    """O[u, v] = sum_n conj(P) psi_n at window n / (H + alpha)   (era.py:336-369)."""
    if H is None:
        H = p_heatmap(P, R, O_shape)
    M, K = P.shape
    num = np.zeros(O_shape, dtype=complex)
    for n, (r0, r1) in enumerate(R):
        for i in range(M):
            for j in range(K):
                num[i - int(r0), j - int(r1)] += exits[n, i, j] * P[i, j].conjugate()
    O = np.zeros(O_shape, dtype=complex)
    for u in range(O_shape[0]):
        for v in range(O_shape[1]):
            O[u, v] = num[u, v] / (H[u, v] + alpha)
    return O, H


def psup_P(exits, O, R, G=None, alpha=1e-10):
    """P[i, j] = sum_n conj(O_n) psi_n / (G + alpha)   (era.py:371-400)."""
    M, K = exits.shape[1:]
    if G is None:
        G = o_heatmap(O, R, (M, K))
    num = np.zeros((M, K), dtype=complex)
    for n, (r0, r1) in enumerate(R):
        for i in range(M):
            for j in range(K):
                num[i, j] += exits[n, i, j] * O[i - int(r0), j - int(r1)].conjugate()
    P = np.zeros((M, K), dtype=complex)
    for i in range(M):
        for j in range(K):
            P[i, j] = num[i, j] / (G[i, j] + alpha)
    return P, G


def _mask_at(mask, i, j):
    return float(mask) if np.ndim(mask) == 0 else float(mask[i][j])


def pmod_1(amp, exits, mask, alpha):
    """Modulus step keeping masked pixels; returns (exits', eMod sum)   (era.py:550-572)."""
    N, M, K = exits.shape
    out = np.zeros_like(exits)
    emod = 0.0
    array_mask = np.ndim(mask) > 0
    for n in range(N):
        F = dft2(exits[n])
        for i in range(M):
            for j in range(K):
                mg = math.sqrt(_abs2(F[i][j])) + alpha
                m = _mask_at(mask, i, j)
                emod += (mg - amp[n, i, j]) ** 2 * m
                fac = amp[n, i, j] / mg
                if array_mask and m == 0:
                    fac = 1.0
                F[i][j] = F[i][j] * fac
        out[n] = dft2(F, inverse=True)
    return out, emod


def pmod_7(amp, b, exits, mask, alpha):
    """Modulus step with background; masked pixels -> 0; b scaled in place
    (era.py:574-595).  Returns (exits', eMod sum)."""
    N, M, K = exits.shape
    out = np.zeros_like(exits)
    emod = 0.0
    for n in range(N):
        F = dft2(exits[n])
        for i in range(M):
            for j in range(K):
                mg = math.sqrt(_abs2(F[i][j]) + b[n, i, j] ** 2 + alpha)
                m = _mask_at(mask, i, j)
                emod += (mg - amp[n, i, j]) ** 2 * m
                fac = m * amp[n, i, j] / mg
                F[i][j] = F[i][j] * fac
                b[n, i, j] = b[n, i, j] * fac
        out[n] = dft2(F, inverse=True)
    return out, emod


def pmod_P(Pamp, P, mask, alpha):
    """Far-field probe modulus, masked pixels kept   (era.py:597-610)."""
    M, K = P.shape
    F = dft2(P)
    array_mask = np.ndim(mask) > 0
    for i in range(M):
        for j in range(K):
            fac = Pamp[i, j] / (math.sqrt(_abs2(F[i][j])) + alpha)
            if array_mask and _mask_at(mask, i, j) == 0:
                fac = 1.0
            F[i][j] = F[i][j] * fac
    return np.array(dft2(F, inverse=True))


def model_error_1(amp, exits, mask, b=None):
    """sum mask (sqrt(|FFT psi|^2 + b^2) - amp)^2, no alpha   (dm.py:337-341)."""
    N, M, K = exits.shape
    err = 0.0
    for n in range(N):
        F = dft2(exits[n])
        for i in range(M):
            for j in range(K):
                bb = 0.0 if b is None else b[n, i, j] ** 2
                mg = math.sqrt(_abs2(F[i][j]) + bb)
                err += _mask_at(mask, i, j) * (mg - amp[n, i, j]) ** 2
    return err


def centre_shift(P):
    """rint(centre of mass of |P|^2) - shape // 2   (era.py:172-176)."""
    M, K = P.shape
    tot = sy = sx = 0.0
    for i in range(M):
        for j in range(K):
            w = _abs2(P[i, j])
            tot += w
            sy += i * w
            sx += j * w
    return round(sy / tot) - M // 2, round(sx / tot) - K // 2


# ---------------------------------------------------------------------------
# preamble (era.py:435-548)
# ---------------------------------------------------------------------------

def preamble(I, R, P, O, OP_iters, mask, background, method):
    I = np.asarray(I, dtype=np.float64)
    N, M, K = I.shape
    if method is None:
        if O is None and P is None:
            method = 3
        elif O is None:
            method = 1
        elif P is None:
            method = 2
        if background is not None:
            method += 3
    update = {1: 'O', 4: 'O', 2: 'P', 5: 'P', 3: 'OP', 6: 'OP'}[method]
    if isinstance(OP_iters, int):
        OP_iters = (OP_iters, 1)
    if O is None:
        U = M + int(max(r[0] for r in R)) - int(min(r[0] for r in R))
        V = K + int(max(r[1] for r in R)) - int(min(r[1] for r in R))
        O = np.ones((U, V), dtype=complex)
    if P is None:
        raise ValueError("the loop oracle needs an explicit probe (the reference draws it unseeded)")
    O = np.array(O, dtype=complex)
    P = np.array(P, dtype=complex)
    # I_norm with the un-shifted mask (era.py:495)
    I_norm = 0.0
    for n in range(N):
        for i in range(M):
            for j in range(K):
                I_norm += _mask_at(mask, i, j) * I[n, i, j]
    amp = np.zeros((N, M, K))
    for n in range(N):
        for i in range(M):
            for j in range(K):
                amp[n, i, j] = math.sqrt(I[n, (i + M // 2) % M, (j + K // 2) % K])
    mask_s = mask if np.ndim(mask) == 0 else ifftshift2(np.asarray(mask))
    # R shifted in place so that each column's maximum is 0 (era.py:504-505)
    m0 = max(int(r[0]) for r in R)
    m1 = max(int(r[1]) for r in R)
    for n in range(len(R)):
        R[n][0] -= m0
        R[n][1] -= m1
    exits = make_exits(O, P, R)
    bg = None
    if method in (4, 5, 6):
        if background is None:
            raise ValueError("the loop oracle needs an explicit background (the reference draws it unseeded)")
        bs = ifftshift2(np.asarray(background, dtype=np.float64))
        bg = np.zeros((N, M, K))
        for n in range(N):
            for i in range(M):
                for j in range(K):
                    bg[n, i, j] = math.sqrt(bs[i, j])
    FP = dft2(P)
    Pamp = np.array([[math.sqrt(_abs2(FP[i][j])) for j in range(K)] for i in range(M)])
    return dict(method=method, update=update, OP_iters=OP_iters, O=O, P=P, amp=amp,
                mask=mask_s, I_norm=I_norm, exits=exits, bg=bg, Pamp=Pamp)


def _flat(update, O, P):
    vals = []
    if update in ('O', 'OP'):
        vals += [complex(v) for v in O.ravel()]
    if update in ('P', 'OP'):
        vals += [complex(v) for v in P.ravel()]
    return vals


def _econ(bak, cur):
    num = sum(_abs2(a - b) for a, b in zip(bak, cur))
    den = sum(_abs2(b) for b in cur)
    return math.sqrt(num / den)


def _frame_mean(b):
    """Every frame of b replaced by the mean over frames (era.py:265-270)."""
    N, M, K = b.shape
    out = np.zeros_like(b)
    for i in range(M):
        for j in range(K):
            s = 0.0
            for n in range(N):
                s += b[n, i, j]
            for n in range(N):
                out[n, i, j] = s / N
    return out


def _sq_dist(a, b):
    return sum(_abs2(x - y) for x, y in zip(a.ravel(), b.ravel()))


def _info(exits, eMods, eCons, eSups, H, bframe):
    N, M, K = exits.shape
    I = np.zeros((N, M, K))
    for n in range(N):
        F = dft2(exits[n])
        for i in range(M):
            for j in range(K):
                I[n, i, j] = _abs2(F[i][j])
    info = dict(I=fftshift2(I), eMod=eMods, eCon=eCons, eSup=eSups, heatmap=H)
    if bframe is not None:
        info['background'] = fftshift2(bframe) ** 2
    return info


# ---------------------------------------------------------------------------
# ERA (era.py:152-291)
# ---------------------------------------------------------------------------

def _consistency(update, i, OP_iters, exits, O, P, R, H, G, alpha, probe_centering):
    """The consistency block of era.py:163-189 / dm.py:56-85; returns
    (O, P, H, G, cm or None)."""
    cm = None
    if update == 'O':
        O, H = psup_O(exits, P, R, O.shape, H, alpha)
    elif update == 'P':
        P, G = psup_P(exits, O, R, G, alpha)
    elif i % OP_iters[1] == 0:
        for _ in range(OP_iters[0]):
            O, H = psup_O(exits, P, R, O.shape, None, alpha)
            P, G = psup_P(exits, O, R, None, alpha)
        if probe_centering:
            cm = centre_shift(P)
            P = roll2(P, -cm[0], -cm[1])
            O = roll2(O, -cm[0], -cm[1])
            H = G = None
    else:
        O, H = psup_O(exits, P, R, O.shape, H, alpha)
    return O, P, H, G, cm


def ERA(I, R, P, O, iters, OP_iters=1, mask=1, background=None, method=None,
        Pmod_probe=False, probe_centering=False, alpha=1.0e-10, callback=None):
    """ERA for one rank; returns (O, P, info) whatever the update mode."""
    st = preamble(I, R, P, O, OP_iters, mask, background, method)
    method, update, OP_iters = st['method'], st['update'], st['OP_iters']
    O, P, amp, mask, I_norm = st['O'], st['P'], st['amp'], st['mask'], st['I_norm']
    exits, bg, Pamp = st['exits'], st['bg'], st['Pamp']
    M, K = P.shape
    H = G = None
    eMods, eCons, eSups = [], [], []
    for i in range(iters):
        bak = _flat(update, O, P)
        if bg is None:
            exits, emod = pmod_1(amp, exits, mask, alpha)
        else:
            exits, emod = pmod_7(amp, bg, exits, mask, alpha)
        O, P, H, G, _ = _consistency(update, i, OP_iters, exits, O, P, R, H, G, alpha,
                                     probe_centering)
        if Pmod_probe is not None and i < Pmod_probe:
            P = pmod_P(Pamp, P, mask, alpha)
        if bg is not None:
            bg = _frame_mean(bg)
        new = make_exits(O, P, R)
        eSups.append(math.sqrt(M * K * _sq_dist(exits, new) / I_norm))
        exits = new
        eCons.append(_econ(bak, _flat(update, O, P)))
        eMods.append(math.sqrt(emod / I_norm))
        if callback is not None:
            callback(i, O, P)
    return O, P, _info(exits, eMods, eCons, eSups, H, None if bg is None else bg[0])


# ---------------------------------------------------------------------------
# DM (dm.py:45-294)
# ---------------------------------------------------------------------------

def DM(I, R, P, O, iters, OP_iters=1, mask=1, background=None, Pmod_probe=False,
       probe_centering=False, method=None, alpha=1.0e-10, callback=None):
    """Difference map for one rank; returns (O, P, info) whatever the update mode."""
    st = preamble(I, R, P, O, OP_iters, mask, background, method)
    method, update, OP_iters = st['method'], st['update'], st['OP_iters']
    O, P, amp, mask, I_norm = st['O'], st['P'], st['amp'], st['mask'], st['I_norm']
    f, bg, Pamp = st['exits'], st['bg'], st['Pamp']
    M, K = P.shape
    N = len(R)
    H = G = None
    eMods, eCons, eSups = [], [], []
    bak = _flat(update, O, P)
    Os, Ps = O, P
    bbar = None
    for i in range(iters):
        if bg is not None:
            f, _ = pmod_7(amp, bg, f, mask, alpha)            # dm.py:171
        O, P, H, G, cm = _consistency(update, i, OP_iters, f, O, P, R, H, G, alpha,
                                      probe_centering)
        if cm is not None:                                    # dm.py:82 / :200
            f = np.array([roll2(f[n], -cm[0], -cm[1]) for n in range(N)])
        if Pmod_probe is not None and i < Pmod_probe:
            P = pmod_P(Pamp, P, mask, alpha)
        f_in = f.copy()
        ex0 = make_exits(O, P, R)                            # Ps f
        if bg is None:
            # f <- f - Ps f + Pm(2 Ps f - f)   (dm.py:91-97)
            arg = np.zeros_like(f)
            for n in range(N):
                for a in range(M):
                    for c in range(K):
                        arg[n, a, c] = 2.0 * ex0[n, a, c] - f[n, a, c]
            pm, _ = pmod_1(amp, arg, mask, alpha)
            for n in range(N):
                for a in range(M):
                    for c in range(K):
                        f[n, a, c] = f[n, a, c] - ex0[n, a, c] + pm[n, a, c]
        else:
            # the same update on (f, b) with b's projection the frame mean (dm.py:208-224)
            bmean = _frame_mean(bg)
            arg = np.zeros_like(f)
            barg = np.zeros_like(bg)
            for n in range(N):
                for a in range(M):
                    for c in range(K):
                        arg[n, a, c] = 2.0 * ex0[n, a, c] - f[n, a, c]
                        barg[n, a, c] = 2.0 * bmean[n, a, c] - bg[n, a, c]
            pm, _ = pmod_7(amp, barg, arg, mask, alpha)
            for n in range(N):
                for a in range(M):
                    for c in range(K):
                        f[n, a, c] = f[n, a, c] - ex0[n, a, c] + pm[n, a, c]
                        bg[n, a, c] = bg[n, a, c] - bmean[n, a, c] + barg[n, a, c]
        eSups.append(math.sqrt(M * K * _sq_dist(f, f_in) / I_norm))

        # metrics: Os, Ps = Psup(f) without touching the cached heatmaps on
        # OP iterations and without rolling f (dm.py:104-137 / :231-271)
        Os, Ps = O.copy(), P.copy()
        if update == 'O':
            Os, H = psup_O(f, Ps, R, O.shape, H, alpha)
        elif update == 'P':
            Ps, G = psup_P(f, Os, R, G, alpha)
        elif i % OP_iters[1] == 0:
            for _ in range(OP_iters[0]):
                Os, _h = psup_O(f, Ps, R, O.shape, None, alpha)
                Ps, _g = psup_P(f, Os, R, None, alpha)
            if probe_centering:
                cs = centre_shift(Ps)
                Ps = roll2(Ps, -cs[0], -cs[1])
                Os = roll2(Os, -cs[0], -cs[1])
        else:
            Os, H = psup_O(f, P, R, O.shape, H, alpha)
        if Pmod_probe is not None and i < Pmod_probe:
            Ps = pmod_P(Pamp, Ps, mask, alpha)
        exm = make_exits(Os, Ps, R)
        if bg is None:
            emod = model_error_1(amp, exm, mask)
        else:
            bbar = _frame_mean(bg)
            emod = model_error_1(amp, exm, mask, bbar)
        cur = _flat(update, Os, Ps)
        eCons.append(_econ(bak, cur))
        eMods.append(math.sqrt(emod / I_norm))
        bak = cur
        if callback is not None:
            callback(i, O, P)
    exf = make_exits(Os, Ps, R)
    return O, P, _info(exf, eMods, eCons, eSups, H, None if bbar is None else bbar[0])
