"""NumPy restatement of the reference difference-map (DM) solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Follows
``ptychography/dm.py`` of andyofmelbourne/Ptychography for one MPI rank and
reuses the restated ERA helpers, as the reference reuses ``era.*``.

Build additions: ``info['eSup']`` is the DM step size
``sqrt(M*K * sum |f_new - f_in|^2 / I_norm)`` where ``f_in`` is the iterate the
DM update starts from (after the leading ``pmod_7`` of methods 4-6) -- the
legacy GPU class's ``error_conv``
(``ptychography/Ptychography_2dsample_2dprobe_farfield_gpu.py:86``) in era.py's
FFT convention; ``callback(i, O, P)`` as in the ERA oracle.
"""

import numpy as np

from . import era as _era


def model_error_1(amp, exits, mask, background=0):
    """DM modulus error ``sum mask (sqrt(|FFT psi|^2 + b^2) - amp)^2``, no alpha (dm.py:337-341)."""
    ex = np.fft.fftn(exits, axes=(-2, -1))
    Mg = np.sqrt((ex.conj() * ex).real + background ** 2)
    return np.sum(mask * (Mg - amp) ** 2)


def _project(update, i, OP_iters, exits, O, P, R, P_heatmap, O_heatmap, alpha,
             probe_centering, metric):
    """The consistency-projection block shared by the main and metric paths.

    Main path: dm.py:56-85 (and :174-202).  Metric path: dm.py:106-130 (and
    :233-257) -- there heatmaps computed in the OP branch are discarded and the
    exits are not rolled.  Returns (O, P, P_heatmap, O_heatmap, cm or None).
    """
    cm = None
    if update == 'O':
        # P may be the per-frame probe stack of sub-pixel positions
        O, P_heatmap = _era.psup_O(exits, P, R, O.shape, P_heatmap, alpha)
    elif update == 'P':
        P, O_heatmap = _era.psup_P(exits, O, R, O_heatmap, alpha)
    else:
        if i % OP_iters[1] == 0:
            for _ in range(OP_iters[0]):
                O, ph = _era.psup_O(exits, P, R, O.shape, None, alpha)
                P, oh = _era.psup_P(exits, O, R, None, alpha)
                if not metric:
                    P_heatmap, O_heatmap = ph, oh
            if probe_centering:
                cm = _era.probe_centre_shift(P)
                P = _era.multiroll(P, -cm)
                O = _era.multiroll(O, -cm)
                if not metric:
                    P_heatmap = O_heatmap = None
        else:
            O, P_heatmap = _era.psup_O(exits, P, R, O.shape, P_heatmap, alpha)
    return O, P, P_heatmap, O_heatmap, cm


def DM(I, R, P, O, iters, OP_iters=1, mask=1, background=None, Pmod_probe=False,
       probe_centering=False, method=None, hardware='cpu', alpha=1.0e-10,
       dtype=None, full_output=True, callback=None, rng=None, resume=None,
       state_callback=None, windows=None):
    """Difference map, one-rank restatement of dm.py:13-333.

    f_{i+1} = f_i - Ps f_i + Pm(2 Ps f_i - f_i)  (dm.py:94-97); metrics use
    Os, Ps = Psup(f_{i+1}) (dm.py:104-137).  Note the positional order of
    ``Pmod_probe`` / ``probe_centering`` / ``method`` differs from ERA's.
    """
    st = _era.preamble(I, R, P, O, iters, OP_iters, mask, background, method,
                       alpha, dtype, rng, windows)
    method, update, OP_iters = st['method'], st['update'], st['OP_iters']
    bg_mode = method in (4, 5, 6)
    i0, P_heatmap, O_heatmap, background = _era._resume(st, resume, bg_mode)
    O, P, amp, Pamp = st['O'], st['P'], st['amp'], st['Pamp']
    R, mask = st['R'], st['mask']
    I_norm, exits = st['I_norm'], st['exits']
    MK = P.shape[0] * P.shape[1]

    eMods, eCons, eSups = [], [], []
    bak = _era._state(update, O, P).copy()
    if resume is not None:
        if resume.get('exits') is not None:
            exits = np.array(resume['exits'], dtype=st['c_dtype'])
        if resume.get('O_ref') is not None:
            bak = _era._state(update, np.asarray(resume['O_ref'], st['c_dtype']),
                              np.asarray(resume['P_ref'], st['c_dtype'])).copy()
    Os, Ps = O, P
    b_0 = np.empty_like(background) if bg_mode else None
    # sub-pixel positions: the frames see per-frame probes (object-only methods)
    Pf = st['Pn']
    if Pf is not None and Pmod_probe is not None and Pmod_probe is not False:
        raise ValueError("sub-pixel positions are modelled without the far-field probe constraint")

    for i in range(i0, i0 + iters):
        if bg_mode:
            # leading modulus projection of f and b (dm.py:171)
            exits, background = _era.pmod_7(amp, background, exits, mask, alpha=alpha)

        O, Pp, P_heatmap, O_heatmap, cm = _project(
            update, i, OP_iters, exits, O, P if Pf is None else Pf, R, P_heatmap, O_heatmap,
            alpha, probe_centering, metric=False)
        if Pf is None:
            P = Pp
        if cm is not None:
            # dm.py:82 / :200 -- the remembered exits follow the probe shift
            exits = _era.multiroll(exits, (0, -cm[0], -cm[1]))

        if Pmod_probe is not None and i < Pmod_probe:
            P = _era.Pmod_P(Pamp, P, mask, alpha)

        f_in = exits.copy()
        ex_0 = _era.make_exits(O, P if Pf is None else Pf, R)
        if bg_mode:
            b_0[:] = np.mean(background, axis=0)          # dm.py:208-213
            exits -= ex_0                                 # f - Ps f
            background -= b_0
            ex_0 -= exits                                 # 2 Ps f - f
            b_0 -= background
            ex_0, b_0 = _era.pmod_7(amp, b_0, ex_0, mask, alpha=alpha)
            exits += ex_0
            background += b_0
        else:
            exits -= ex_0                                 # dm.py:94-97
            ex_0 -= exits
            ex_0 = _era.pmod_1(amp, ex_0, mask, alpha=alpha)
            exits += ex_0

        d = exits - f_in
        eSups.append(float(np.sqrt(MK * np.sum((d * d.conj()).real) / I_norm)))

        # metrics (dm.py:104-137 / :231-271)
        Os, Psp, P_heatmap, O_heatmap, _ = _project(
            update, i, OP_iters, exits, O.copy(), P.copy() if Pf is None else Pf, R, P_heatmap,
            O_heatmap, alpha, probe_centering, metric=True)
        Ps = Psp if Pf is None else P.copy()
        if Pmod_probe is not None and i < Pmod_probe:
            Ps = _era.Pmod_P(Pamp, Ps, mask, alpha)

        ex_m = _era.make_exits(Os, Ps if Pf is None else Pf, R)
        if bg_mode:
            b_0 = np.empty_like(background)
            b_0[:] = np.mean(background, axis=0)          # dm.py:263-268
            eMod = model_error_1(amp, ex_m, mask, b_0)
        else:
            eMod = model_error_1(amp, ex_m, mask)

        cur = _era._state(update, Os, Ps)
        eCons.append(_era._econ(bak, cur))
        eMods.append(np.sqrt(eMod / I_norm))
        bak = cur.copy()
        if callback is not None:
            callback(i, O, P)
        if state_callback is not None:
            state_callback(i, dict(O=O.copy(), P=P.copy(), exits=exits.copy(),
                                   background=None if background is None else background.copy(),
                                   P_heatmap=_era._copy(P_heatmap), O_heatmap=_era._copy(O_heatmap),
                                   O_ref=Os.copy(), P_ref=Ps.copy()))

    if not full_output:
        if update == 'O':
            return O
        if update == 'P':
            return P
        return O, P

    ex_f = _era.make_exits(Os, Ps if Pf is None else Pf, R)
    info = {}
    info['I'] = np.fft.fftshift(np.abs(np.fft.fftn(ex_f, axes=(-2, -1))) ** 2,
                                axes=(-2, -1))
    info['eMod'] = eMods
    info['eCon'] = eCons
    info['eSup'] = eSups
    info['heatmap'] = P_heatmap
    if bg_mode:
        info['background'] = np.fft.fftshift(b_0[0]) ** 2
    if update == 'O':
        return O, info
    if update == 'P':
        return P, info
    return O, P, info
