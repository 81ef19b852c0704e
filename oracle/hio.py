"""Oracle: hybrid input-output (HIO) on the ERA path's projections (TEST
INFRASTRUCTURE ONLY).

Restates the legacy method ``Ptychography.HIO(iters, beta, update)`` of
``ptychography/Ptychography_2dsample_2dprobe_farfield.py:93-127`` with the
update of ``ptychography/utility_Ptych.py:14-17``::

    out  = Pmod(exit)
    exit = exit + beta * Psup((1 + 1/beta) * out - 1/beta * exit) - beta * out

The exits start as ``makeExits(sample, probe, coords)`` = O * P
(``Ptychography_2dsample_2dprobe_farfield.py:29``).  ``Psup`` by ``update``
(``:97-111``):

* ``'sample'``: the exits of (psup_O(z), P);
* ``'probe'``: the exits of (O, psup_P(z));
* ``'both'``: ``temp_probe(temp_sample(z))`` -- O from z, then P from the exits
  of (new O, old P), and the exits of (new O, new P).

Both O and P keep the values the last ``Psup`` left (the class updates them in
place, ``inPlace=True``).  The projections are the ERA path's own (``pmod_1``,
``psup_O``, ``psup_P`` of ptychography/era.py, with the ERA preamble's
amplitude, mask and R conventions, era.py:495-505), not the legacy class's
(its ``Psup_sample`` thresholds the heatmap and its ``Pmod`` clips |psi_hat|
instead of adding alpha).  eMod per iteration is era.py's modulus error of the
exits entering the iteration, sqrt(sum mask (|fft psi| + alpha - amp)^2 /
I_norm) (era.py:561-563, 210); the legacy class evaluates its own every 4
iterations.  ``callback(i, O, P)`` sees the iterates after iteration i.
"""

import numpy as np

from .era import make_exits, pmod_1, psup_O, psup_P

UPDATES = ('sample', 'probe', 'both')


def HIO(I, R, P, O, iters, beta=1.0, update='sample', mask=1, alpha=1.0e-10,
        full_output=True, dtype=None, callback=None):
    """Legacy HIO method (see the module note); returns (O, P, info) or (O, P)."""
    if update not in UPDATES:
        raise ValueError(f"update must be one of {UPDATES}")
    if beta == 0:
        raise ValueError("beta must be non-zero (the update divides by it)")
    if dtype is None:
        dtype = 'single' if I.dtype == np.float32 else 'double'
    r_dtype, c_dtype = (np.float32, np.complex64) if dtype == 'single' else (np.float64, np.complex128)
    if O is None:
        shape = (I.shape[1] + R[:, 0].max() - R[:, 0].min(),
                 I.shape[2] + R[:, 1].max() - R[:, 1].min())
        O = np.ones(shape, dtype=c_dtype)
    O = O.astype(c_dtype)
    P = P.astype(c_dtype)
    I_norm = np.sum(mask * I)                                           # era.py:495
    amp = np.fft.ifftshift(np.sqrt(I).astype(r_dtype), axes=(-2, -1))   # era.py:497-498
    mask = np.asarray(mask)
    if mask.ndim > 0:
        mask = np.fft.ifftshift(mask)                                   # era.py:499
    R[:, 0] -= R[:, 0].max()                                            # era.py:504-505
    R[:, 1] -= R[:, 1].max()

    b = r_dtype(beta)
    f = make_exits(O, P, R)                                             # farfield.py:29
    P_heatmap = O_heatmap = None
    eMods = []
    for i in range(iters):
        m, e = pmod_1(amp, f, mask, alpha=alpha, eMod_calc=True)       # utility_Ptych.py:15
        eMods.append(float(np.sqrt(e / I_norm)))
        z = (1 + 1 / b) * m - f / b                                     # :16, Psup's argument
        if update == 'sample':                                          # farfield.py:97-99
            O, P_heatmap = psup_O(z, P, R, O.shape, P_heatmap, alpha)
        elif update == 'probe':                                         # :101-103
            P, O_heatmap = psup_P(z, O, R, O_heatmap, alpha)
        else:                                                           # :105-109
            O, _ = psup_O(z, P, R, O.shape, None, alpha)
            P, _ = psup_P(make_exits(O, P, R), O, R, None, alpha)
        O = O.astype(c_dtype)
        P = P.astype(c_dtype)
        f = (f + b * make_exits(O, P, R) - b * m).astype(c_dtype)      # :16
        if callback is not None:
            callback(i, O, P)
    info = {'eMod': eMods, 'exits': f}
    if full_output:
        return O, P, info
    return O, P
