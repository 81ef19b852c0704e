"""``HIO`` -- hybrid input-output on the GPU.

The reference keeps HIO in its legacy class: ``Ptychography.HIO(iters, beta,
update)`` (ptychography/Ptychography_2dsample_2dprobe_farfield.py:93-127)
iterates the update of ptychography/utility_Ptych.py:14-17 on exit waves that
start as O * P (``:29``)::

    m = Pmod(f)
    f <- f + beta * Psup((1 + 1/beta) m - f / beta) - beta * m

with ``Psup`` chosen by ``update`` (``:97-111``): 'sample' (object from its
argument, probe fixed), 'probe' (probe from its argument, object fixed) or
'both' (object from its argument, then the probe from the exits of the new
object and the old probe).  This module offers it with the ERA call surface:
``HIO(I, R, P, O, iters, beta=1., update='sample', mask=1, alpha=1e-10,
full_output=True)`` returns ``(O, P, info)`` (``info['eMod']`` per iteration,
``info['exits']``) or ``(O, P)``, shifts ``R`` in place and takes ``O=None``
like ``ERA`` (ptychography/era.py:478-505).  The projections are the ERA
path's (pmod_1 with the preamble's centred amplitudes and mask, psup_O /
psup_P with alpha), and eMod is era.py's modulus error of the exits entering
each iteration (era.py:561-563, 210).

Every step runs on the path's kernels: ``pty_hio_modulus`` (the row / column
FFT stages with the masked modulus step and eMod partials; the iterate is kept
without its Psup term, which both row stages add back on the fly from the
object and probe, and the update's two linear combinations are applied as
Pmod(f) leaves the inverse row FFT) and ``pty_psup_O`` / ``pty_psup_P``
(deterministic gathers, heatmap kept while its factor is fixed);
``pty_add_exits`` completes the exits once for ``info['exits']``.
"""

import math

import numpy as np
import torch

from .engine import Engine, require_gpu
from .era import _as_numpy

__all__ = ["HIO"]

UPDATES = ('sample', 'probe', 'both')


def HIO(I, R, P, O, iters, beta=1.0, update='sample', mask=1, alpha=1.0e-10, full_output=True,
        device=None, callback=None):
    """Legacy ``Ptychography.HIO`` (see the module note) on the GPU ``device``."""
    if update not in UPDATES:
        raise ValueError(f"update must be one of {UPDATES}")
    if not (beta != 0 and math.isfinite(beta)):
        raise ValueError("beta must be finite and non-zero (the update divides by it)")
    if not isinstance(R, np.ndarray):
        raise TypeError("R must be a numpy integer array (it is shifted in place)")
    dev = require_gpu(device)
    N, M, K = (int(s) for s in I.shape)
    P = _as_numpy(P)
    if P.shape != (M, K):
        raise ValueError(f"P has shape {P.shape}, frames are {(M, K)}")
    c_dtype = np.complex128 if P.dtype == np.complex128 else np.complex64
    if O is None:
        O = np.ones((M + int(R[:, 0].max() - R[:, 0].min()), K + int(R[:, 1].max() - R[:, 1].min())),
                    dtype=c_dtype)
    O = _as_numpy(O)
    U, V = O.shape
    R[:, 0] -= R[:, 0].max()
    R[:, 1] -= R[:, 1].max()
    eng = Engine(N, M, K, U, V, R, device=dev)
    try:
        mask_c = mask_s = None
        if not (np.isscalar(mask) and mask == 1):
            mk = _as_numpy(mask)
            if mk.shape != (M, K):
                raise ValueError(f"mask must be 1 or an array of shape {(M, K)}")
            mc = mk.astype(np.uint8)
            mask_c = torch.from_numpy(np.ascontiguousarray(mc)).to(dev)
            mask_s = torch.from_numpy(np.ascontiguousarray(np.fft.ifftshift(mc))).to(dev)
        I_t = I.to(dev) if isinstance(I, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(I, dtype=np.float32))
        amp, I_norm = eng.prepare_amp(I_t, mask_c)
        eng.set_data(amp, mask_s, I_norm)
        O_t = torch.from_numpy(np.ascontiguousarray(O.astype(np.complex64))).to(dev)
        P_t = torch.from_numpy(np.ascontiguousarray(P.astype(np.complex64))).to(dev)
        # F holds the iterate without its pending Psup term: f = F + w O P, where w
        # is 0 at the start (F = O P) and beta after every iteration
        F = eng.make_exits(O_t, P_t)
        m = torch.empty_like(F)
        H = G = None
        w = 0.0
        eMods = []
        for i in range(int(iters)):
            # m: Psup's argument; F: f - beta Pmod(f)
            e = eng.hio_modulus(F, m, O_t, P_t, w, beta, alpha)
            eMods.append(math.sqrt(e / I_norm))
            if update == 'sample':
                O_t, H = eng.psup_O(m, P_t, H, alpha=alpha)
            elif update == 'probe':
                P_t, G = eng.psup_P(m, O_t, G, alpha=alpha)
            else:
                O_t, _ = eng.psup_O(m, P_t, None, alpha=alpha)
                P_t, _ = eng.psup_P(eng.make_exits(O_t, P_t), O_t, None, alpha=alpha)
            w = beta
            if callback is not None:
                callback(i, O_t.cpu().numpy().astype(c_dtype), P_t.cpu().numpy().astype(c_dtype))
        O_out = O_t.cpu().numpy().astype(c_dtype)
        P_out = P_t.cpu().numpy().astype(c_dtype)
        if not full_output:
            return O_out, P_out
        if w != 0.0:
            eng.add_exits(F, O_t, P_t, w)
        return O_out, P_out, {'eMod': eMods, 'exits': F.cpu().numpy().astype(c_dtype)}
    finally:
        eng.close()
