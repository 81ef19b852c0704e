"""``Back_projection`` -- the back-projection initial object on the GPU.

Drop-in for ``ptychography/back_projection.py:4-88``: each frame's exit wave
is the inverse FFT of the measured modulus with the phase of the probe's far
field, and the object is their overlap projection ``psup_O``.  Same arguments
and returns (``O`` or ``(O, info)`` with ``'exits'`` and ``'heatmap'``), same
in-place shift of ``R``.  The reference's file does not run as written; this
follows its steps with the ERA preamble's conventions (sqrt(I) and the mask
ifftshifted so their DC matches ``fftn(P)``'s, ptychography/era.py:495-499).
Computed by ``pty_back_projection`` (one FFT of P, a spectra kernel, the
batched inverse FFT and the deterministic object gather).  Where the probe's
far field vanishes (outside a hard pupil) its phase is round-off, as in the
reference: any two implementations differ there.
"""

import numpy as np
import torch

from .engine import Engine, require_gpu
from .era import _as_numpy

__all__ = ["Back_projection"]


def Back_projection(I, R, P, O, mask=1, alpha=1.0e-10, full_output=True, device=None):
    """ptychography/back_projection.py:4-88 on the GPU ``device``."""
    dev = require_gpu(device)
    if not isinstance(R, np.ndarray):
        raise TypeError("R must be a numpy integer array (it is shifted in place)")
    N, M, K = (int(s) for s in I.shape)
    P = _as_numpy(P)
    c_dtype = np.complex128 if P.dtype == np.complex128 else np.complex64
    r_dtype = np.float64 if c_dtype == np.complex128 else np.float32
    if P.shape != (M, K):
        raise ValueError(f"P has shape {P.shape}, frames are {(M, K)}")
    if O is None:
        O = np.ones((M + int(R[:, 0].max() - R[:, 0].min()), K + int(R[:, 1].max() - R[:, 1].min())),
                    dtype=c_dtype)
    U, V = _as_numpy(O).shape
    R[:, 0] -= R[:, 0].max()
    R[:, 1] -= R[:, 1].max()
    eng = Engine(N, M, K, U, V, R, device=dev)
    try:
        mask_c = mask_s = None
        if not (np.isscalar(mask) and mask == 1):
            m = _as_numpy(mask)
            if m.shape != (M, K):
                raise ValueError(f"mask must be 1 or an array of shape {(M, K)}")
            mc = m.astype(np.uint8)
            mask_c = torch.from_numpy(np.ascontiguousarray(mc)).to(dev)
            mask_s = torch.from_numpy(np.ascontiguousarray(np.fft.ifftshift(mc))).to(dev)
        I_t = I.to(dev) if isinstance(I, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(I, dtype=np.float32))
        amp, I_norm = eng.prepare_amp(I_t, mask_c)
        eng.set_data(amp, mask_s, I_norm)
        P_t = torch.from_numpy(np.ascontiguousarray(P.astype(np.complex64))).to(dev)
        O_t, H, ex = eng.back_projection(P_t, alpha=alpha, exits=full_output)
        O_out = O_t.cpu().numpy().astype(c_dtype)
        if not full_output:
            return O_out
        return O_out, {'exits': ex.cpu().numpy().astype(c_dtype),
                       'heatmap': H.cpu().numpy().astype(r_dtype)}
    finally:
        eng.close()
