"""Oracle: back-projection initial guess (TEST INFRASTRUCTURE ONLY).

Restates ``Back_projection`` of ``ptychography/back_projection.py:4-88``: each
frame's exit wave is the inverse FFT of the measured modulus with the phase of
the probe's far field, ``exits = ifftn(mask * sqrt(I) * exp(i angle(fftn(P))))``
(``:75-79``), and the object is their overlap projection ``psup_O``
(``:81``).  The reference as written does not run (Python 2 ``print R``
at ``:73``, ``era.psup_O_1`` which era.py does not define, and the centred
``sqrt(I)`` combined with the uncentred ``fftn(P)``); this restatement keeps its
steps with the conventions of the ERA preamble (``ptychography/era.py:495-505``):
sqrt(I) and the mask are ifftshifted so their DC sits where ``fftn(P)``'s does,
and R is shifted in place so that its maximum is 0.
"""

import numpy as np

from .era import psup_O


def Back_projection(I, R, P, O, mask=1, alpha=1.0e-10, full_output=True):
    """ptychography/back_projection.py:4-88 (see the module note on conventions)."""
    if O is None:   # :62-69
        shape = (I.shape[1] + R[:, 0].max() - R[:, 0].min(),
                 I.shape[2] + R[:, 1].max() - R[:, 1].min())
        O = np.ones(shape, dtype=P.dtype)
    R[:, 0] -= R[:, 0].max()   # :71-72
    R[:, 1] -= R[:, 1].max()
    amp = np.fft.ifftshift(np.sqrt(I.astype(np.float64)), axes=(-2, -1))   # :75, era.py:498
    m = np.asarray(mask)
    if m.ndim > 0:
        m = np.fft.ifftshift(m)                                             # era.py:499
    P_F = np.fft.fftn(P)                                                    # :77
    exits = m * amp * np.exp(1j * np.angle(P_F))                            # :78
    exits = np.fft.ifftn(exits, axes=(-2, -1))                              # :79
    O, P_heatmap = psup_O(exits, P, R, O.shape, None, alpha=alpha)          # :81
    if full_output:                                                         # :83-88
        return O, {'exits': exits, 'heatmap': P_heatmap}
    return O
