"""CPU oracle for the ERA / difference-map iteration loop.

TEST INFRASTRUCTURE ONLY.  This package is a Python 3 / NumPy restatement of
the reference's ``ptychography/era.py`` and ``ptychography/dm.py`` (the
reference itself is Python 2 and imports mpi4py at module scope, so it cannot
be imported or run here).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline -- never as part of the shipped GPU path.

Every function cites the reference ``file:line`` it follows.  The MPI layer of
the reference is restated for one rank (``size == 1``): every ``Allreduce`` /
``reduce`` / ``bcast`` / ``scatter`` is the identity.

Pinning status (what checks each function against something other than its
own formula; the tests live in ``tests/test_oracle_*.py``):

* ``multiroll``                 -- nested ``np.roll`` and a brute-force index loop.
* ``make_exits``                -- brute-force index loop; periodic-roll formula;
                                   the "slice" window mode is bit-identical
                                   to the reference's roll.
* ``psup_O`` / ``psup_P``       -- brute-force loops, the adjoint identity, the
                                   exact-data fixed point.
* ``make_P_heatmap`` / ``make_O_heatmap`` -- adjoint applied to |P|^2 / |O|^2.
* ``pmod_1`` / ``Pmod_1``       -- direct O(n^2) DFT on small frames, closed-form
                                   modulus invariants, idempotence.
* ``pmod_7`` / ``Pmod_7``       -- closed-form invariants (masked pixels -> 0,
                                   background scaled by the same factor).
* ``Pmod_P``                    -- modulus invariant on the probe.
* ``center_of_mass``            -- ``scipy.ndimage.center_of_mass``.
* ``back_projection.Back_projection`` -- exact-data closed form (flat object:
                                   the exits are P, O = H / (H + alpha)) and the
                                   exit spectra's modulus (mask * sqrt I) and
                                   phase (the probe's far field).
* ``ERA`` / ``DM`` (methods 1-6) -- ``tests/loop_oracle.py``, a second
                                   restatement of the reference loops with
                                   explicit per-pixel loops and a direct DFT
                                   (no code shared with this package): same O,
                                   P, eMod, eCon, eSup, info['I'], heatmap and
                                   background after every iteration for methods
                                   1-6, OP_iters (1,2) / (2,3) (the stale-heatmap
                                   branch), probe centring, the Pmod_probe
                                   window and a mask (``tests/test_oracle_loops.py``,
                                   which also shows that four planted mistakes
                                   break the comparison).  Properties: ERA
                                   eMod_{i+1} <= eSup_i <= eMod_i (methods 1-3),
                                   non-increasing eMod with eSup_i <= eMod_i
                                   (methods 4-6), the exact-data fixed point
                                   (ERA 1; DM 1, 3 and 4 with a known
                                   background), the centring roll against
                                   scipy's centroid, the Pmod_probe window.
* ``hio.HIO``                   -- Fienup's classic input-output rule, to which
                                   HIO reduces when Psup is a 0/1 support
                                   projection (binary probe, frames that do not
                                   overlap; one frame and a binary object for
                                   'probe'); the exact-data fixed point; the
                                   first eMod equals ERA's.
* ``numpy.fft`` itself           -- parity unpinned against the reference's own
                                   (unpinned) numpy version; pinned here by the
                                   direct DFT and closed forms.
"""

from .era import (ERA, preamble, pmod_1, Pmod_1, pmod_7, Pmod_7, Pmod_P, psup_O,
                  psup_P, make_P_heatmap, make_O_heatmap, make_exits, multiroll,
                  center_of_mass, probe_centre_shift, use_windows)
from .dm import DM, model_error_1
from .hio import HIO

__all__ = ["ERA", "DM", "preamble", "pmod_1", "Pmod_1", "pmod_7", "Pmod_7",
           "Pmod_P", "psup_O", "psup_P", "make_P_heatmap", "make_O_heatmap",
           "make_exits", "multiroll", "center_of_mass", "probe_centre_shift",
           "model_error_1", "HIO", "use_windows"]
