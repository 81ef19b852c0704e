"""``DM`` -- difference-map ptychography on one MI355X.

Drop-in for ``Ptychography.DM`` (ptychography/dm.py:13-333).  Note the
reference's positional order differs from ERA's: ``Pmod_probe``,
``probe_centering`` and ``method`` follow ``background``.  Returns the main
iterate (O, P) as the reference does (dm.py:316-318); ``info['I']`` comes from
the metric-path projections Os, Ps (dm.py:298-299).  ``info['eSup']`` is the DM
step size sqrt(M K sum |f_{i+1} - f_i|^2 / I_norm).
"""

from .era import ALGO_DM, solve

__all__ = ["DM"]


def DM(I, R, P, O, iters, OP_iters=1, mask=1, background=None, Pmod_probe=False,
       probe_centering=False, method=None, hardware='cpu', alpha=1.0e-10, dtype=None,
       full_output=True, device=None, callback=None, rng=None, init=None, init_net=None, graphs=False,
       check_finite=True, progress=False, refine_positions=0):
    """Difference map (ptychography/dm.py:13); build-only keywords as for ``ERA``."""
    return solve(ALGO_DM, I, R, P, O, iters, OP_iters=OP_iters, mask=mask,
                 background=background, method=method, Pmod_probe=Pmod_probe,
                 probe_centering=probe_centering, alpha=alpha, dtype=dtype,
                 full_output=full_output, device=device, callback=callback, rng=rng,
                 init=init, init_net=init_net, graphs=graphs, check_finite=check_finite,
                 progress=progress, refine_positions=refine_positions)
