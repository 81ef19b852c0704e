"""Several independent scans advanced together on one GPU.

``ERA_many`` / ``DM_many`` take a list of per-scan argument dicts (the
arguments of ``ERA`` / ``DM``: ``I``, ``R``, ``P``, ``O`` and any keyword) and
return the list of what ``ERA`` / ``DM`` would return for each scan.  Each
scan keeps its own engine (frame size, object, method); their iterations run
through one ``pty_batch``: the scans' kernels go to separate lane streams so
small scans fill the device side by side, and with ``graphs=True`` one
iteration of every scan is captured as a single hipGraph and replayed.  Per
scan the results are identical to running that scan alone.  No counterpart in
the reference, whose loop handles one scan per call.
"""

import ctypes

import numpy as np

from ._abi import PtyOpts, call, load
from .era import ALGO_DM, ALGO_ERA, finish, prepare

__all__ = ["Batch", "ERA_many", "DM_many", "solve_many"]


class Batch:
    """A ``pty_batch`` over prepared engines (which stay owned by the caller)."""

    def __init__(self, engines, graphs=False):
        if not engines:
            raise ValueError("a batch needs at least one engine")
        self.engines = list(engines)
        ctxs = (ctypes.c_void_p * len(self.engines))(*[e._ctx.value for e in self.engines])
        self._b = ctypes.c_void_p()
        call("pty_batch_create", ctypes.byref(self._b), len(self.engines), ctxs)
        if graphs:
            call("pty_batch_set_graphs", self._b, 1)

    def run(self, opts, iters, iter0=0, begin=True):
        """Advance every scan ``iters`` iterations; returns one (iters, 3) array
        of eMod, eCon, eSup per scan."""
        n = len(self.engines)
        if len(opts) != n:
            raise ValueError("one pty_opts per scan")
        for e in self.engines:
            e._sync_torch()
        oa = (PtyOpts * n)(*opts)
        out = [np.zeros((max(iters, 0), 3), np.float64) for _ in range(n)]
        ptrs = (ctypes.c_void_p * n)(*[(o.ctypes.data if iters > 0 else None) for o in out])
        call("pty_batch_run", self._b, oa, int(iter0), int(iters), int(bool(begin)), ptrs)
        return out

    def graph_status(self):
        """(graphs on, a capture failed and the batch fell back to eager launches)."""
        en, fb = ctypes.c_int32(), ctypes.c_int32()
        call("pty_batch_get_graphs", self._b, ctypes.byref(en), ctypes.byref(fb))
        return bool(en.value), bool(fb.value)

    def close(self):
        if self._b:
            load().pty_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_many(algorithm, scans, iters, graphs=False, tune=False, full_output=True, device=None,
               check_finite=True, **common):
    """Run ``algorithm`` (0 = ERA, 1 = DM) on every scan of ``scans`` together.

    ``scans``: list of dicts of ERA / DM arguments (``I``, ``R``, ``P``, ``O``,
    optional keywords); ``common`` keywords apply to every scan.  ``tune=True``
    lets each engine pick its fastest launch configuration first (``pty_tune``).
    Each scan's ``R`` is shifted in place as ``ERA`` does."""
    jobs = []
    try:
        for sc in scans:
            kw = dict(common)
            kw.update(sc)
            I, R, P, O = kw.pop("I"), kw.pop("R"), kw.pop("P", None), kw.pop("O", None)
            jobs.append(prepare(algorithm, I, R, P, O, iters, device=device, **kw))
        if tune:
            for j in jobs:
                j.eng.tune(algorithm, j.method, **j.run_kwargs())
        batch = Batch([j.eng for j in jobs], graphs=graphs)
        try:
            metrics = batch.run([j.opts() for j in jobs], iters, begin=True)
            status = batch.graph_status()
        finally:
            batch.close()
    except BaseException:
        for j in jobs:
            j.eng.close()
        raise
    out = []
    for k, (j, m) in enumerate(zip(jobs, metrics)):
        j.graphs = False   # the batch, not the scan's engine, replayed the graphs
        r = finish(j, m, full_output, check_finite)
        if full_output and graphs:
            r[-1]['graphs'] = {'enabled': status[0], 'fell_back': status[1]}
        out.append(r)
    return out


def ERA_many(scans, iters, **kw):
    """``ERA`` on several scans at once; see ``solve_many``."""
    return solve_many(ALGO_ERA, scans, iters, **kw)


def DM_many(scans, iters, **kw):
    """``DM`` on several scans at once; see ``solve_many``."""
    return solve_many(ALGO_DM, scans, iters, **kw)
