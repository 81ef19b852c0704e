"""Several independent scans advanced together on one GPU.

``ERA_many`` / ``DM_many`` take a list of per-scan argument dicts (the
arguments of ``ERA`` / ``DM``: ``I``, ``R``, ``P``, ``O`` and any keyword) and
return the list of what ``ERA`` / ``DM`` would return for each scan.

Scans of one frame size that run the same ERA step (methods 1-3 without
background, centring or Pmod_probe; method 3 with OP_iters (1, 1); integer
positions) are packed into one engine (``Engine.set_scans``): their frames
form one stack and their objects row bands of one object, so every stage of
an iteration is one launch over all of them.  The packed engines and the
scans that do not pack run through one ``pty_batch``: each engine's kernels go
to a lane stream so that frame sizes run side by side, and with
``graphs=True`` one iteration of every engine is captured as a single hipGraph
and replayed.  Per scan the object and probe are identical to running that
scan alone (metrics to FP64 rounding).  No counterpart in the reference, whose
loop handles one scan per call.
"""

import ctypes

import numpy as np
import torch

from ._abi import PtyOpts, call, load
from .engine import Engine, require_gpu
from .era import (ALGO_DM, ALGO_ERA, Job, finish, is_fractional, mask_array, pmod_probe_until,
                  prepare, resolve_method, result_dtypes, start_state)

__all__ = ["Batch", "ERA_many", "DM_many", "solve_many", "pack_key"]

# object bands of packed scans start on multiples of this many rows (the
# object-update tiles' height, so no tile holds two scans)
ROW_ALIGN = 16

_PACK_KEYWORDS = {"OP_iters", "mask", "background", "method", "Pmod_probe", "probe_centering",
                  "alpha", "dtype", "rng", "init", "init_net", "refine_positions"}


def pack_key(algorithm, kw):
    """Scans whose keys are equal (and not None) can share a packed engine:
    ERA methods 1-3 on integer positions, without background, probe
    centring, Pmod_probe, learned start or position refinement, method 3 with
    OP_iters (1, 1); the key holds the frame shape, method and alpha."""
    if algorithm != ALGO_ERA or set(kw) - _PACK_KEYWORDS - {"I", "R", "P", "O"}:
        return None
    if kw.get("background") is not None or kw.get("probe_centering"):
        return None
    if pmod_probe_until(kw.get("Pmod_probe", False)) > 0:
        return None
    if kw.get("init") is not None or kw.get("refine_positions"):
        return None
    R, I = kw.get("R"), kw.get("I")
    if not isinstance(R, np.ndarray) or is_fractional(R) or I is None or len(I.shape) != 3:
        return None
    try:
        method, _ = resolve_method(kw.get("O"), kw.get("P"), None, kw.get("method"))
    except ValueError:
        return None
    OP = kw.get("OP_iters", 1)
    OP = (int(OP), 1) if isinstance(OP, (int, np.integer)) else (int(OP[0]), int(OP[1]))
    if method == 3 and OP != (1, 1):
        return None
    return (int(I.shape[1]), int(I.shape[2]), method, float(kw.get("alpha", 1.0e-10)))


class _Pack:
    """Scans of one frame size in one packed engine; ``view(s)`` stands in
    for scan s's own engine where ``finish`` reads its results."""

    def __init__(self, algorithm, scans, iters, device):
        dev = require_gpu(device)
        prep = []
        for kw in scans:
            kw = dict(kw)
            I, R, P, O = kw.pop("I"), kw.pop("R"), kw.pop("P", None), kw.pop("O", None)
            method, update = resolve_method(O, P, None, kw.get("method"))
            N, M, K = (int(v) for v in I.shape)
            r_dtype, c_dtype = result_dtypes(I, kw.get("dtype"))
            O, P = start_state(R, P, O, M, K, c_dtype, kw.get("rng"))
            R[:, 0] -= R[:, 0].max()   # in place, as ERA does (era.py:504-505)
            R[:, 1] -= R[:, 1].max()
            OP = kw.get("OP_iters", 1)
            OP = (int(OP), 1) if isinstance(OP, (int, np.integer)) else (int(OP[0]), int(OP[1]))
            prep.append(dict(I=I, R=R, O=O, P=P, N=N, method=method, update=update,
                             r_dtype=r_dtype, c_dtype=c_dtype, OP=OP,
                             alpha=float(kw.get("alpha", 1.0e-10)),
                             mask=mask_array(kw.get("mask", 1), M, K)))
        M, K = (int(v) for v in prep[0]["I"].shape[1:])
        self.fstart = np.cumsum([0] + [p["N"] for p in prep])
        rstart = [0]
        for p in prep:
            rstart.append(rstart[-1] + -(-p["O"].shape[0] // ROW_ALIGN) * ROW_ALIGN)
        self.rstart = np.asarray(rstart)
        U, V = int(rstart[-1]), max(p["O"].shape[1] for p in prep)
        Rp = np.concatenate([p["R"].astype(np.int32) - np.array([[r0, 0]], np.int32)
                             for p, r0 in zip(prep, rstart)])
        self.eng = eng = Engine(len(Rp), M, K, U, V, Rp, device=dev)
        try:
            eng.set_scans(self.fstart, self.rstart)
            I_all = torch.cat([(p["I"] if isinstance(p["I"], torch.Tensor)
                                else torch.from_numpy(np.ascontiguousarray(p["I"], np.float32)))
                               .to(dev, torch.float32) for p in prep])
            mask_c = mask_s = None
            if any(p["mask"] is not None for p in prep):
                mc = np.stack([np.ones((M, K), np.uint8) if p["mask"] is None else p["mask"]
                               for p in prep])
                mask_c = torch.from_numpy(mc).to(dev)
                mask_s = torch.from_numpy(np.ascontiguousarray(
                    np.fft.ifftshift(mc, axes=(1, 2)))).to(dev)
            amp, _ = eng.prepare_amp(I_all, mask_c)
            del I_all
            eng.set_data(amp, mask_s, 0.0)
            O_all = torch.zeros((U, V), dtype=torch.complex64, device=dev)
            for p, r0 in zip(prep, rstart):
                u, v = p["O"].shape
                O_all[r0:r0 + u, :v] = torch.from_numpy(
                    np.ascontiguousarray(p["O"].astype(np.complex64))).to(dev)
            P_all = torch.from_numpy(np.ascontiguousarray(
                np.stack([p["P"].astype(np.complex64) for p in prep]))).to(dev)
            eng.set_state(O_all, P_all)
        except BaseException:
            eng.close()
            raise
        self.O_all, self.P_all = O_all, P_all
        self.prep = prep
        self.iters = iters
        self.algorithm = algorithm
        self._I = None
        self._H = None
        self._open = len(prep)

    def opts(self):
        p = self.prep[0]
        return Engine.opts(self.algorithm, p["method"], OP_iters=p["OP"], alpha=p["alpha"])

    def job(self, s):
        """Scan s as a ``Job`` for ``finish`` (its O and P are views of the packed state)."""
        p, r0 = self.prep[s], int(self.rstart[s])
        u, v = p["O"].shape
        return Job(algorithm=self.algorithm, eng=_ScanView(self, s), O_t=self.O_all[r0:r0 + u, :v],
                   P_t=self.P_all[s], method=p["method"], update=p["update"],
                   r_dtype=p["r_dtype"], c_dtype=p["c_dtype"], bg_init=None, background=None,
                   OP_iters=p["OP"], until=0, probe_centering=False, alpha=p["alpha"],
                   init_info=None, iters=self.iters, graphs=False, R_win=p["R"], d=None, refine=0)

    def intensity(self):
        if self._I is None:
            self._I = self.eng.intensity()
        return self._I

    def heatmap(self):
        if self._H is None:
            self._H = self.eng.heatmap()
        return self._H

    def release(self):
        self._open -= 1
        if self._open <= 0:
            self._I = self._H = None
            self.eng.close()


class _ScanView:
    """The parts of an ``Engine`` that ``finish`` uses, for one packed scan."""

    def __init__(self, pack, s):
        self.pack, self.s = pack, s

    def intensity(self):
        f = self.pack.fstart
        return self.pack.intensity()[int(f[self.s]):int(f[self.s + 1])]

    def heatmap(self):
        H = self.pack.heatmap()
        if H is None:
            return None
        r0 = int(self.pack.rstart[self.s])
        u, v = self.pack.prep[self.s]["O"].shape
        return H[r0:r0 + u, :v]

    def graph_status(self):
        return False, False

    def close(self):
        self.pack.release()


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
        """Advance every engine ``iters`` iterations; returns one (iters, 3)
        array of eMod, eCon, eSup per engine ((iters, S, 3) for packed ones)."""
        n = len(self.engines)
        if len(opts) != n:
            raise ValueError("one pty_opts per scan")
        for e in self.engines:
            e._sync_torch()
        oa = (PtyOpts * n)(*opts)
        out = [np.zeros((max(iters, 0), 3) if e.S == 1 else (max(iters, 0), e.S, 3), np.float64)
               for e in self.engines]
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
               check_finite=True, pack=True, **common):
    """Run ``algorithm`` (0 = ERA, 1 = DM) on every scan of ``scans`` together.

    ``scans``: list of dicts of ERA / DM arguments (``I``, ``R``, ``P``, ``O``,
    optional keywords); ``common`` keywords apply to every scan.  ``pack``
    (default) packs scans with equal ``pack_key`` into one engine each.
    ``tune=True`` lets each unpacked engine pick its fastest launch
    configuration first (``pty_tune``).  Each scan's ``R`` is shifted in place
    as ``ERA`` does."""
    kws = []
    for sc in scans:
        kw = dict(common)
        kw.update(sc)
        kws.append(kw)
    groups = {}
    if pack:
        for i, kw in enumerate(kws):
            key = pack_key(algorithm, kw)
            if key is not None:
                groups.setdefault(key, []).append(i)
    packed = [g for g in groups.values() if len(g) > 1]
    in_pack = {i for g in packed for i in g}
    jobs, packs = {}, []
    try:
        for g in packed:
            packs.append((g, _Pack(algorithm, [kws[i] for i in g], iters, device)))
        for i, kw in enumerate(kws):
            if i in in_pack:
                continue
            kw = dict(kw)
            I, R, P, O = kw.pop("I"), kw.pop("R"), kw.pop("P", None), kw.pop("O", None)
            jobs[i] = prepare(algorithm, I, R, P, O, iters, device=device, **kw)
        if tune:
            for j in jobs.values():
                j.eng.tune(algorithm, j.method, **j.run_kwargs())
        singles = sorted(jobs)
        engines = [pk.eng for _, pk in packs] + [jobs[i].eng for i in singles]
        opts = [pk.opts() for _, pk in packs] + [jobs[i].opts() for i in singles]
        batch = Batch(engines, graphs=graphs)
        try:
            metrics = batch.run(opts, iters, begin=True)
            status = batch.graph_status()
        finally:
            batch.close()
    except BaseException:
        for j in jobs.values():
            j.eng.close()
        for _, pk in packs:
            pk.eng.close()
        raise
    per_scan = {}
    for (g, pk), m in zip(packs, metrics):
        for s, i in enumerate(g):
            per_scan[i] = (pk.job(s), m[:, s, :])
    for i, m in zip(singles, metrics[len(packs):]):
        per_scan[i] = (jobs[i], m)
    out = []
    for i in range(len(kws)):
        j, m = per_scan[i]
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
