"""``ERA`` -- error-reduction ptychography on one MI355X.

Drop-in for ``Ptychography.ERA`` (ptychography/era.py:13-333): same positional
and keyword arguments, same returns ``(O, info)`` / ``(P, info)`` /
``(O, P, info)`` (or without ``info`` when ``full_output=False``), same side
effect on the caller's ``R`` (shifted in place so each column's maximum is 0,
ptychography/era.py:504-505).  ``info`` holds ``'I'``, ``'eMod'``, ``'eCon'``,
``'heatmap'`` and ``'background'`` as in the reference, plus ``'eSup'``.

All iteration arithmetic runs in complex64 in the gfx950 kernels of
``libptyhip.so``; the error sums and the object normalisation use FP64.  The
returned arrays have the precision the reference would use (complex128 for
float64 intensities, complex64 for float32).
"""

import math

import numpy as np
import torch

from .engine import Engine, require_gpu

__all__ = ["ERA", "solve", "update_progress"]

ALGO_ERA = 0
ALGO_DM = 1


def resolve_method(O, P, background, method):
    """Method and update target as the reference's preamble decides (era.py:440-456)."""
    if method is None:
        if O is None and P is None:
            method = 3
        elif O is None:
            method = 1
        elif P is None:
            method = 2
        else:
            raise ValueError("method=None with both O and P given is undefined in the reference "
                             "(its 'update' is never set); pass method explicitly")
        if background is not None:
            method += 3
    method = int(method)
    if method not in (1, 2, 3, 4, 5, 6):
        raise ValueError("method must be one of 1..6")
    update = {1: 'O', 4: 'O', 2: 'P', 5: 'P', 3: 'OP', 6: 'OP'}[method]
    return method, update


def pmod_probe_until(Pmod_probe):
    """Iterations i with ``Pmod_probe is not None and i < Pmod_probe`` are [0, until)."""
    if Pmod_probe is None or Pmod_probe is False:
        return 0
    p = float(Pmod_probe)
    if math.isinf(p):
        return 1 << 62 if p > 0 else 0
    return max(0, int(math.ceil(p)))


def _as_numpy(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _random(rng, shape):
    return np.random.random(shape) if rng is None else rng.random(shape)


def result_dtypes(I, dtype):
    """(real, complex) numpy dtypes of the returns: the reference's precision
    rule (float32 intensities -> complex64, else complex128) or ``dtype``."""
    if dtype is None:
        single = (I.dtype == torch.float32) if isinstance(I, torch.Tensor) else \
            (np.asarray(I).dtype == np.float32)
        dtype = 'single' if single else 'double'
    if dtype == 'single':
        return np.float32, np.complex64
    if dtype == 'double':
        return np.float64, np.complex128
    raise ValueError("dtype must be None, 'single' or 'double'")


def start_state(R, P, O, M, K, c_dtype, rng):
    """Starting O and P as the preamble makes them (era.py:475-493): a flat
    object covering every window, a random probe."""
    if O is None:
        shape = (M + int(np.ceil(R[:, 0].max() - R[:, 0].min())),
                 K + int(np.ceil(R[:, 1].max() - R[:, 1].min())))
        O = np.ones(shape, dtype=c_dtype)
    if P is None:
        P = _random(rng, (M, K)) + 1J * _random(rng, (M, K))
    O = _as_numpy(O).astype(c_dtype)
    P = _as_numpy(P).astype(c_dtype)
    if P.shape != (M, K):
        raise ValueError(f"P has shape {P.shape}, frames are {(M, K)}")
    return O, P


def mask_array(mask, M, K):
    """The centred good-pixel mask as uint8 (None for the reference's mask=1)."""
    if np.isscalar(mask) and mask == 1:
        return None
    m = _as_numpy(mask)
    if m.shape != (M, K):
        raise ValueError(f"mask must be 1 or an array of shape {(M, K)}")
    if not np.all((m == 0) | (m == 1)):
        raise ValueError("mask values must be 0/1 (True = good pixel)")
    return m.astype(np.uint8)


class Job:
    """One prepared scan: the engine with data and state bound, and what the
    returns need (``prepare`` / ``finish``)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def run_kwargs(self):
        return dict(OP_iters=self.OP_iters, Pmod_probe_until=self.until,
                    probe_centering=self.probe_centering, alpha=self.alpha)

    def opts(self):
        return Engine.opts(self.algorithm, self.method, **self.run_kwargs())


def prepare(algorithm, I, R, P, O, iters, OP_iters=1, mask=1, background=None, method=None,
            Pmod_probe=False, probe_centering=False, alpha=1.0e-10, dtype=None, device=None,
            rng=None, init=None, init_net=None, graphs=False, refine_positions=0):
    """Host preamble of ERA / DM (era.py:435-548) on a new engine; returns a ``Job``."""
    dev = require_gpu(device)
    if algorithm not in (ALGO_ERA, ALGO_DM):
        raise ValueError("algorithm must be 0 (ERA) or 1 (DM)")
    if iters < 0:
        raise ValueError("iters must be >= 0")
    method, update = resolve_method(O, P, background, method)
    if init not in (None, 'flat', 'learned'):
        raise ValueError("init must be None, 'flat' or 'learned'")
    if init == 'learned' and init_net is None:
        raise ValueError("init='learned' needs init_net (a trained learned.InitNet)")
    if init == 'learned' and update == 'P':
        raise ValueError("init='learned' seeds O, but this method updates only P")
    if isinstance(OP_iters, (int, np.integer)):
        OP_iters = (int(OP_iters), 1)
    OP_iters = (int(OP_iters[0]), int(OP_iters[1]))

    I_is_tensor = isinstance(I, torch.Tensor)
    N, M, K = (int(s) for s in I.shape)
    r_dtype, c_dtype = result_dtypes(I, dtype)

    if not isinstance(R, np.ndarray):
        raise TypeError("R must be a numpy array (it is shifted in place, era.py:504-505)")
    # float positions with fractional parts: integer windows plus per-frame
    # Fourier-shifted probes (object-only methods, no far-field probe constraint)
    refine_positions = int(refine_positions or 0)
    if refine_positions < 0:
        raise ValueError("refine_positions must be >= 0")
    # refining positions runs the sub-pixel model (from zero shifts for integer R)
    frac = is_fractional(R) or refine_positions > 0
    what = "position refinement" if refine_positions else "sub-pixel positions"
    if frac and method not in (1, 4):
        raise ValueError(f"{what}: modelled for the object-only methods 1 and 4")
    if frac and pmod_probe_until(Pmod_probe) > 0:
        raise ValueError(f"{what}: modelled without the far-field probe constraint")
    if frac and init == 'learned':
        raise ValueError(f"{what}: init='learned' stitches patches at integer positions only")
    O, P = start_state(R, P, O, M, K, c_dtype, rng)

    # in-place shift of the caller's R (era.py:504-505)
    R[:, 0] -= R[:, 0].max()
    R[:, 1] -= R[:, 1].max()

    U, V = O.shape
    R_win, d = split_positions(R) if frac else (R, None)
    eng = Engine(N, M, K, U, V, R_win, device=dev)
    try:
        mask_c = None
        mask_s = None
        mc = mask_array(mask, M, K)
        if mc is not None:
            mask_c = torch.from_numpy(np.ascontiguousarray(mc)).to(dev)
            mask_s = torch.from_numpy(np.ascontiguousarray(np.fft.ifftshift(mc))).to(dev)
        I_t = I.to(dev) if I_is_tensor else torch.from_numpy(np.ascontiguousarray(I, dtype=np.float32))
        amp, I_norm = eng.prepare_amp(I_t, mask_c)
        eng.set_data(amp, mask_s, I_norm)

        O_t = torch.from_numpy(np.ascontiguousarray(O.astype(np.complex64))).to(dev)
        P_t = torch.from_numpy(np.ascontiguousarray(P.astype(np.complex64))).to(dev)
        init_info = None
        if init == 'learned' and (M, K) != (init_net.frame, init_net.frame):
            init_info = dict(used_learned=False,
                             reason=f"the network takes {init_net.frame}x{init_net.frame} frames, "
                                    f"the scan has {M}x{K}")
        elif init == 'learned':
            from .learned import learned_start
            O_t, init_info = learned_start(init_net, eng, I_t.to(dev, torch.float32).contiguous(),
                                           P_t, O_t, alpha=alpha)
        del I_t
        eng.set_state(O_t, P_t)
        if frac:
            eng.set_subpixel(d)

        bg_init = None
        if method >= 4:
            if background is None:
                b = (_random(rng, (N, M, K)).astype(r_dtype) + 0.1).astype(np.float32)
                eng.set_background(torch.from_numpy(b), per_frame=True)
            else:
                B = _as_numpy(background)
                b = np.sqrt(np.fft.ifftshift(B)).astype(np.float32)  # every axis, as era.py:543
                if b.ndim == 2:
                    eng.set_background(torch.from_numpy(np.ascontiguousarray(b)), per_frame=False)
                else:
                    eng.set_background(torch.from_numpy(np.ascontiguousarray(
                        np.broadcast_to(b, (N, M, K)))), per_frame=True)
            bg_init = b[0] if b.ndim == 3 else b

        until = pmod_probe_until(Pmod_probe)
        if until > 0:
            eng.set_pamp(None)   # |FFT(P0)| (era.py:546-547)
        if graphs:
            eng.set_graphs(True)
    except BaseException:
        eng.close()
        raise
    return Job(algorithm=algorithm, eng=eng, O_t=O_t, P_t=P_t, method=method, update=update,
               r_dtype=r_dtype, c_dtype=c_dtype, bg_init=bg_init, background=background,
               OP_iters=OP_iters, until=until, probe_centering=probe_centering, alpha=alpha,
               init_info=init_info, iters=iters, graphs=bool(graphs), R_win=R_win, d=d,
               refine=refine_positions)


def is_fractional(R):
    """True for float positions with a non-integer entry."""
    return np.issubdtype(R.dtype, np.floating) and bool(np.any(R != np.rint(R)))


def position_step(terms, d, damping=1e-3, max_step=0.5, max_shift=1.0):
    """One Gauss-Newton step on every frame's shift from its position terms
    (``Engine.position_terms``: f, gradient (y, x), matrix (yy, yx, xx)):
    delta = -(H + damping tr(H) I)^-1 g, each component clipped to
    +-max_step, the new shift to +-max_shift.  With the object fixed the
    modulus error is a sum of per-frame terms, so the frames step
    independently."""
    T = np.asarray(terms, dtype=np.float64)
    g = T[:, 1:3]
    H = np.empty((len(T), 2, 2))
    H[:, 0, 0], H[:, 0, 1], H[:, 1, 0], H[:, 1, 1] = T[:, 3], T[:, 4], T[:, 4], T[:, 5]
    tr = T[:, 3] + T[:, 5]
    ok = tr > 0
    delta = np.zeros_like(g)
    if np.any(ok):
        A = H[ok] + (damping * tr[ok])[:, None, None] * np.eye(2)
        delta[ok] = -np.linalg.solve(A, g[ok][..., None])[..., 0]
    delta = np.clip(delta, -max_step, max_step)
    return np.clip(np.asarray(d, dtype=np.float64) + delta, -max_shift, max_shift)


def split_positions(R):
    """Windows and remainders of positions R <= 0 (after the in-place shift):
    frame n's window starts at s = -R[n]; w = rint(s) is its integer window
    (returned as the int32 position -w) and d = s - w in [-1/2, 1/2] the shift
    of the probe it sees."""
    s = -np.asarray(R, dtype=np.float64)
    w = np.rint(s)
    return (-w).astype(np.int32), s - w


def _round_half_away(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def update_progress(progress, algorithm, i, emod, esup, stream=None):
    """The reference's one-line progress bar (ptychography/era.py:612-629),
    written to ``stream`` (default ``sys.stdout``) and flushed.

    ``"\r{algorithm}: [{15-character bar}] {percent}% {i} {emod} {esup} {status}"``
    plus five spaces; status is "Done...\r\n" at progress >= 1, "Halt...\r\n"
    below 0, and an error note for a non-number.  The reference runs under
    Python 2, so the bar length rounds half away from zero and floats print
    with 12 significant digits, as Python 2's ``str`` does.  ERA / DM pass
    (i / max(1, iters - 1), name, i, eCon, eMod) (era.py:212, dm.py:152).
    """
    import sys
    bar = 15
    status = ""
    if isinstance(progress, int):
        progress = float(progress)
    if not isinstance(progress, float):
        progress = 0
        status = "error: progress var must be float\r\n"
    if progress < 0:
        progress = 0
        status = "Halt...\r\n"
    if progress >= 1:
        progress = 1
        status = "Done...\r\n"
    block = _round_half_away(bar * progress)

    def fmt(v):
        return "%.12g" % v if isinstance(v, (float, np.floating)) else str(v)
    text = "\r{0}: [{1}] {2}% {3} {4} {5} {6} {7}".format(
        algorithm, "#" * block + "-" * (bar - block), int(progress * 100), i, fmt(emod),
        fmt(esup), status, " " * 5)
    out = sys.stdout if stream is None else stream
    out.write(text)
    out.flush()
    return text


def check_metrics(metrics):
    """Raise if an iteration's eMod is not finite (NaN / inf in the data or the
    state): the reference would carry the NaNs on silently (era.py:561-571)."""
    e = np.asarray(metrics)[:, 0] if len(metrics) else np.zeros(0)
    bad = np.nonzero(~np.isfinite(e))[0]
    if len(bad):
        raise FloatingPointError(f"eMod is not finite from iteration {int(bad[0])} on: the "
                                 "intensities or the state hold NaN / inf")


def finish(job, metrics, full_output=True, check_finite=True):
    """Returns of ERA / DM (era.py:315-332) from a run job; closes its engine."""
    eng = job.eng
    try:
        if check_finite:
            check_metrics(metrics)
        c_dtype, r_dtype = job.c_dtype, job.r_dtype
        O_out = job.O_t.cpu().numpy().astype(c_dtype)
        P_out = job.P_t.cpu().numpy().astype(c_dtype)
        update = job.update
        if not full_output:
            return {'O': O_out, 'P': P_out}[update] if update != 'OP' else (O_out, P_out)

        info = {}
        if job.init_info is not None:
            info['init'] = job.init_info
        if job.graphs:
            en, fb = eng.graph_status()
            info['graphs'] = {'enabled': en, 'fell_back': fb}
        info['I'] = eng.intensity().cpu().numpy().astype(r_dtype)
        info['eMod'] = [float(x) for x in metrics[:, 0]]
        info['eCon'] = [float(x) for x in metrics[:, 1]]
        info['eSup'] = [float(x) for x in metrics[:, 2]]
        H = eng.heatmap()
        info['heatmap'] = None if H is None else H.cpu().numpy().astype(r_dtype)
        if job.refine:
            info['R'] = np.asarray(job.R_win, np.float64) - job.d
        background = job.background
        if job.method >= 4:
            if job.iters > 0:
                b = eng.background().cpu().numpy().astype(r_dtype)
            else:
                b = np.asarray(job.bg_init, dtype=r_dtype)
            info['background'] = np.fft.fftshift(b) ** 2
        elif background is not None and job.algorithm == ALGO_ERA:
            # era.py:311-314 reports a passed background even when it is unused
            B = _as_numpy(background)
            info['background'] = np.fft.fftshift(B[0] if B.ndim == 3 else B) ** 2
        if update == 'O':
            return O_out, info
        if update == 'P':
            return P_out, info
        return O_out, P_out, info
    finally:
        eng.close()


def solve(algorithm, I, R, P, O, iters, OP_iters=1, mask=1, background=None, method=None,
          Pmod_probe=False, probe_centering=False, alpha=1.0e-10, dtype=None,
          full_output=True, device=None, callback=None, rng=None, init=None, init_net=None,
          graphs=False, check_finite=True, progress=False, refine_positions=0):
    """Shared body of ERA and DM: host preamble (era.py:435-548), device loop, returns.

    ``init='learned'`` (with ``init_net``, a trained ``learned.InitNet``) replaces
    the starting object by the network's stitched estimate when its
    first-iteration eMod beats that of the ``O`` passed in; ``info['init']``
    records both eMods, the choice and the time spent (or, for frames of
    another size than the network's, that the plain start was kept and why).

    ``refine_positions=k`` (object-only methods 1 and 4) refines the frame
    positions: after every k iterations each frame's shift takes one
    Gauss-Newton step on its modulus error with the object fixed
    (``position_step``), within +-1 pixel of the window its starting position
    rounds to; ``info['R']`` returns the refined positions (same convention
    as the shifted ``R``)."""
    job = prepare(algorithm, I, R, P, O, iters, OP_iters=OP_iters, mask=mask,
                  background=background, method=method, Pmod_probe=Pmod_probe,
                  probe_centering=probe_centering, alpha=alpha, dtype=dtype, device=device,
                  rng=rng, init=init, init_net=init_net, graphs=graphs,
                  refine_positions=refine_positions)
    try:
        eng, kw = job.eng, job.run_kwargs()
        name = 'ERA' if algorithm == ALGO_ERA else 'DM'
        if progress:
            print('algrithm progress iteration convergence modulus error')   # era.py:151
        per_iter = callback is not None or progress
        rows = []
        i = 0
        while i < iters:
            n = iters - i
            if job.refine:
                n = min(n, job.refine - i % job.refine)
            if per_iter:
                n = 1
            rows.append(eng.run(algorithm, job.method, n, iter0=i, begin=(i == 0), **kw))
            if progress:
                update_progress(i / max(1.0, float(iters - 1)), name, i, float(rows[-1][0, 1]),
                                float(rows[-1][0, 0]))
            if callback is not None:
                callback(i, job.O_t.cpu().numpy().astype(job.c_dtype),
                         job.P_t.cpu().numpy().astype(job.c_dtype))
            i += n
            if job.refine and i % job.refine == 0 and i < iters:
                job.d = position_step(eng.position_terms(), job.d)
                eng.set_subpixel(job.d)
        metrics = np.concatenate(rows, axis=0) if rows else np.zeros((0, 3))
        if iters == 0:
            eng.run(algorithm, job.method, 0, iter0=0, begin=True, **kw)
    except BaseException:
        job.eng.close()
        raise
    return finish(job, metrics, full_output, check_finite)


def ERA(I, R, P, O, iters, OP_iters=1, mask=1, background=None, method=None, Pmod_probe=False,
        probe_centering=False, hardware='cpu', alpha=1.0e-10, dtype=None, full_output=True,
        device=None, callback=None, rng=None, init=None, init_net=None, graphs=False,
        check_finite=True, progress=False, refine_positions=0):
    """Error-reduction algorithm (ptychography/era.py:13).  ``hardware`` is accepted and
    ignored, as in the reference; the work always runs on the GPU ``device``.

    Build-only keywords: ``device`` (torch device, default the current GPU),
    ``callback(i, O, P)`` after every iteration, ``rng`` (object with
    ``.random(shape)``) in place of ``np.random`` for the default P / background,
    ``init='learned'`` with ``init_net`` for the gated learned starting object,
    ``graphs=True`` to replay each kind of iteration as a captured hipGraph
    (bit-identical to launching the kernels one by one; a failed capture falls
    back to eager launches, reported in ``info['graphs']``),
    ``check_finite`` (default True) to raise ``FloatingPointError`` when eMod
    becomes NaN / inf instead of returning NaNs, ``progress=True`` to print
    the reference's header and per-iteration progress bar (``update_progress``;
    the reference always prints them, era.py:151,212), ``refine_positions=k``
    to refine the positions every k iterations (``solve``).  ``R`` may hold
    float positions (sub-pixel model, methods 1 and 4).
    """
    return solve(ALGO_ERA, I, R, P, O, iters, OP_iters=OP_iters, mask=mask,
                 background=background, method=method, Pmod_probe=Pmod_probe,
                 probe_centering=probe_centering, alpha=alpha, dtype=dtype,
                 full_output=full_output, device=device, callback=callback, rng=rng,
                 init=init, init_net=init_net, graphs=graphs, check_finite=check_finite,
                 progress=progress, refine_positions=refine_positions)
