"""Single-step (teacher-forced) parity of every ERA / DM method.

Multi-iteration comparisons mix two things: the kernels' own error in one
step and the growth of round-off over many steps (DM is not a contraction and
amplifies any perturbation, the reference's own complex64 mode included).
This test separates them.  The complex128 oracle runs 20 iterations and
records the state it carries into each one: O, P, the cached heatmaps (the
reference's ``P_heatmap`` / ``O_heatmap``, ptychography/era.py:143-189), the
per-frame background, and for DM the iterate f and the eCon reference
(ptychography/dm.py:35-38,158-160).  For every k < 20 the state entering
iteration k, rounded to complex64 / float32, goes into the engine through the
C ABI (``pty_set_state``, ``pty_set_heatmap``, ``pty_set_iterate``,
``pty_set_background``); the engine runs iteration k alone, and the oracle
runs iteration k from the same rounded state in complex128.

Bar per step (stated here; the north star asks for 1e-5 on the iterates over
the first 20 iterations and 1 % on the final metrics):

* O on the illuminated region (sum |P|^2 >= 1e-6 of its maximum), P, f and
  the per-frame background: relative L2 error <= STEP_TOL;
* eMod and eSup: relative error <= METRIC_TOL; eCon (a difference of
  successive iterates, so its relative error is the iterate error over eCon)
  within METRIC_TOL relative or 3 STEP_TOL absolute.
"""

import numpy as np
import pytest
import torch

import oracle
from ptychography_amd import era as pera
from tests.scans import broadband_start, rel_l2, small_scan

pytestmark = pytest.mark.gpu

STEP_TOL = 2e-6
METRIC_TOL = 1e-5
ITERS = 20

CASES = [
    # (algorithm, method, OP_iters, probe_centering, Pmod_probe)
    ("ERA", 1, 1, False, False),
    ("ERA", 1, 1, False, 3),
    ("ERA", 2, 1, False, False),
    ("ERA", 3, 1, False, False),
    ("ERA", 3, (2, 3), True, 3),
    ("ERA", 4, 1, False, False),
    ("ERA", 5, 1, False, 2),
    ("ERA", 6, (1, 2), True, False),
    ("DM", 1, 1, False, False),
    ("DM", 2, 1, False, False),
    ("DM", 3, 1, False, False),
    ("DM", 3, (1, 3), True, 3),
    ("DM", 4, 1, False, False),
    ("DM", 5, 1, False, 2),
    ("DM", 6, (1, 2), True, False),
    ("DM", 6, 1, False, False),
]


def _scan(bg):
    kw = dict(frame=32, grid=(5, 5), step=6, jitter=1, beamstop_radius=2.0, dead_frac=0.05)
    if bg:
        kw["background_frac"] = 0.02
    return small_scan(**kw)


def _c64(x):
    return None if x is None else np.asarray(x).astype(np.complex64)


def _f32(x):
    return None if x is None else np.asarray(x).astype(np.float32)


def _t(x, dev):
    return None if x is None else torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("alg,method,OP,centre,pmodp", CASES,
                         ids=[f"{c[0]}{c[1]}-OP{c[2]}-c{int(c[3])}-p{int(c[4])}" for c in CASES])
def test_single_step_parity(gpu_device, alg, method, OP, centre, pmodp):
    print(f"single-step parity {alg} method {method} OP {OP} centring {centre} "
          f"Pmod_probe {pmodp}", flush=True)
    bg = method >= 4
    s = _scan(bg)
    I = s.I.astype(np.float32).astype(np.float64)        # what the engine sees
    rng = np.random.default_rng(21)
    P0 = s.P * (1.0 + 0.2 * rng.standard_normal(s.P.shape))
    O0 = broadband_start(s.O.shape)
    B0 = np.full(I.shape[1:], 0.5 * I.mean()) if bg else None
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    kw = dict(OP_iters=OP, mask=s.mask, background=B0, method=method, Pmod_probe=pmodp,
              probe_centering=centre)

    # the complex128 trajectory and the state entering each iteration
    states = []
    fn(I, s.R.copy(), P0.copy(), O0.copy(), ITERS, state_callback=lambda i, st: states.append(st),
       **kw)
    Rs = s.R - s.R.max(axis=0)
    f0 = oracle.make_exits(O0, P0, Rs) if alg == "DM" else None
    b0 = None
    if bg:
        b0 = np.broadcast_to(np.sqrt(np.fft.ifftshift(B0)), I.shape).copy()
    start = dict(O=O0, P=P0, exits=f0, background=b0, P_heatmap=None, O_heatmap=None,
                 O_ref=O0, P_ref=P0)
    entering = [start] + states[:-1]

    algorithm = pera.ALGO_ERA if alg == "ERA" else pera.ALGO_DM
    job = pera.prepare(algorithm, s.I, s.R.copy(), P0.copy(), O0.copy(), ITERS, OP_iters=OP,
                       mask=s.mask, background=B0, method=method, Pmod_probe=pmodp,
                       probe_centering=centre, device=gpu_device)
    eng = job.eng
    dev = gpu_device
    Pamp = np.abs(np.fft.fftn(P0.astype(np.complex64).astype(np.complex128)))
    region = oracle.make_P_heatmap(s.P, Rs, s.O.shape)
    region = region >= 1e-6 * region.max()
    worst = dict(O=0.0, P=0.0, f=0.0, b=0.0, eMod=0.0, eSup=0.0, eCon=0.0)
    try:
        for k, st in enumerate(entering):
            O32, P32 = _c64(st["O"]), _c64(st["P"])
            job.O_t.copy_(_t(O32, dev))
            job.P_t.copy_(_t(P32, dev))
            eng.set_state(job.O_t, job.P_t)
            eng.set_heatmap(_t(st["P_heatmap"], dev), _t(st["O_heatmap"], dev))
            b32 = _f32(st["background"])
            if bg:
                eng.set_background(_t(b32, dev), per_frame=True)
            f32 = Or32 = Pr32 = None
            if alg == "DM":
                f32, Or32, Pr32 = _c64(st["exits"]), _c64(st["O_ref"]), _c64(st["P_ref"])
                eng.set_iterate(_t(f32, dev), _t(Or32, dev), _t(Pr32, dev))
            m = eng.run(algorithm, method, 1, iter0=k, begin=False, **job.run_kwargs())[0]
            Og = job.O_t.cpu().numpy()
            Pg = job.P_t.cpu().numpy()
            fg, bgs = eng.get_iterate(f=alg == "DM", background=bg)

            # the oracle's iteration k from the same rounded state
            got = []
            resume = dict(i0=k, P_heatmap=st["P_heatmap"], O_heatmap=st["O_heatmap"],
                          background=None if b32 is None else b32.astype(np.float64), Pamp=Pamp)
            if alg == "DM":
                resume.update(exits=f32.astype(np.complex128), O_ref=Or32.astype(np.complex128),
                              P_ref=Pr32.astype(np.complex128))
            out = fn(I, s.R.copy(), P32.astype(np.complex128), O32.astype(np.complex128), 1,
                     resume=resume, state_callback=lambda i, x: got.append(x), **kw)
            ref, info = got[0], out[-1]

            e = dict(O=rel_l2(Og[region], ref["O"][region]), P=rel_l2(Pg, ref["P"]))
            if alg == "DM":
                e["f"] = rel_l2(fg.cpu().numpy(), ref["exits"])
            if bg:
                e["b"] = rel_l2(bgs.cpu().numpy(), ref["background"])
            for key, col in (("eMod", 0), ("eCon", 1), ("eSup", 2)):
                r = info[key][0]
                d = abs(m[col] - r)
                e[key] = d / max(abs(r), 1e-300)
                if key == "eCon":
                    assert d <= max(METRIC_TOL * abs(r), 3 * STEP_TOL), \
                        f"step {k}: eCon {m[col]:.6e} vs {r:.6e}"
                else:
                    assert e[key] <= METRIC_TOL, f"step {k}: {key} {m[col]:.8e} vs {r:.8e}"
            for key in ("O", "P", "f", "b"):
                if key in e:
                    assert e[key] <= STEP_TOL, f"step {k}: {key} rel L2 {e[key]:.2e}"
            for key, v in e.items():
                worst[key] = max(worst[key], v)
    finally:
        eng.close()
    print("worst single-step errors:", " ".join(f"{k} {v:.1e}" for k, v in worst.items()),
          flush=True)
