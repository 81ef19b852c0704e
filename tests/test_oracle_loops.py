"""Pin the loop semantics of ``oracle.ERA`` / ``oracle.DM``.

Two kinds of check, both on the CPU:

1. Trajectories.  ``tests/loop_oracle.py`` restates the reference loops
   (ptychography/era.py:152-291, ptychography/dm.py:45-294) a second time
   with explicit per-pixel loops and a direct DFT, written from the reference
   text and sharing no code with ``oracle/``.  On tiny inputs both must give
   the same O, P, eMod, eCon, eSup, info['I'], info['heatmap'] and
   info['background'] after every iteration, for methods 1-6, the OP_iters
   schedule (including its stale-heatmap branch), probe centring, the
   Pmod_probe window and a mask.

2. Properties the mathematics fixes, chosen so that a plausible slip fails
   one of them:
   * ERA is alternating projections (or, for object + probe, a block
     coordinate descent) between the modulus set and the overlap constraint:
     eMod_{i+1} <= eSup_i <= eMod_i for methods 1-3 (this pins eSup's M*K
     factor and its use of the updated exits) and eMod non-increasing with
     eSup_i <= eMod_i for methods 4-6 (the background projection is the frame
     mean, orthogonal to the exit-wave part);
   * DM keeps exact data at a fixed point, with a known background too;
   * a centring step rolls P and O by -(rint(centre of mass of |P|^2) - shape // 2),
     with scipy's centroid;
   * the far-field probe constraint holds exactly on the iterations
     i < Pmod_probe and not after.
"""

import numpy as np
import pytest

import oracle
from tests import loop_oracle as lo


def _problem(seed, M=8, K=8, N=4, masked=False, bg=False, offcentre=False, spread=4):
    rng = np.random.default_rng(seed)
    R = rng.integers(0, spread + 1, size=(N, 2)).astype(np.int64)
    R[0] = (spread, spread)          # ptp = spread on both axes
    U, V = M + spread, K + spread
    I = rng.random((N, M, K)) * 3.0 + 0.2
    P = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    if offcentre:
        y, x = np.mgrid[:M, :K]
        P = P * (0.25 + np.exp(-((y - 2.0) ** 2 + (x - 5.0) ** 2) / 4.0))
    O = 1.0 + 0.3 * (rng.standard_normal((U, V)) + 1j * rng.standard_normal((U, V)))
    mask = 1
    if masked:
        mask = rng.random((M, K)) > 0.2
    B = rng.random((M, K)) + 0.2 if bg else None
    return dict(I=I, R=R, P=P, O=O, mask=mask, background=B)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _run_oracle(alg, pb, iters, **kw):
    traj = []
    cb = lambda i, O, P: traj.append((O.copy(), P.copy()))
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    out = fn(pb["I"], pb["R"].copy(), pb["P"].copy(), pb["O"].copy(), iters, mask=pb["mask"],
             background=pb["background"], callback=cb, **kw)
    info = out[-1]
    return traj, info


def _run_loops(alg, pb, iters, **kw):
    traj = []
    cb = lambda i, O, P: traj.append((O.copy(), P.copy()))
    fn = lo.ERA if alg == "ERA" else lo.DM
    R = [list(r) for r in pb["R"]]
    _, _, info = fn(pb["I"], R, pb["P"].copy(), pb["O"].copy(), iters, mask=pb["mask"],
                    background=pb["background"], callback=cb, **kw)
    return traj, info


CASES = [
    # (algorithm, method, OP_iters, probe_centering, Pmod_probe, masked, offcentre)
    ("ERA", 1, 1, False, False, True, False),
    ("ERA", 2, 1, False, False, True, False),
    ("ERA", 3, 1, False, False, True, False),
    ("ERA", 3, (2, 3), False, False, True, False),       # stale-heatmap branch
    ("ERA", 3, (1, 2), True, 2, True, True),             # centring + Pmod_P window
    ("ERA", 1, 1, False, 3, True, False),                # Pmod_P with a cached heatmap
    ("ERA", 2, 1, False, 2, False, False),
    ("ERA", 4, 1, False, False, True, False),
    ("ERA", 5, 1, False, 2, True, False),
    ("ERA", 6, (2, 3), True, 2, True, True),
    ("DM", 1, 1, False, False, True, False),
    ("DM", 2, 1, False, False, True, False),
    ("DM", 3, 1, False, False, True, False),
    ("DM", 3, (2, 3), False, 3, True, False),
    ("DM", 3, (1, 2), True, 2, True, True),
    ("DM", 1, 1, False, 2, False, False),
    ("DM", 4, 1, False, False, True, False),
    ("DM", 5, 1, False, 2, True, False),
    ("DM", 6, (1, 2), True, 2, True, True),
    ("DM", 6, (2, 3), False, False, False, False),
]


@pytest.mark.parametrize("alg,method,OP,centre,pmodp,masked,offc", CASES)
def test_oracle_matches_pure_loop_restatement(alg, method, OP, centre, pmodp, masked, offc):
    pb = _problem(seed=100 + method + 10 * (alg == "DM"), masked=masked, bg=method >= 4,
                  offcentre=offc)
    iters = 6
    kw = dict(OP_iters=OP, method=method, Pmod_probe=pmodp, probe_centering=centre)
    t_o, i_o = _run_oracle(alg, pb, iters, **kw)
    t_l, i_l = _run_loops(alg, pb, iters, **kw)
    assert len(t_o) == len(t_l) == iters
    for k, ((Oo, Po), (Ol, Pl)) in enumerate(zip(t_o, t_l)):
        assert _rel(Oo, Ol) < 1e-11, f"O differs at iteration {k}: {_rel(Oo, Ol):.2e}"
        assert _rel(Po, Pl) < 1e-11, f"P differs at iteration {k}: {_rel(Po, Pl):.2e}"
    for key in ("eMod", "eCon", "eSup"):
        a, b = np.array(i_o[key]), np.array(i_l[key])
        assert np.allclose(a, b, rtol=1e-10, atol=1e-14), f"{key}: {a} vs {b}"
    assert _rel(i_o["I"], i_l["I"]) < 1e-11
    if i_l["heatmap"] is None:
        assert i_o["heatmap"] is None
    else:
        assert _rel(i_o["heatmap"], i_l["heatmap"]) < 1e-12
    if method >= 4:
        assert _rel(i_o["background"], i_l["background"]) < 1e-11


def test_loop_dft_is_the_numpy_transform():
    # the two checkers share no transform code: pin the loop DFT's sign and scaling
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 8)) + 1j * rng.standard_normal((6, 8))
    assert np.allclose(np.array(lo.dft2(x)), np.fft.fftn(x), atol=1e-12)
    assert np.allclose(np.array(lo.dft2(x, inverse=True)), np.fft.ifftn(x), atol=1e-13)
    assert np.array_equal(lo.ifftshift2(x), np.fft.ifftshift(x))
    assert np.array_equal(lo.fftshift2(x), np.fft.fftshift(x))
    assert np.array_equal(lo.roll2(x, 2, -3), np.roll(x, (2, -3), axis=(0, 1)))


# ---------------------------------------------------------------------------
# properties fixed by the mathematics
# ---------------------------------------------------------------------------

def _scan(bg=False, poisson=True):
    from tests.scans import small_scan
    kw = dict(frame=32, grid=(4, 4), step=6, poisson=poisson)
    if bg:
        kw["background_frac"] = 0.05
    return small_scan(**kw)


def _start(s, seed=13):
    from tests.scans import broadband_start
    return broadband_start(s.O.shape, seed=seed, eps=0.1)


@pytest.mark.parametrize("method", [1, 2, 3])
def test_era_sandwich_emod_esup(method):
    s = _scan()
    rng = np.random.default_rng(4)
    P0 = s.P * (1.0 + 0.2 * rng.standard_normal(s.P.shape))
    O0 = _start(s)
    mask = np.ones(s.I.shape[1:], bool)
    mask[rng.random(mask.shape) < 0.1] = False          # masked pixels are left free
    out = oracle.ERA(s.I, s.R.copy(), P0, O0, 15, mask=mask, method=method)
    info = out[-1]
    e, sup = np.array(info["eMod"]), np.array(info["eSup"])
    tol = 1e-9
    # the projection onto the overlap constraint moves no further than the last
    # modulus projection did, and the next modulus error is at most that move
    assert np.all(sup <= e * (1 + tol)), (sup, e)
    assert np.all(e[1:] <= sup[:-1] * (1 + tol)), (e, sup)
    assert e[-1] < e[0]


@pytest.mark.parametrize("method", [4, 5, 6])
def test_era_background_emod_non_increasing(method):
    s = _scan(bg=True)
    rng = np.random.default_rng(5)
    P0 = s.P * (1.0 + 0.2 * rng.standard_normal(s.P.shape))
    O0 = _start(s)
    B0 = np.full(s.I.shape[1:], 0.5 * s.I.mean())
    out = oracle.ERA(s.I, s.R.copy(), P0, O0, 15, background=B0, method=method)
    info = out[-1]
    e, sup = np.array(info["eMod"]), np.array(info["eSup"])
    assert np.all(np.diff(e) <= 1e-9 * e[:-1]), e
    assert np.all(sup <= e * (1 + 1e-9)), (sup, e)
    assert e[-1] < e[0]


@pytest.mark.parametrize("method", [3, 4])
def test_dm_exact_data_fixed_point_with_probe_and_background(method):
    s = _scan(poisson=False)
    I = s.I.copy()
    B = None
    if method == 4:
        rng = np.random.default_rng(6)
        B = 0.02 * I.mean() * (1.0 + rng.random(I.shape[1:]))
        I = I + B[None]
    out = oracle.DM(I, s.R.copy(), s.P.copy(), s.O.copy(), 4, background=B, method=method)
    info = out[-1]
    O = out[0]
    H = info["heatmap"]
    assert max(info["eMod"]) < 1e-6
    assert max(info["eSup"]) < 1e-6
    if method == 3:
        P = out[1]
        assert np.allclose(P, s.P, rtol=1e-6, atol=1e-9 * np.abs(s.P).max())
        H = oracle.make_P_heatmap(s.P, s.R - s.R.max(axis=0), s.O.shape)
    cov = H > 1e-6 * H.max()
    assert np.allclose(O[cov], s.O[cov], rtol=1e-6, atol=1e-8)
    if method == 4:
        assert np.allclose(info["background"], np.fft.fftshift(np.fft.ifftshift(B)), rtol=1e-6)


@pytest.mark.parametrize("alg", ["ERA", "DM"])
def test_centring_step_rolls_by_the_rounded_centroid(alg):
    import scipy.ndimage
    s = _scan()
    M = s.P.shape[0]
    P0 = np.roll(s.P, (5, -3), axis=(0, 1))           # a displaced probe
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    runs = []
    for centre in (False, True):
        got = []
        cb = lambda i, O, P: got.append((O.copy(), P.copy()))
        fn(s.I, s.R.copy(), P0, _start(s), 1, method=3, probe_centering=centre, callback=cb)
        runs.append(got[0])
    (Oa, Pa), (Ob, Pb) = runs
    # scipy's centroid, rounded half to even, relative to shape // 2
    cm = np.rint(scipy.ndimage.center_of_mass(np.abs(Pa) ** 2)).astype(int) - M // 2
    assert np.any(cm != 0)
    assert np.array_equal(Pb, np.roll(Pa, tuple(-cm), axis=(0, 1)))
    assert np.array_equal(Ob, np.roll(Oa, tuple(-cm), axis=(0, 1)))


@pytest.mark.parametrize("alg", ["ERA", "DM"])
def test_pmod_probe_window(alg):
    s = _scan()
    rng = np.random.default_rng(7)
    P0 = s.P * (1.0 + 0.3 * rng.standard_normal(s.P.shape))
    Pamp = np.abs(np.fft.fftn(P0))
    Ps = []
    cb = lambda i, O, P: Ps.append(P.copy())
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    fn(s.I, s.R.copy(), P0, _start(s), 4, method=3, Pmod_probe=2, callback=cb)
    for i, P in enumerate(Ps):
        dev = np.abs(np.abs(np.fft.fftn(P)) - Pamp).max() / Pamp.max()
        if i < 2:
            assert dev < 1e-9, (i, dev)       # constraint applied on i < Pmod_probe
        else:
            assert dev > 1e-4, (i, dev)       # and not after


# ---------------------------------------------------------------------------
# planted mistakes (in the oracle, on the CPU): each must break the comparison
# ---------------------------------------------------------------------------

def _max_traj_diff(alg, pb, iters, **kw):
    t_o, i_o = _run_oracle(alg, pb, iters, **kw)
    t_l, i_l = _run_loops(alg, pb, iters, **kw)
    d = max(max(_rel(a[0], b[0]), _rel(a[1], b[1])) for a, b in zip(t_o, t_l))
    for key in ("eMod", "eSup"):
        d = max(d, _rel(i_o[key], i_l[key]))
    return d


def _planted_pmod7_keeps_masked(mp):
    import oracle.era as oe
    good = oe.Pmod_7

    def bad(amp, background, exits, mask=1, alpha=1e-10, eMod_calc=False):
        Mg = np.sqrt((exits.conj() * exits).real + background ** 2 + alpha)
        eMod = np.sum((Mg - amp) ** 2 * mask) if eMod_calc else None
        fac = np.where(mask, amp / Mg, 1.0)          # masked pixels kept, as in Pmod_1
        exits *= fac
        background *= fac
        return exits, background, eMod
    mp.setattr(oe, "Pmod_7", bad)
    return good


def _planted_fresh_heatmap(mp):
    import oracle.era as oe
    good = oe.psup_O
    mp.setattr(oe, "psup_O", lambda ex, P, R, shape, H=None, alpha=1e-10: good(ex, P, R, shape, None, alpha))


def _planted_probe_window_transposed(mp):
    import oracle.era as oe
    good = oe.psup_P

    def bad(exits, O, R, O_heatmap=None, alpha=1e-10):
        return good(exits, O, np.asarray(R)[:, ::-1], O_heatmap, alpha)
    mp.setattr(oe, "psup_P", bad)


def _planted_centre_sign(mp):
    import oracle.era as oe
    good = oe.probe_centre_shift
    mp.setattr(oe, "probe_centre_shift", lambda P: -good(P))


PLANTED = [
    (_planted_pmod7_keeps_masked, ("DM", 4, 1, False, False, True, False)),
    (_planted_fresh_heatmap, ("ERA", 3, (2, 3), False, False, True, False)),
    (_planted_probe_window_transposed, ("ERA", 3, 1, False, False, True, False)),
    (_planted_centre_sign, ("DM", 3, (1, 2), True, 2, True, True)),
]


@pytest.mark.parametrize("plant,case", PLANTED, ids=[p[0].__name__ for p in PLANTED])
def test_planted_oracle_mistakes_are_caught(monkeypatch, plant, case):
    alg, method, OP, centre, pmodp, masked, offc = case
    pb = _problem(seed=100 + method + 10 * (alg == "DM"), masked=masked, bg=method >= 4,
                  offcentre=offc)
    kw = dict(OP_iters=OP, method=method, Pmod_probe=pmodp, probe_centering=centre)
    assert _max_traj_diff(alg, pb, 6, **kw) < 1e-10
    plant(monkeypatch)
    assert _max_traj_diff(alg, pb, 6, **kw) > 1e-6


# ---------------------------------------------------------------------------
# the oracle's resume hook (used by the single-step GPU parity tests)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", [c for c in CASES if c[2] != 1 or c[1] in (1, 4)],
                         ids=lambda c: f"{c[0]}{c[1]}-{c[2]}-{c[3]}-{c[4]}")
def test_oracle_resume_reproduces_the_run(case):
    alg, method, OP, centre, pmodp, masked, offc = case
    pb = _problem(seed=200 + method + 10 * (alg == "DM"), masked=masked, bg=method >= 4,
                  offcentre=offc)
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    kw = dict(OP_iters=OP, method=method, Pmod_probe=pmodp, probe_centering=centre,
              mask=pb["mask"], background=pb["background"])
    states = []
    full = fn(pb["I"], pb["R"].copy(), pb["P"].copy(), pb["O"].copy(), 6,
              state_callback=lambda i, st: states.append(st), **kw)
    st = states[2]
    resume = dict(i0=3, P_heatmap=st["P_heatmap"], O_heatmap=st["O_heatmap"],
                  background=st["background"])
    if alg == "DM":
        resume.update(exits=st["exits"], O_ref=st["O_ref"], P_ref=st["P_ref"])
    part = fn(pb["I"], pb["R"].copy(), st["P"], st["O"], 3, resume=resume, **kw)
    for a, b in zip(full[:-1], part[:-1]):
        assert np.array_equal(a, b)
    for key in ("eMod", "eCon", "eSup"):
        assert full[-1][key][3:] == part[-1][key]
