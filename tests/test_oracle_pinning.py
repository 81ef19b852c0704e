"""Pin the NumPy oracle against brute force, closed forms and invariants.

The reference (Python 2 + mpi4py at import) cannot run here and holds no
golden vectors for this path, so each oracle function is checked against
something other than its own formula: explicit index loops, a direct O(n^2)
DFT, scipy, adjoint identities, fixed points and monotonicity.
"""

import numpy as np
import pytest
import scipy.ndimage

import oracle
from tests.scans import small_scan


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dft2(x):
    """Direct O(n^2) 2-D DFT over the last two axes (numpy sign convention)."""
    M, K = x.shape[-2:]
    FM = np.exp(-2j * np.pi * np.outer(np.arange(M), np.arange(M)) / M)
    FK = np.exp(-2j * np.pi * np.outer(np.arange(K), np.arange(K)) / K)
    return np.einsum("ij,...jk,lk->...il", FM, x, FK)


def _idft2(x):
    M, K = x.shape[-2:]
    return np.conj(_dft2(np.conj(x))) / (M * K)


def _tiny_geometry(rng, M=4, K=5, N=4, U=9, V=10):
    off = np.stack([rng.integers(0, U - M + 1, N), rng.integers(0, V - K + 1, N)], 1)
    off[0] = 0
    R = -off
    R -= R.max(axis=0)  # already max 0 (off[0] = 0)
    return R, U, V


# ---------------------------------------------------------------------------
def test_multiroll_equals_nested_roll_and_index_formula():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 7))
    for s in [(1, 3), (-2, 4), (0, -9), (13, 0)]:
        y = oracle.multiroll(x, s)
        assert np.array_equal(y, np.roll(np.roll(x, s[0], 0), s[1], 1))
        for i in range(5):
            for j in range(7):
                assert y[i, j] == x[(i - s[0]) % 5, (j - s[1]) % 7]
    # the docstring example of the reference (era.py:666-679)
    x = np.arange(20).reshape(4, 5)
    assert oracle.multiroll(x, [1, 3]).tolist() == [[17, 18, 19, 15, 16], [2, 3, 4, 0, 1],
                                                    [7, 8, 9, 5, 6], [12, 13, 14, 10, 11]]


def test_make_exits_brute_force():
    rng = np.random.default_rng(1)
    R, U, V = _tiny_geometry(rng)
    O = _rand_c(rng, (U, V))
    P = _rand_c(rng, (4, 5))
    ex = oracle.make_exits(O, P, R)
    for n, (r0, r1) in enumerate(R):
        for i in range(4):
            for j in range(5):
                # equal up to the rounding of a (possibly fused) complex product
                assert abs(ex[n, i, j] - O[i - r0, j - r1] * P[i, j]) <= 1e-15 * abs(ex[n, i, j])


def test_psup_O_brute_force_and_adjoint():
    rng = np.random.default_rng(2)
    R, U, V = _tiny_geometry(rng)
    P = _rand_c(rng, (4, 5))
    psi = _rand_c(rng, (len(R), 4, 5))
    alpha = 1e-3
    O, H = oracle.psup_O(psi, P, R, (U, V), None, alpha)
    num = np.zeros((U, V), complex)
    den = np.zeros((U, V))
    for n, (r0, r1) in enumerate(R):
        for i in range(4):
            for j in range(5):
                num[i - r0, j - r1] += np.conj(P[i, j]) * psi[n, i, j]
                den[i - r0, j - r1] += abs(P[i, j]) ** 2
    assert np.allclose(H, den, rtol=1e-14, atol=0)
    assert np.allclose(O, num / (den + alpha), rtol=1e-13, atol=1e-15)
    # adjoint identity <make_exits(X, P), psi> = <X, numerator(psi)>
    X = _rand_c(rng, (U, V))
    lhs = np.vdot(oracle.make_exits(X, P, R), psi)
    rhs = np.vdot(X, O * (H + alpha))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # heatmap = adjoint applied to |P|^2
    ones = np.ones((U, V), complex)
    h_adj = np.zeros((U, V))
    for n, (r0, r1) in enumerate(R):
        h_adj[-r0:4 - r0, -r1:5 - r1] += (np.abs(P) ** 2)
    assert np.allclose(H, h_adj)
    assert np.isclose(np.sum(H), len(R) * np.sum(np.abs(P) ** 2))
    del ones


def test_psup_P_brute_force_and_heatmap():
    rng = np.random.default_rng(3)
    R, U, V = _tiny_geometry(rng)
    O = _rand_c(rng, (U, V))
    psi = _rand_c(rng, (len(R), 4, 5))
    P, G = oracle.psup_P(psi, O, R, None, 1e-10)
    num = np.zeros((4, 5), complex)
    den = np.zeros((4, 5))
    for n, (r0, r1) in enumerate(R):
        for i in range(4):
            for j in range(5):
                num[i, j] += np.conj(O[i - r0, j - r1]) * psi[n, i, j]
                den[i, j] += abs(O[i - r0, j - r1]) ** 2
    assert np.allclose(G, den, rtol=1e-13)
    assert np.allclose(P, num / (den + 1e-10), rtol=1e-12)


def test_psup_fixed_points():
    rng = np.random.default_rng(4)
    R, U, V = _tiny_geometry(rng, N=6)
    O = _rand_c(rng, (U, V))
    P = _rand_c(rng, (4, 5))
    psi = oracle.make_exits(O, P, R)
    O2, H = oracle.psup_O(psi, P, R, (U, V), None, 0.0 + 1e-300)
    cov = H > 0
    assert np.allclose(O2[cov], O[cov], rtol=1e-12)
    assert np.all(O2[~cov] == 0)
    P2, _ = oracle.psup_P(psi, O, R, None, 1e-300)
    assert np.allclose(P2, P, rtol=1e-12)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (16, 4)])
def test_numpy_fft_matches_direct_dft(shape):
    rng = np.random.default_rng(5)
    x = _rand_c(rng, (3,) + shape)
    assert np.allclose(np.fft.fftn(x, axes=(-2, -1)), _dft2(x), atol=1e-11)
    assert np.allclose(np.fft.ifftn(x, axes=(-2, -1)), _idft2(x), atol=1e-12)


def test_pmod_1_against_direct_dft_and_invariants():
    rng = np.random.default_rng(6)
    N, M, K = 3, 8, 8
    psi = _rand_c(rng, (N, M, K))
    amp = np.abs(_rand_c(rng, (N, M, K)))
    mask = rng.random((M, K)) > 0.3
    out, emod = oracle.pmod_1(amp, psi.copy(), mask, alpha=1e-10, eMod_calc=True)
    # brute force with the direct DFT
    f = _dft2(psi)
    mg = np.abs(f) + 1e-10
    e_ref = np.sum(mask * (mg - amp) ** 2)
    fac = np.where(mask, amp / mg, 1.0)
    ref = _idft2(f * fac)
    assert np.allclose(out, ref, atol=1e-12)
    assert abs(emod - e_ref) < 1e-10 * e_ref
    # invariants: |F psi'| = amp on good pixels, F psi' = F psi on masked pixels
    fo = np.fft.fftn(out, axes=(-2, -1))
    assert np.allclose(np.abs(fo)[:, mask], amp[:, mask], rtol=1e-8)
    assert np.allclose(fo[:, ~mask], f[:, ~mask], atol=1e-10)
    # idempotent (up to alpha)
    out2 = oracle.pmod_1(amp, out.copy(), mask, alpha=1e-10)
    assert np.allclose(out2, out, atol=1e-8)


def test_pmod_7_invariants():
    rng = np.random.default_rng(7)
    N, M, K = 2, 8, 8
    psi = _rand_c(rng, (N, M, K))
    amp = np.abs(_rand_c(rng, (N, M, K))) + 0.5
    b = rng.random((N, M, K)) + 0.1
    mask = rng.random((M, K)) > 0.25
    out, b2, emod = oracle.pmod_7(amp, b.copy(), psi.copy(), mask, alpha=0.0, eMod_calc=True)
    fo = np.fft.fftn(out, axes=(-2, -1))
    tot = np.abs(fo) ** 2 + b2 ** 2
    assert np.allclose(tot[:, mask], amp[:, mask] ** 2, rtol=1e-9)
    assert np.abs(fo[:, ~mask]).max() < 1e-12 and np.all(b2[:, ~mask] == 0)
    f = _dft2(psi)
    fac = (np.sqrt(np.abs(f) ** 2 + b ** 2))
    assert np.allclose(b2[:, mask], (b * amp / fac)[:, mask])
    assert abs(emod - np.sum(mask * (fac - amp) ** 2)) < 1e-10 * emod


def test_pmod_P_invariant():
    rng = np.random.default_rng(8)
    P = _rand_c(rng, (8, 8))
    Pamp = np.abs(_rand_c(rng, (8, 8)))
    mask = rng.random((8, 8)) > 0.2
    Q = oracle.Pmod_P(Pamp, P, mask, 1e-12)
    fq = np.fft.fftn(Q)
    assert np.allclose(np.abs(fq)[mask], Pamp[mask], rtol=1e-9)
    assert np.allclose(fq[~mask], np.fft.fftn(P)[~mask])


def test_center_of_mass_matches_scipy_and_shift_rounding():
    rng = np.random.default_rng(9)
    a = rng.random((12, 9)) ** 4
    assert np.allclose(oracle.center_of_mass(a), scipy.ndimage.center_of_mass(a), rtol=1e-13)
    P = np.zeros((8, 8), complex)
    P[6, 1] = 1.0
    # cm = (6, 1) - shape // 2 = (2, -3); rolling by -cm centres the peak at (4, 4)
    assert oracle.probe_centre_shift(P).tolist() == [2, -3]
    assert np.argwhere(np.abs(oracle.multiroll(P, [-2, 3])) > 0).tolist() == [[4, 4]]
    # rint rounds half to even
    P = np.zeros((8, 8), complex)
    P[4, 4] = P[5, 4] = 1.0          # centroid row 4.5 -> 4 (even)
    assert oracle.probe_centre_shift(P).tolist() == [0, 0]


# ---------------------------------------------------------------------------
def test_era_method1_emod_non_increasing():
    s = small_scan(frame=32, grid=(4, 4), step=6, poisson=True)
    O, info = oracle.ERA(s.I, s.R.copy(), s.P, None, 25, method=1)
    e = np.array(info["eMod"])
    # alternating projections: distance to the modulus set cannot grow
    assert np.all(np.diff(e) <= 1e-9 * e[:-1])


def test_era_exact_data_fixed_point():
    s = small_scan(frame=32, grid=(4, 4), step=6, poisson=False)
    R = s.R.copy()
    O0 = s.O.copy()
    O, info = oracle.ERA(s.I, R, s.P, O0, 3, method=1)
    H = info["heatmap"]
    cov = H > 1e-6 * H.max()
    assert max(info["eMod"]) < 1e-6
    assert np.allclose(O[cov], s.O[cov], rtol=1e-6, atol=1e-8)


def test_dm_exact_data_fixed_point():
    s = small_scan(frame=32, grid=(4, 4), step=6, poisson=False)
    O, info = oracle.DM(s.I, s.R.copy(), s.P, s.O.copy(), 3, method=1)
    H = info["heatmap"]
    cov = H > 1e-6 * H.max()
    assert max(info["eMod"]) < 1e-6
    assert max(info["eSup"]) < 1e-6
    assert np.allclose(O[cov], s.O[cov], rtol=1e-6, atol=1e-8)


def test_preamble_side_effects_and_defaults():
    s = small_scan(frame=16, grid=(3, 3), step=4)
    R = s.R.copy() - 3
    st = oracle.preamble(s.I, R, None, None, 1, 2, 1, None, None, 1e-10, None)
    assert st["method"] == 3 and st["update"] == "OP" and st["OP_iters"] == (2, 1)
    assert R.max(axis=0).tolist() == [0, 0]                  # shifted in place
    assert st["O"].shape == (16 + np.ptp(R[:, 0]), 16 + np.ptp(R[:, 1]))
    assert np.all(st["O"] == 1)
    assert st["c_dtype"] == np.complex128
    st = oracle.preamble(s.I.astype(np.float32), s.R.copy(), s.P, None, 1, 1, 1, np.ones((16, 16)),
                         None, 1e-10, None)
    assert st["method"] == 4 and st["c_dtype"] == np.complex64
    assert np.allclose(st["background"], 1.0) and st["background"].shape == s.I.shape


@pytest.mark.parametrize("alg,method", [("ERA", 3), ("DM", 3), ("DM", 6)])
def test_slice_windows_bit_identical_to_roll(alg, method):
    s = small_scan(frame=16, grid=(3, 3), step=4, background_frac=0.05 if method > 3 else 0.0)
    fn = oracle.ERA if alg == "ERA" else oracle.DM
    bg = None if method < 4 else np.full(s.I.shape[1:], 0.3)
    outs = []
    for mode in ("roll", "slice"):
        with oracle.use_windows(mode):
            outs.append(fn(s.I, s.R.copy(), s.P * 1.1, None, 4, method=method, background=bg,
                           probe_centering=True))
    for a, b in zip(outs[0][:-1], outs[1][:-1]):
        assert np.array_equal(a, b)
    for key in ("eMod", "eCon", "eSup"):
        assert outs[0][-1][key] == outs[1][-1][key]
    assert np.array_equal(outs[0][-1]["I"], outs[1][-1]["I"])
    assert oracle.era.WINDOWS == "roll"
