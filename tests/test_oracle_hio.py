"""Pinning the HIO oracle (oracle/hio.py) without retyping its update.

* Fienup's classic input-output rule: when Psup is a 0/1 support projection,
  HIO's update f + b Psup((1 + 1/b) m - f/b) - b m, m = Pmod(f), is
  f <- m on the support and f <- f - b m off it.  With a binary probe and
  frames that do not overlap, psup_O gives O = z / (1 + alpha) on the
  illuminated pixels, so Psup(z) = P * O is that projection ('sample'); with
  one frame and a binary object, psup_P does the same ('probe').  A wrong
  coefficient or sign in the update breaks one of the two rules.
* Exact-data fixed point: data formed from (O, P) leave O, P and the exits in
  place for every update mode.
* The first eMod is ERA's (same modulus step on the same starting exits).
"""

import numpy as np
import pytest

import oracle
from oracle.era import pmod_1

ALPHA = 1e-10


def _amp_mask(I, mask):
    amp = np.fft.ifftshift(np.sqrt(I), axes=(-2, -1))
    m = np.asarray(mask)
    return amp, (np.fft.ifftshift(m) if m.ndim > 0 else m)


def _classic(f, I, mask, support, beta):
    amp, ms = _amp_mask(I, mask)
    m = pmod_1(amp, f, ms, alpha=ALPHA)
    return np.where(support, m, f - beta * m)


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def test_sample_update_is_fienups_rule_for_a_support_projection():
    rng = np.random.default_rng(0)
    M = 8
    off = np.array([[0, 0], [0, 8], [8, 3]])            # windows that do not overlap
    R = -off
    U, V = 16, 16
    P = (rng.random((M, M)) > 0.4).astype(np.complex128)    # binary probe
    O = rng.standard_normal((U, V)) + 1j * rng.standard_normal((U, V))
    I = rng.random((3, M, M)) * 4.0 + 0.5
    mask = np.ones((M, M), bool)
    mask[0, 1] = False                                   # a masked pixel keeps its value
    beta = 0.7
    f0 = oracle.make_exits(O, P, R)
    support = np.broadcast_to(P != 0, f0.shape)
    f1_expected = _classic(f0, I, mask, support, beta)
    f2_expected = _classic(f1_expected, I, mask, support, beta)
    for iters, want in ((1, f1_expected), (2, f2_expected)):
        _, _, info = oracle.HIO(I, R.copy(), P, O, iters, beta=beta, update='sample', mask=mask,
                                alpha=ALPHA)
        assert _rel(info['exits'], want) < 1e-9


def test_probe_update_is_fienups_rule_for_a_support_projection():
    rng = np.random.default_rng(1)
    M = 8
    R = np.zeros((1, 2), int)
    O = (rng.random((M, M)) > 0.5).astype(np.complex128)    # binary object, one frame
    P = rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))
    I = rng.random((1, M, M)) * 4.0 + 0.5
    beta = 1.3
    f0 = oracle.make_exits(O, P, R)
    support = (O != 0)[None]
    f1 = _classic(f0, I, 1, support, beta)
    _, _, info = oracle.HIO(I, R.copy(), P, O, 1, beta=beta, update='probe', alpha=ALPHA)
    assert _rel(info['exits'], f1) < 1e-9


@pytest.mark.parametrize("update", ["sample", "probe", "both"])
def test_exact_data_fixed_point(update):
    rng = np.random.default_rng(2)
    M = 16
    g = np.arange(3) * 5
    off = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    R = -off
    U = V = M + 10
    yy, xx = np.mgrid[:M, :M] - M / 2
    P = np.exp(-(yy ** 2 + xx ** 2) / 20.0) * np.exp(0.3j * xx)
    O = np.exp(1j * rng.standard_normal((U, V)) * 0.5) * (0.8 + 0.2 * rng.random((U, V)))
    ex = oracle.make_exits(O, P, R)
    I = np.abs(np.fft.fftshift(np.fft.fftn(ex, axes=(-2, -1)), axes=(-2, -1))) ** 2
    O2, P2, info = oracle.HIO(I, R.copy(), P, O, 5, beta=0.9, update=update, alpha=ALPHA)
    assert max(info['eMod']) < 1e-8
    assert _rel(P2, P) < 1e-8
    H = oracle.make_P_heatmap(P, R, O.shape)
    lit = H > 1e-3 * H.max()
    # psup_O returns O H / (H + alpha): relative change alpha / H <= 1e-7 here
    assert _rel(O2[lit], O[lit]) < 1e-6
    assert _rel(info['exits'], ex) < 1e-7      # the same alpha / H, weighted by |P|


def test_first_emod_is_eras():
    rng = np.random.default_rng(3)
    M = 8
    off = np.array([[0, 0], [0, 4], [4, 0], [4, 4]])
    R = -off
    P = rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))
    O = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    I = rng.random((4, M, M)) + 0.1
    _, _, ih = oracle.HIO(I, R.copy(), P, O, 2, beta=0.8)
    _, ie = oracle.ERA(I, R.copy(), P, O, 2, method=1)
    assert abs(ih['eMod'][0] - ie['eMod'][0]) <= 1e-12 * ie['eMod'][0]


def test_argument_errors():
    I = np.ones((1, 4, 4))
    R = np.zeros((1, 2), int)
    P = np.ones((4, 4), complex)
    with pytest.raises(ValueError):
        oracle.HIO(I, R.copy(), P, None, 1, update='object')
    with pytest.raises(ValueError):
        oracle.HIO(I, R.copy(), P, None, 1, beta=0.0)
