"""The seeded generator keeps the reference's forward_sim conventions
(ptychography/forward_models.py:119-192)."""

import numpy as np

import oracle
from ptychography_amd import sim


def test_recipe_shapes():
    s = sim.make_scan("c1")
    assert s.I.shape == (144, 64, 64)
    assert s.O.shape == (256, 256)
    assert s.R.max(axis=0).tolist() == [0, 0]
    assert (-s.R.min(axis=0) + 64 <= 256).all()
    sp = sim.RECIPES["c3"]
    assert sp.grid == (64, 64) and sp.frame == 256 and sp.step == 40 and sp.jitter == 3


def test_intensities_are_centred_fft_of_exits():
    s = sim.make_scan("c1", noise=False, grid=(3, 3))
    ex = oracle.make_exits(s.O, s.P, s.R)
    I = np.fft.ifftshift(np.abs(np.fft.fftn(ex, axes=(-2, -1))) ** 2, axes=(-2, -1))
    assert np.allclose(s.I, I, rtol=1e-10)
    assert np.isclose(s.I.sum() / len(s.I), sim.RECIPES["c1"].photons, rtol=1e-12)


def test_seeded_and_poisson():
    a = sim.make_scan("c1", grid=(2, 2))
    b = sim.make_scan("c1", grid=(2, 2))
    assert np.array_equal(a.I, b.I) and np.array_equal(a.R, b.R)
    assert np.all(a.I == np.round(a.I)) and a.I.min() >= 0


def test_mask_and_background():
    s = sim.make_scan(sim.RECIPES["c2"].with_(grid=(2, 2)))
    assert s.mask.dtype == bool and s.mask.shape == (128, 128)
    assert not s.mask[64, 64]                         # beamstop at the centre
    assert 0.05 < 1 - s.mask.mean() < 0.2             # ~10 % masked
    t = sim.make_scan(sim.RECIPES["c3"].with_(grid=(2, 2), poisson=False))
    assert t.B.shape == (256, 256)
    assert np.isclose(t.B.mean(), 0.01 * (t.I - t.B[None]).mean(), rtol=1e-6)


def test_probe_pupil_is_band_limited():
    P = sim.make_probe(64, 16, 0.02)
    f = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(P)))
    q = np.hypot(*np.meshgrid(np.arange(64) - 32, np.arange(64) - 32, indexing="ij"))
    assert np.abs(f[q >= 16.5]).max() < 1e-10 * np.abs(f).max()
