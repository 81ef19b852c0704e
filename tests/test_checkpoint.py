"""Checkpoint round trip (host only)."""

import numpy as np

from ptychography_amd import checkpoint


def test_checkpoint_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    O = rng.standard_normal((20, 24)) + 1j * rng.standard_normal((20, 24))
    P = (rng.standard_normal((8, 8)) + 1j).astype(np.complex64)
    info = {"eMod": [0.5, 0.25], "eCon": [1.0, 0.1], "eSup": [0.3, 0.2],
            "background": rng.random((8, 8))}
    path = tmp_path / "run.npz"
    checkpoint.save(path, O=O, P=P, info=info, iteration=2)
    d = checkpoint.load(path)
    assert d["iteration"] == 2
    assert np.array_equal(d["O"], O) and d["O"].dtype == O.dtype
    assert np.array_equal(d["P"], P) and d["P"].dtype == P.dtype
    assert np.array_equal(d["background"], info["background"])
    assert d["eMod"].tolist() == info["eMod"]
