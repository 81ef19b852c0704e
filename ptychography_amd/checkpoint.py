"""Checkpoint / resume of an ERA / DM run.

The reference resumes by passing a run's outputs back in as the next run's
initial guesses (test.py:48-49, DM then ERA); ERA / DM here accept the same.
``save`` writes what that needs -- O, P, the background, the iteration count
and the metric history -- to one ``.npz`` file (no pickled objects), ``load``
reads it back.
"""

import numpy as np

__all__ = ["save", "load"]


def save(path, O=None, P=None, info=None, iteration=0, background=None):
    arrays = {"iteration": np.int64(iteration)}
    if O is not None:
        arrays["O"] = np.asarray(O)
    if P is not None:
        arrays["P"] = np.asarray(P)
    if background is None and info is not None and info.get("background") is not None:
        background = info["background"]
    if background is not None:
        arrays["background"] = np.asarray(background)
    for k in ("eMod", "eCon", "eSup"):
        if info is not None and k in info:
            arrays[k] = np.asarray(info[k], np.float64)
    np.savez(path, **arrays)


def load(path):
    """dict with the keys that were saved (O, P, background, eMod, eCon, eSup,
    iteration)."""
    with np.load(path, allow_pickle=False) as d:
        out = {k: d[k] for k in d.files}
    out["iteration"] = int(out["iteration"])
    return out
