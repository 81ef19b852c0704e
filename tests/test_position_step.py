"""The host rule of one position-refinement step (``era.position_step``)."""

import numpy as np

from ptychography_amd.era import position_step


def _terms(g, H, f=1.0):
    return np.array([[f, g[0], g[1], H[0][0], H[0][1], H[1][1]]])


def test_step_is_the_damped_gauss_newton_solution():
    H = [[4.0, 1.0], [1.0, 2.0]]
    g = [0.2, -0.1]
    d = position_step(_terms(g, H), np.zeros((1, 2)), damping=0.0)
    assert np.allclose(d[0], -np.linalg.solve(np.array(H), g))
    dd = position_step(_terms(g, H), np.zeros((1, 2)), damping=0.5)
    assert np.allclose(dd[0], -np.linalg.solve(np.array(H) + 0.5 * 6.0 * np.eye(2), g))


def test_step_and_shift_are_clipped():
    d = position_step(_terms([-100.0, 100.0], [[1.0, 0.0], [0.0, 1.0]]), np.array([[0.8, -0.1]]))
    assert np.allclose(d, [[1.0, -0.6]])      # +0.5 capped at the 1-pixel bound; -0.5


def test_frames_without_signal_stay():
    d0 = np.array([[0.3, -0.2]])
    assert np.array_equal(position_step(_terms([0.0, 0.0], [[0.0, 0.0], [0.0, 0.0]]), d0), d0)
