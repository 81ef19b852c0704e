"""Launch-shape controls change how work is spread, not what is computed.

Every (column block, frame) item of the column kernel and every frame row of
the row kernels is computed the same way whatever the grid, the work-group
width or the frame blocking of ERA's modulus step, so the iterates must be
bit-identical; only the order in which the FP64 metric partials are added
changes (bar: 1e-12 relative).
"""

import numpy as np
import pytest
import torch

from ptychography_amd.engine import Engine
from tests.scans import broadband_start, small_scan

pytestmark = pytest.mark.gpu


def _run(s, dev, setup, method=1, iters=6):
    N, M, K = s.I.shape
    eng = Engine(N, M, K, *s.O.shape, s.R.astype(np.int32), device=dev)
    setup(eng)
    amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), None)
    eng.set_data(amp, None, I_norm)
    O_t = torch.from_numpy(broadband_start(s.O.shape).astype(np.complex64)).to(dev)
    P_t = torch.from_numpy(s.P.astype(np.complex64)).to(dev)
    eng.set_state(O_t, P_t)
    m = eng.run(0, method, iters, begin=True)
    out = (m, O_t.cpu().numpy(), P_t.cpu().numpy())
    eng.close()
    return out


SHAPES = {
    "default": lambda e: None,
    "persistent columns": lambda e: e.set_launch(3, 2, 0),
    "256-thread columns": lambda e: e.set_launch(3, 0, 256),
    "rows 1 x CUs, columns 8 x CUs": lambda e: e.set_launch(1, 8, 0),
    "frame blocks of 5": lambda e: e.set_chunk(5),
}


@pytest.mark.parametrize("frame,method", [(256, 1), (192, 3), (64, 1)])
def test_launch_shapes_give_identical_iterates(frame, method, gpu_device):
    print(f"launch shapes, frame {frame}, method {method}", flush=True)
    s = small_scan(frame=frame, grid=(4, 5), step=frame // 5, jitter=2)
    ref = _run(s, gpu_device, SHAPES["default"], method)
    for name, setup in SHAPES.items():
        out = _run(s, gpu_device, setup, method)
        assert np.array_equal(out[1], ref[1]), name
        assert np.array_equal(out[2], ref[2]), name
        np.testing.assert_allclose(out[0], ref[0], rtol=1e-12, atol=0, err_msg=name)
