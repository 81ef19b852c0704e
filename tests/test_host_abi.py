"""CPU tests: the C ABI loads and exports every declared symbol, argument
validation fails with the documented codes before any device work, and the
host-side preamble logic of the facade matches the reference's rules."""

import ctypes
import pathlib
import re

import numpy as np
import pytest

from ptychography_amd import _abi
from ptychography_amd.era import pmod_probe_until, resolve_method

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "ptyhip.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pty_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _abi.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} missing from libptyhip.so"
        assert n in _abi.SIGNATURES, f"{n} has no ctypes signature"
    assert set(_abi.SIGNATURES) == set(names)


def test_opts_struct_layout_matches_header():
    assert ctypes.sizeof(_abi.PtyOpts) == 32
    assert _abi.PtyOpts.pmod_probe_until.offset == 16
    assert _abi.PtyOpts.alpha.offset == 28


def test_supported_frames():
    lib = _abi.load()
    for L in _abi.SUPPORTED_FRAMES:
        assert lib.pty_supported_frame(L) == 1
    for L in (0, 8, 48, 100, 200, 512):
        assert lib.pty_supported_frame(L) == 0


def _create(N, M, K, U, V, R):
    lib = _abi.load()
    ctx = ctypes.c_void_p()
    Rc = np.ascontiguousarray(R, dtype=np.int32)
    code = lib.pty_create(ctypes.byref(ctx), N, M, K, U, V, Rc.ctypes.data_as(ctypes.c_void_p))
    return code, lib.pty_last_error().decode()


def test_create_validates_before_touching_the_device():
    R = np.zeros((2, 2), np.int32)
    assert _create(2, 100, 64, 200, 64, R)[0] == _abi.PTY_EUNSUP
    assert _create(0, 64, 64, 64, 64, R)[0] == _abi.PTY_EINVAL
    assert _create(2, 64, 64, 32, 64, R)[0] == _abi.PTY_EINVAL          # object smaller than frame
    code, msg = _create(2, 64, 64, 70, 70, np.array([[0, 0], [-10, 0]], np.int32))
    assert code == _abi.PTY_EINVAL and "outside the object" in msg      # window 10..74 > 70
    code, msg = _create(2, 64, 64, 70, 70, np.array([[0, 0], [3, 0]], np.int32))
    assert code == _abi.PTY_EINVAL                                       # R > 0 after the shift
    lib = _abi.load()
    assert lib.pty_create(None, 2, 64, 64, 64, 64, None) == _abi.PTY_EINVAL


def test_null_context_calls_fail_cleanly():
    lib = _abi.load()
    assert lib.pty_destroy(None) == _abi.PTY_OK
    assert lib.pty_set_stream(None, None) == _abi.PTY_EINVAL
    assert lib.pty_run(None, None, 0, 1, 1, None) == _abi.PTY_EINVAL
    assert lib.pty_fft2(None, None, 1, 0) == _abi.PTY_EINVAL


def test_resolve_method_follows_preamble():
    assert resolve_method(None, None, None, None) == (3, 'OP')
    assert resolve_method(None, 1, None, None) == (1, 'O')
    assert resolve_method(1, None, None, None) == (2, 'P')
    assert resolve_method(None, 1, np.ones(2), None) == (4, 'O')
    assert resolve_method(None, None, np.ones(2), None) == (6, 'OP')
    assert resolve_method(1, 1, None, 5) == (5, 'P')
    with pytest.raises(ValueError):
        resolve_method(1, 1, None, None)
    with pytest.raises(ValueError):
        resolve_method(None, None, None, 7)


def test_pmod_probe_window_matches_reference_comparison():
    # the reference applies Pmod_P when `Pmod_probe is not None and i < Pmod_probe`
    for p in [False, None, True, 0, 1, 3, 2.5, np.inf, -1]:
        until = pmod_probe_until(p)
        for i in range(8):
            expect = p is not None and i < p
            assert (i < until) == bool(expect), (p, i)


def test_facade_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import ptychography_amd as pty
    I = np.ones((2, 16, 16))
    R = np.array([[0, 0], [-2, -2]])
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        pty.ERA(I, R, np.ones((16, 16)), None, 1, method=1)


def test_update_progress_reproduces_the_reference_line():
    import io
    from ptychography_amd import update_progress

    def line(*a):
        buf = io.StringIO()
        return update_progress(*a, stream=buf), buf.getvalue()

    t, written = line(0.5, 'ERA', 3, 0.25, 0.125)
    assert t == written == "\rERA: [########-------] 50% 3 0.25 0.125  " + " " * 5
    # 15 / 6 = 2.5 rounds half away from zero (Python 2), 1/6 -> 16 %
    assert line(1.0 / 6, 'DM', 1, 0.5, 0.5)[0].startswith("\rDM: [###------------] 16% 1 ")
    assert line(1.0, 'ERA', 9, 0.1, 0.2)[0] == ("\rERA: [###############] 100% 9 0.1 0.2 "
                                                 "Done...\r\n" + " " * 6)
    assert "Halt...\r\n" in line(-0.5, 'ERA', 0, 1.0, 1.0)[0]
    assert "error: progress var must be float" in line("x", 'ERA', 0, 1.0, 1.0)[0]
    # Python 2's str(float): 12 significant digits
    assert " 0.123456789012 " in line(0.0, 'ERA', 0, 0.1234567890123456, 2.0)[0]
    assert line(3, 'ERA', 0, 1.0, 1.0)[0].startswith("\rERA: [###############] 100%")
