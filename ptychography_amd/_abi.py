"""ctypes binding of ``libptyhip.so`` (declared in ``include/ptyhip.h``).

The library is built in-tree by ``ptychography_amd._build``.  Loading fails
loudly -- there is no CPU fallback for the iteration path.
"""

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libptyhip.so"

PTY_OK = 0
PTY_EINVAL = -1
PTY_ENOMEM = -2
PTY_EHIP = -3
PTY_ESTATE = -4
PTY_EUNSUP = -5

_CODES = {PTY_EINVAL: "EINVAL", PTY_ENOMEM: "ENOMEM", PTY_EHIP: "EHIP",
          PTY_ESTATE: "ESTATE", PTY_EUNSUP: "EUNSUP"}

# every symbol include/ptyhip.h declares: name -> (restype, argtypes)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double


class PtyTuneResult(ctypes.Structure):
    _fields_ = [("fused", ctypes.c_int32), ("frame_threads", ctypes.c_int32),
                ("frame_slots", ctypes.c_int32), ("rows_mult", ctypes.c_int32),
                ("cols_mult", ctypes.c_int32), ("cols_wg", ctypes.c_int32),
                ("chunk_frames", ctypes.c_int32),
                ("row_gather", ctypes.c_int32), ("fft_plan", ctypes.c_int32),
                ("candidates", ctypes.c_int32),
                ("best_ms", ctypes.c_float), ("default_ms", ctypes.c_float)]


class PtyOpts(ctypes.Structure):
    """Mirror of ``pty_opts``."""
    _fields_ = [("algorithm", _i32), ("method", _i32), ("op_n", _i32), ("op_every", _i32),
                ("pmod_probe_until", _i64), ("probe_centering", _i32), ("alpha", _f32)]


SIGNATURES = {
    "pty_supported_frame": (ctypes.c_int, [_i32]),
    "pty_last_error": (ctypes.c_char_p, []),
    "pty_create": (ctypes.c_int, [ctypes.POINTER(_vp), _i32, _i32, _i32, _i32, _i32, _vp]),
    "pty_destroy": (ctypes.c_int, [_vp]),
    "pty_set_stream": (ctypes.c_int, [_vp, _vp]),
    "pty_prepare_amp": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(_f64)]),
    "pty_set_data": (ctypes.c_int, [_vp, _vp, _vp, _f64]),
    "pty_set_state": (ctypes.c_int, [_vp, _vp, _vp]),
    "pty_set_background": (ctypes.c_int, [_vp, _vp, _i32]),
    "pty_set_pamp": (ctypes.c_int, [_vp, _vp]),
    "pty_set_subpixel": (ctypes.c_int, [_vp, _vp]),
    "pty_set_scans": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp]),
    "pty_set_plan": (ctypes.c_int, [_vp, _i32]),
    "pty_get_plan": (ctypes.c_int, [_vp, _vp, _vp]),
    "pty_position_terms": (ctypes.c_int, [_vp, _vp]),
    "pty_run": (ctypes.c_int, [_vp, ctypes.POINTER(PtyOpts), _i64, _i32, _i32, _vp]),
    "pty_intensity": (ctypes.c_int, [_vp, _vp]),
    "pty_heatmap": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_i32)]),
    "pty_background": (ctypes.c_int, [_vp, _vp]),
    "pty_set_heatmap": (ctypes.c_int, [_vp, _vp, _vp]),
    "pty_set_iterate": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pty_get_iterate": (ctypes.c_int, [_vp, _vp, _vp]),
    "pty_set_path": (ctypes.c_int, [_vp, _i32, _i32, _i32]),
    "pty_set_graphs": (ctypes.c_int, [_vp, _i32]),
    "pty_profile": (ctypes.c_int, [_vp, _i32]),
    "pty_profile_read": (ctypes.c_int, [_vp, _vp, _vp]),
    "pty_fft2": (ctypes.c_int, [_vp, _vp, _i32, _i32]),
    "pty_make_exits": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pty_pmod1": (ctypes.c_int, [_vp, _vp, _f32, ctypes.POINTER(_f64)]),
    "pty_pmod1_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _i32, _f32, _vp, _vp]),
    "pty_psup_O": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _f32]),
    "pty_psup_P": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _f32]),
    "pty_back_projection": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _f32]),
    "pty_hio_modulus": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _f32, _f32, _f32,
                                       ctypes.POINTER(_f64)]),
    "pty_add_exits": (ctypes.c_int, [_vp, _vp, _vp, _f32, _vp]),
    "pty_set_launch": (ctypes.c_int, [_vp, _i32, _i32, _i32]),
    "pty_set_chunk": (ctypes.c_int, [_vp, _i32]),
    "pty_set_row_gather": (ctypes.c_int, [_vp, _i32]),
    "pty_get_row_gather": (ctypes.c_int, [_vp, _vp, _vp]),
    "pty_tune": (ctypes.c_int, [_vp, _vp, _i32, _vp]),
    "pty_batch_create": (ctypes.c_int, [_vp, _i32, _vp]),
    "pty_batch_destroy": (ctypes.c_int, [_vp]),
    "pty_batch_set_graphs": (ctypes.c_int, [_vp, _i32]),
    "pty_batch_get_graphs": (ctypes.c_int, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "pty_get_graphs": (ctypes.c_int, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "pty_batch_run": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _i32, _i32, _vp]),
    "pty_nn_last_error": (ctypes.c_char_p, []),
    "pty_nn_gemm_probe": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _i32, _vp]),
    "pty_nn_conv3x3_fwd": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32,
                                          _i32, _i32, _vp]),
    "pty_nn_conv3x3_fwd_intensity": (ctypes.c_int, [_vp, _f32, _vp, _vp, _vp, _vp, _i32, _i32, _i32,
                                                    _i32, _i32, _vp]),
    "pty_nn_conv3x3_dgrad": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp]),
    "pty_nn_weight_t": (ctypes.c_int, [_vp, _vp, _i32, _i32, _vp]),
    "pty_nn_conv3x3_wgrad": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32,
                                            _i32, _vp]),
    "pty_nn_wgrad_splits": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32]),
    "pty_nn_relu_bwd": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "pty_nn_pool2": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp]),
    "pty_nn_up2": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp]),
    "pty_nn_mse": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp, _f32, _vp, _vp, ctypes.c_size_t,
                                  _vp]),
    "pty_nn_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _f32, _f32, _f32, _f32,
                                   _i32, _f32, _vp]),
    "pty_nn_to_bf16": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "pty_nn_batch_gather": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _i32, ctypes.c_uint32, _vp,
                                           _vp, _vp]),
    "pty_nn_pair_overlap": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp]),
    "pty_nn_phase_sync": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp]),
    "pty_nn_phase_exits": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp]),
}

SUPPORTED_FRAMES = (16, 32, 64, 96, 128, 192, 256)
STAGES = ("rows_fwd", "cols", "rows_inv", "gather", "probe", "reduce", "misc")

_lib = None


class PtyError(RuntimeError):
    """A ptyhip call returned a negative status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({_CODES.get(code, code)}): {msg}")
        self.code = code


def load(path=None):
    """Load the library and declare every exported signature (no GPU needed)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m ptychography_amd._build` "
                           "(the ERA/DM path has no CPU fallback)")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(fn_name, code):
    if code != PTY_OK:
        lib = load()
        msg = lib.pty_nn_last_error() if fn_name.startswith("pty_nn_") else lib.pty_last_error()
        raise PtyError(fn_name, code, msg.decode() if msg else "")


def call(fn_name, *args):
    code = getattr(load(), fn_name)(*args)
    check(fn_name, code)
    return code
