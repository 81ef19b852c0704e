"""``write_cxi`` / ``read_cxi`` -- the reference's result files.

Same names, arguments and returns as ptychography/display.py:21-117: groups
``input`` (I, M, P, O, R, B) and ``output`` (I, P, O, R, B, eMod), the two
``I`` stacks gzip-compressed when ``compress`` is set, ``None`` entries left
out on writing and returned as ``None`` on reading.  The files are written
through the HDF5 C library (``libptycxi.so``, ``csrc/cxi.cpp``) in the layout
h5py gives the same numpy arrays (complex as an {r, i} compound, bool as an
int8 enum), so they open in h5py and in the reference's own reader.

Differences from the reference, where its code does not run as written:
``read_cxi`` opens the path it is given (the reference passes an undefined
name to ``h5py.File``, display.py:64) and takes a path, not an open h5py file.
"""

import ctypes
import os
import pathlib

import numpy as np

__all__ = ["write_cxi", "read_cxi"]

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libptycxi.so"

F32, F64, C64, C128, I32, I64, BOOL = 1, 2, 3, 4, 5, 6, 7
ENOENT = -6
MAXDIM = 8
_NP = {F32: np.float32, F64: np.float64, C64: np.complex64, C128: np.complex128,
       I32: np.int32, I64: np.int64, BOOL: np.bool_}

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m ptychography_amd._build`")
        lib = ctypes.CDLL(str(LIB_PATH))
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        lib.pty_cxi_last_error.restype = ctypes.c_char_p
        lib.pty_cxi_last_error.argtypes = []
        lib.pty_cxi_create.restype = ctypes.c_int
        lib.pty_cxi_create.argtypes = [ctypes.c_char_p]
        lib.pty_cxi_write.restype = ctypes.c_int
        lib.pty_cxi_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, i32, vp, vp, i32]
        lib.pty_cxi_info.restype = ctypes.c_int
        lib.pty_cxi_info.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, vp, vp]
        lib.pty_cxi_read.restype = ctypes.c_int
        lib.pty_cxi_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, vp, i64]
        _lib = lib
    return _lib


def _check(code, what):
    if code != 0:
        raise OSError(f"{what}: {_load().pty_cxi_last_error().decode()}")


def _layout(a):
    """(layout code, contiguous array in that layout) for a numpy value."""
    a = np.asarray(a)
    # np.require keeps rank 0 (ascontiguousarray would make a scalar 1-D)
    if a.dtype == np.bool_:
        return BOOL, np.require(a, np.uint8, "C")
    if np.issubdtype(a.dtype, np.complexfloating):
        code = C64 if a.dtype == np.complex64 else C128
    elif np.issubdtype(a.dtype, np.floating):
        code = F32 if a.dtype == np.float32 else F64
    elif np.issubdtype(a.dtype, np.integer):
        code = I32 if a.dtype.itemsize <= 4 else I64
    else:
        raise TypeError(f"cannot store dtype {a.dtype}")
    return code, np.require(a, _NP[code], "C")


def _write(path, name, value, gzip=0):
    code, a = _layout(value)
    if a.ndim > MAXDIM:
        raise ValueError(f"{name}: rank {a.ndim} above {MAXDIM}")
    dims = np.asarray(a.shape, np.int64)
    _check(_load().pty_cxi_write(path, name.encode(), code, a.ndim,
                                 dims.ctypes.data_as(ctypes.c_void_p),
                                 a.ctypes.data_as(ctypes.c_void_p), int(gzip)), f"writing {name}")


def _read(path, name, max_rows=-1):
    """The dataset as a numpy array (a numpy scalar for rank 0), or None if absent."""
    lib = _load()
    code = ctypes.c_int32()
    ndim = ctypes.c_int32()
    dims = np.zeros(MAXDIM, np.int64)
    r = lib.pty_cxi_info(path, name.encode(), ctypes.byref(code), ctypes.byref(ndim),
                         dims.ctypes.data_as(ctypes.c_void_p))
    if r == ENOENT:
        return None
    _check(r, f"inspecting {name}")
    shape = [int(d) for d in dims[:ndim.value]]
    if shape and max_rows >= 0:
        shape[0] = min(shape[0], int(max_rows))
    out = np.empty(shape, np.uint8 if code.value == BOOL else _NP[code.value])
    if out.size:
        _check(lib.pty_cxi_read(path, name.encode(), code.value,
                                out.ctypes.data_as(ctypes.c_void_p), int(max_rows)),
               f"reading {name}")
    if code.value == BOOL:
        out = out.astype(np.bool_)
    return out[()] if out.ndim == 0 else out


def _remove_existing(fnam):
    """display.py:5-19: the directory must exist; an existing file is replaced."""
    out_dir = os.path.split(os.path.realpath(fnam))[0]
    if not os.path.exists(out_dir):
        raise ValueError('specified path does not exist: ', out_dir)
    if not os.path.isdir(out_dir):
        raise ValueError('specified path is not a path: ', out_dir)
    if os.path.exists(fnam):
        print('\n', fnam, 'file already exists, deleting the old one and making a new one')
        os.remove(fnam)


def write_cxi(I_in, I_out, P_in, P_out, O_in, O_out, R_in, R_out, B_in, B_out, M, eMod,
              fnam='output.cxi', compress=True):
    """Write the input and output of a reconstruction (display.py:21-56).

    Replaces an existing file.  ``I_in`` / ``I_out`` are stored with gzip when
    ``compress``; ``M`` is always written (the scalar 1 when there is no mask);
    any other argument that is ``None`` is left out."""
    _remove_existing(fnam)
    path = os.fsencode(fnam)
    _check(_load().pty_cxi_create(path), f"creating {fnam}")
    gz = 4 if compress else 0   # h5py's default gzip level
    _write(path, "input/I", I_in, gz)
    _write(path, "input/M", M)
    for key, val in (("P", P_in), ("O", O_in), ("R", R_in), ("B", B_in)):
        if val is not None:
            _write(path, "input/" + key, val)
    _write(path, "output/I", I_out, gz)
    for key, val in (("P", P_out), ("O", O_out), ("R", R_out), ("B", B_out), ("eMod", eMod)):
        if val is not None:
            _write(path, "output/" + key, val)


def read_cxi(fnam='output.cxi', maxlen=100):
    """Read a file written by ``write_cxi`` (display.py:59-117): returns
    (I_in, I_out, P_in, P_out, O_in, O_out, R_in, R_out, B_in, B_out, M, eMod),
    the two ``I`` stacks cut to their first ``maxlen`` frames, absent entries
    as ``None``."""
    path = os.fsencode(fnam)
    I_in = _read(path, "input/I", maxlen)
    M = _read(path, "input/M")
    P_in, O_in, R_in, B_in = (_read(path, "input/" + k) for k in ("P", "O", "R", "B"))
    I_out = _read(path, "output/I", maxlen)
    P_out, O_out, R_out, B_out, eMod = (_read(path, "output/" + k)
                                        for k in ("P", "O", "R", "B", "eMod"))
    return I_in, I_out, P_in, P_out, O_in, O_out, R_in, R_out, B_in, B_out, M, eMod
