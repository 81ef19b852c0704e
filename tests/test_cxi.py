"""CXI result files (ptychography/display.py:21-117) written and read through
the HDF5 C library.

Checks: every layout the reference's arrays take (float32/64, complex64/128,
int64 positions, the bool mask and its scalar stand-in 1) round-trips exactly,
with and without gzip; ``maxlen`` cuts the frame stacks; absent entries come
back as None; an existing file is replaced.  The file itself is checked with
``h5dump``, the HDF5 library's own dumper (an independent reader): group and
dataset names, h5py's complex (compound r / i) and bool (int8 enum) layouts,
the deflate filter on the frame stacks, and the stored values of eMod.
"""

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from ptychography_amd import read_cxi, write_cxi
from tests.scans import small_scan

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


def _arrays(rng):
    N, M = 5, 16
    return dict(
        I_in=rng.random((N, M, M)) * 1e4,
        I_out=(rng.random((N, M, M)) * 1e4).astype(np.float32),
        P_in=(rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))),
        P_out=(rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))).astype(np.complex64),
        O_in=rng.standard_normal((30, 31)) + 1j * rng.standard_normal((30, 31)),
        O_out=rng.standard_normal((30, 31)) + 1j * rng.standard_normal((30, 31)),
        R_in=-rng.integers(0, 14, size=(N, 2)).astype(np.int64),
        R_out=None, B_in=rng.random((M, M)), B_out=None,
        M=rng.random((M, M)) > 0.1, eMod=list(rng.random(7)))


def _write(path, a, compress=True):
    write_cxi(a["I_in"], a["I_out"], a["P_in"], a["P_out"], a["O_in"], a["O_out"], a["R_in"],
              a["R_out"], a["B_in"], a["B_out"], a["M"], a["eMod"], fnam=str(path), compress=compress)


KEYS = ("I_in", "I_out", "P_in", "P_out", "O_in", "O_out", "R_in", "R_out", "B_in", "B_out", "M",
        "eMod")


@pytest.mark.parametrize("compress", [True, False])
def test_round_trip_every_layout(tmp_path, compress):
    a = _arrays(np.random.default_rng(1))
    f = tmp_path / "run.cxi"
    _write(f, a, compress)
    got = dict(zip(KEYS, read_cxi(str(f))))
    for k in KEYS:
        if a[k] is None:
            assert got[k] is None
            continue
        want = np.asarray(a[k])
        assert got[k].dtype == want.dtype, k
        assert np.array_equal(got[k], want), k
    # maxlen keeps the first frames of both stacks
    got = read_cxi(str(f), maxlen=2)
    assert np.array_equal(got[0], a["I_in"][:2]) and np.array_equal(got[1], a["I_out"][:2])


def test_scalar_mask_and_replacement(tmp_path, capsys):
    a = _arrays(np.random.default_rng(2))
    a["M"] = 1                                   # forward_sim's mask when there is none
    f = tmp_path / "run.cxi"
    _write(f, a)
    _write(f, a)                                 # replaces the file
    assert "already exists" in capsys.readouterr().out
    M = read_cxi(str(f))[10]
    assert np.ndim(M) == 0 and int(M) == 1
    with pytest.raises(ValueError):
        _write(tmp_path / "missing_dir" / "x.cxi", a)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_h5dump_reads_the_file(tmp_path):
    a = _arrays(np.random.default_rng(3))
    f = tmp_path / "run.cxi"
    _write(f, a)
    head = subprocess.run([H5DUMP, "-H", "-p", str(f)], capture_output=True, text=True, check=True).stdout
    for grp in ("input", "output"):
        assert f'GROUP "{grp}"' in head
    for ds in ("I", "M", "P", "O", "R", "B"):
        assert f'DATASET "{ds}"' in head
    assert 'DATASET "eMod"' in head
    # h5py's layouts: complex as compound {r, i}, bool as an int8 enum
    assert re.search(r'H5T_COMPOUND \{\s*H5T_IEEE_F64LE "r";\s*H5T_IEEE_F64LE "i";', head)
    assert re.search(r'H5T_COMPOUND \{\s*H5T_IEEE_F32LE "r";\s*H5T_IEEE_F32LE "i";', head)
    assert re.search(r'H5T_ENUM \{\s*H5T_STD_I8LE;\s*"FALSE"\s*0;\s*"TRUE"\s*1;', head)
    assert head.count("COMPRESSION DEFLATE { LEVEL 4 }") == 2        # the two frame stacks
    # values through the dumper: eMod at full precision
    body = subprocess.run([H5DUMP, "-d", "/output/eMod", "-m", "%.17g", "-y", "-w", "0", str(f)],
                          capture_output=True, text=True, check=True).stdout
    data = body.split("DATA {", 1)[1].split("}", 1)[0]
    vals = [float(v) for v in data.replace(",", " ").split()]
    assert vals == list(a["eMod"])
    body = subprocess.run([H5DUMP, "-d", "/input/R", "-y", "-w", "0", str(f)],
                          capture_output=True, text=True, check=True).stdout
    ints = [int(v) for v in body.split("DATA {", 1)[1].split("}", 1)[0].replace(",", " ").split()]
    assert ints == a["R_in"].ravel().tolist()


def test_reference_call_after_a_reconstruction(tmp_path):
    # test.py:55-56: the inputs, info['I'] and the retrieved object and probe
    s = small_scan(frame=16, grid=(3, 3), step=4)
    O, P, info = oracle.ERA(s.I, s.R.copy(), s.P, None, 3, method=3)
    f = tmp_path / "output_method1.cxi"
    write_cxi(s.I, info['I'], s.P, P, s.O, O, s.R, None, None, None, s.mask, info['eMod'],
              fnam=str(f))
    I_in, I_out, P_in, P_out, O_in, O_out, R_in, R_out, B_in, B_out, M, eMod = read_cxi(str(f))
    assert np.array_equal(I_in, s.I) and np.array_equal(I_out, info['I'])
    assert np.array_equal(P_out, P) and np.array_equal(O_out, O)
    assert np.array_equal(eMod, np.array(info['eMod']))
    assert R_out is None and B_in is None


def test_every_declared_cxi_symbol_is_exported():
    import pathlib
    from ptychography_amd import cxi
    header = (pathlib.Path(__file__).resolve().parents[1] / "include" / "ptycxi.h").read_text()
    names = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pty_\w+)\s*\(", header, re.M)))
    assert len(names) == 5
    lib = cxi._load()
    for n in names:
        assert hasattr(lib, n), n
