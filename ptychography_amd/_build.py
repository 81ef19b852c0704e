"""Build the gfx950 shared library ``libptyhip.so`` in-tree with hipcc.

Usage: ``python -m ptychography_amd._build`` (or ``__graft_entry__.build()``).
"""

import os
import pathlib
import subprocess
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
SRCS = [HERE / "csrc" / "ptyhip.hip", HERE / "csrc" / "nn.hip"]
LIB = HERE / "libptyhip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PTY_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# the CXI file reader / writer: host C++ over the HDF5 C library of the image
HDF5_ROOT = pathlib.Path(os.environ.get("PTY_HDF5_ROOT", "/opt/conda"))
CXI_SRC = HERE / "csrc" / "cxi.cpp"
CXI_LIB = HERE / "libptycxi.so"


def build_cxi(verbose=False):
    """Compile ``csrc/cxi.cpp`` into ``libptycxi.so`` against libhdf5 (always recompiles)."""
    tmp = CXI_LIB.with_suffix(".so.tmp%d" % os.getpid())
    lib = HDF5_ROOT / "lib"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{HDF5_ROOT / 'include'}",
           f"-I{ROOT / 'include'}", str(CXI_SRC), f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}",
           "-o", str(tmp)]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        if tmp.exists():
            tmp.unlink()
        raise RuntimeError("g++ failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    os.replace(tmp, CXI_LIB)
    if verbose:
        print(f"built {CXI_LIB} in {time.time() - t0:.1f}s", file=sys.stderr)
    return CXI_LIB


def build(verbose=False):
    """Compile ``csrc/ptyhip.hip`` and ``csrc/nn.hip`` into ``libptyhip.so`` (always recompiles)."""
    tmp = LIB.with_suffix(".so.tmp%d" % os.getpid())
    cmd = [HIPCC, *FLAGS, f"-I{ROOT / 'include'}", f"-I{HERE / 'csrc'}", *map(str, SRCS), "-o",
           str(tmp)]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        if tmp.exists():
            tmp.unlink()
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_cxi(verbose=True)
