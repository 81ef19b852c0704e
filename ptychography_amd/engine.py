"""Device-side engine: one ptyhip context plus the torch tensors it works on.

PyTorch is used only to own device memory and to name the stream; all
arithmetic of the iteration runs in the HIP kernels of ``libptyhip.so``.
"""

import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import PtyTuneResult, PtyOpts, call


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def require_gpu(device=None):
    """Return a ROCm torch device or raise: the ERA / DM path has no CPU fallback."""
    _abi.load()
    if not torch.cuda.is_available():
        raise RuntimeError("ptychography_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "the ERA/DM iteration has no CPU fallback")
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise ValueError(f"device must be a GPU device, got {dev}")
    return dev


# Path override for every new Engine: None = library default (the staged
# kernels), False = staged kernels, True = the fused per-frame kernel where the
# frame shape supports it.
FUSED = None


class Engine:
    """A scan geometry on one GPU: frames, object window offsets and workspace.

    ``R`` follows the reference convention AFTER its preamble shift
    (ptychography/era.py:504-505): ``R <= 0`` and frame n sees
    ``O[-R[n,0]:M-R[n,0], -R[n,1]:K-R[n,1]]``.
    """

    def __init__(self, N, M, K, U, V, R, device=None):
        self.device = require_gpu(device)
        self.N, self.M, self.K, self.U, self.V = int(N), int(M), int(K), int(U), int(V)
        Ra = np.asarray(R)
        if np.issubdtype(Ra.dtype, np.floating) and np.any(Ra != np.rint(Ra)):
            raise ValueError("window positions must be integers; fractional parts go through "
                             "set_subpixel")
        Rc = np.ascontiguousarray(Ra.astype(np.int32).reshape(self.N, 2))
        self._R = Rc
        self._ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            call("pty_create", ctypes.byref(self._ctx), self.N, self.M, self.K, self.U, self.V,
                 Rc.ctypes.data_as(ctypes.c_void_p))
            stream = torch.cuda.current_stream(self.device)
            call("pty_set_stream", self._ctx, ctypes.c_void_p(stream.cuda_stream))
        self._keep = {}
        self.S = 1
        if FUSED is not None:
            self.set_path(FUSED)

    @property
    def R(self):
        """The (N, 2) int32 window offsets (R <= 0) the engine was built with."""
        return self._R

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if self._ctx:
            _abi.load().pty_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_torch(self):
        # the context runs on torch's current stream; make prior torch work visible
        torch.cuda.current_stream(self.device).synchronize()

    # -- binding ------------------------------------------------------------
    def set_scans(self, frame_start, row_start):
        """Pack scans: scan s owns frames [frame_start[s], frame_start[s+1]) and
        the object rows [row_start[s], row_start[s+1]) (``pty_set_scans``).
        Afterwards masks and the probe are per scan ([S, M, K]) and runs
        return (iters, S, 3) metrics.  Unbinds the data."""
        fs = np.ascontiguousarray(frame_start, dtype=np.int32)
        rs = np.ascontiguousarray(row_start, dtype=np.int32)
        if fs.shape != rs.shape or fs.ndim != 1 or len(fs) < 2:
            raise ValueError("frame_start and row_start: S + 1 entries each")
        self._sync_torch()
        call("pty_set_scans", self._ctx, len(fs) - 1, fs.ctypes.data_as(ctypes.c_void_p),
             rs.ctypes.data_as(ctypes.c_void_p), None)
        self.S = len(fs) - 1
        self._keep.pop("amp", None)
        self._keep.pop("mask", None)

    def prepare_amp(self, I, mask_centred=None):
        """amp = ifftshift(sqrt(I)), I_norm = sum(mask * I) (ptychography/era.py:495-498).
        Packed scans: mask_centred is [S, M, K] and I_norm a list of S sums."""
        I = I.to(self.device, torch.float32).contiguous()
        amp = torch.empty_like(I)
        m = None
        if mask_centred is not None:
            m = mask_centred.to(self.device, torch.uint8).contiguous()
        self._sync_torch()
        out = (ctypes.c_double * self.S)()
        call("pty_prepare_amp", self._ctx, _ptr(I), _ptr(m), _ptr(amp),
             ctypes.cast(out, ctypes.POINTER(ctypes.c_double)))
        if self.S > 1:
            return amp, [float(v) for v in out]
        return amp, float(out[0])

    def set_data(self, amp, mask_shifted, I_norm):
        self._keep["amp"] = amp
        self._keep["mask"] = mask_shifted
        call("pty_set_data", self._ctx, _ptr(amp), _ptr(mask_shifted), float(I_norm))

    def set_state(self, O, P):
        assert O.dtype == torch.complex64 and P.dtype == torch.complex64
        assert O.is_contiguous() and P.is_contiguous()
        pshape = (self.M, self.K) if self.S == 1 else (self.S, self.M, self.K)
        assert tuple(O.shape) == (self.U, self.V) and tuple(P.shape) == pshape
        self._keep["O"], self._keep["P"] = O, P
        self._sync_torch()
        call("pty_set_state", self._ctx, _ptr(O), _ptr(P))

    def set_background(self, b, per_frame):
        b = b.to(self.device, torch.float32).contiguous()
        self._sync_torch()
        call("pty_set_background", self._ctx, _ptr(b), int(bool(per_frame)))

    def set_pamp(self, pamp=None):
        if pamp is not None:
            pamp = pamp.to(self.device, torch.float32).contiguous()
            self._sync_torch()
        call("pty_set_pamp", self._ctx, _ptr(pamp))

    def set_subpixel(self, d=None):
        """Fractional position parts d (N x 2: (y, x) per frame, |d| <= 1):
        frame n sees the probe Fourier-shifted by d[n]; None = integer positions."""
        if d is None:
            call("pty_set_subpixel", self._ctx, None)
            return
        d = np.ascontiguousarray(d, dtype=np.float64)
        if d.shape != (self.N, 2):
            raise ValueError(f"sub-pixel shifts must have shape {(self.N, 2)}, got {d.shape}")
        self._sync_torch()
        call("pty_set_subpixel", self._ctx, d.ctypes.data_as(ctypes.c_void_p))

    def position_terms(self):
        """(N, 6) float64 per frame: modulus error f_n, its gradient in the
        frame's shift (y, x) and the Gauss-Newton matrix (yy, yx, xx)."""
        out = np.zeros((self.N, 6), dtype=np.float64)
        self._sync_torch()
        call("pty_position_terms", self._ctx, out.ctypes.data_as(ctypes.c_void_p))
        return out

    # -- running ------------------------------------------------------------
    @staticmethod
    def opts(algorithm, method, OP_iters=(1, 1), Pmod_probe_until=0, probe_centering=False,
             alpha=1e-10):
        """The ``pty_opts`` of one run."""
        return PtyOpts(algorithm=int(algorithm), method=int(method), op_n=int(OP_iters[0]),
                       op_every=int(OP_iters[1]), pmod_probe_until=int(Pmod_probe_until),
                       probe_centering=int(bool(probe_centering)), alpha=float(alpha))

    def run(self, algorithm, method, iters, iter0=0, begin=False, OP_iters=(1, 1),
            Pmod_probe_until=0, probe_centering=False, alpha=1e-10, metrics=True):
        """Run ``iters`` iterations; returns an (iters, 3) array of eMod, eCon, eSup
        ((iters, S, 3) for packed scans)."""
        o = self.opts(algorithm, method, OP_iters, Pmod_probe_until, probe_centering, alpha)
        shape = (max(iters, 0), 3) if self.S == 1 else (max(iters, 0), self.S, 3)
        out = np.zeros(shape, dtype=np.float64) if metrics else None
        ptr = out.ctypes.data_as(ctypes.c_void_p) if (metrics and iters > 0) else None
        call("pty_run", self._ctx, ctypes.byref(o), int(iter0), int(iters), int(bool(begin)), ptr)
        return out

    def set_launch(self, rows_mult=3, cols_mult=0, cols_wg=0):
        """Persistent grid sizes of the staged row / column kernels (x CUs; 0 =
        per-kernel default) and threads per column work-group (256, 512 for
        256^2 frames, or 0 = automatic)."""
        call("pty_set_launch", self._ctx, int(rows_mult), int(cols_mult), int(cols_wg))

    def set_chunk(self, frames):
        """ERA modulus step over blocks of `frames` frames (0 = all at once)."""
        call("pty_set_chunk", self._ctx, int(frames))

    def set_row_gather(self, enable=True):
        """Fuse the inverse row FFT into ERA's object update and into DM-with-background's
        leading projection (where supported)."""
        call("pty_set_row_gather", self._ctx, int(bool(enable)))

    def row_gather_status(self):
        """(geometry supports the fused row gather, currently enabled)."""
        sup, en = ctypes.c_int32(), ctypes.c_int32()
        call("pty_get_row_gather", self._ctx, ctypes.byref(sup), ctypes.byref(en))
        return bool(sup.value), bool(en.value)

    def tune(self, algorithm, method, trial_iters=3, OP_iters=(1, 1), Pmod_probe_until=0,
             probe_centering=False, alpha=1e-10):
        """Time every launch configuration on this scan and keep the fastest
        (state restored; the next run must pass ``begin=True``)."""
        o = self.opts(algorithm, method, OP_iters, Pmod_probe_until, probe_centering, alpha)
        r = PtyTuneResult()
        call("pty_tune", self._ctx, ctypes.byref(o), int(trial_iters), ctypes.byref(r))
        return {f: getattr(r, f) for f, _ in PtyTuneResult._fields_}

    PLAN_ROWS, PLAN_COLS, PLAN_ROWG, PLAN_BGERR = 1, 2, 4, 8

    def set_plan(self, mask):
        """Line-FFT plans per stage (``pty_set_plan``): a set bit (PLAN_ROWS,
        PLAN_COLS, PLAN_ROWG, PLAN_BGERR) selects the side's alternative radix
        split for the row kernels, the column kernels, the fused row gather or
        DM-with-background's model-error pass."""
        call("pty_set_plan", self._ctx, int(mask))

    def plan(self):
        """(current plan mask, the frame has alternative plans)."""
        v, alt = ctypes.c_int32(), ctypes.c_int32()
        call("pty_get_plan", self._ctx, ctypes.byref(v), ctypes.byref(alt))
        return int(v.value), bool(alt.value)

    def set_path(self, fused, threads=0, slots=0):
        """Select the fused per-frame kernel (True) or the staged kernels (False)."""
        call("pty_set_path", self._ctx, int(bool(fused)), int(threads), int(slots))

    def set_graphs(self, enable=True):
        """Replay each kind of iteration as a captured hipGraph."""
        call("pty_set_graphs", self._ctx, int(bool(enable)))

    def graph_status(self):
        """(graphs on, a capture failed and the run fell back to eager launches)."""
        en, fb = ctypes.c_int32(), ctypes.c_int32()
        call("pty_get_graphs", self._ctx, ctypes.byref(en), ctypes.byref(fb))
        return bool(en.value), bool(fb.value)

    def profile(self, enable=True):
        call("pty_profile", self._ctx, int(bool(enable)))

    def profile_read(self):
        """{stage: (total_ms, launches)} since profiling was enabled."""
        ms = np.zeros(len(_abi.STAGES), np.float64)
        cnt = np.zeros(len(_abi.STAGES), np.int64)
        call("pty_profile_read", self._ctx, ms.ctypes.data_as(ctypes.c_void_p),
             cnt.ctypes.data_as(ctypes.c_void_p))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(_abi.STAGES)}

    def intensity(self):
        out = torch.empty((self.N, self.M, self.K), dtype=torch.float32, device=self.device)
        call("pty_intensity", self._ctx, _ptr(out))
        return out

    def heatmap(self):
        H = torch.empty((self.U, self.V), dtype=torch.float64, device=self.device)
        valid = ctypes.c_int32()
        call("pty_heatmap", self._ctx, _ptr(H), ctypes.byref(valid))
        return H if valid.value else None

    def background(self):
        b = torch.empty((self.M, self.K), dtype=torch.float32, device=self.device)
        call("pty_background", self._ctx, _ptr(b))
        return b

    # -- resuming from a recorded state ---------------------------------------
    def set_heatmap(self, H=None, G=None):
        """Cached P heatmap H (U x V) and O heatmap G (M x K), float64; None = not cached."""
        H = None if H is None else H.to(self.device, torch.float64).contiguous()
        G = None if G is None else G.to(self.device, torch.float64).contiguous()
        self._sync_torch()
        call("pty_set_heatmap", self._ctx, _ptr(H), _ptr(G))

    def set_iterate(self, f=None, O_ref=None, P_ref=None):
        """DM iterate f (N x M x K complex64) and the eCon reference (previous
        metric object and probe; None = the bound O and P)."""
        ts = [None if t is None else t.to(self.device, torch.complex64).contiguous()
              for t in (f, O_ref, P_ref)]
        self._sync_torch()
        call("pty_set_iterate", self._ctx, *[_ptr(t) for t in ts])

    def get_iterate(self, f=True, background=False):
        """(f, per-frame background) copies; either may be skipped (None)."""
        ft = torch.empty((self.N, self.M, self.K), dtype=torch.complex64, device=self.device) \
            if f else None
        bt = torch.empty((self.N, self.M, self.K), dtype=torch.float32, device=self.device) \
            if background else None
        call("pty_get_iterate", self._ctx, _ptr(ft), _ptr(bt))
        return ft, bt

    # -- building blocks ----------------------------------------------------
    def fft2(self, data, inverse=False):
        """In-place batched 2-D FFT of a (n, M, K) complex64 tensor."""
        assert data.dtype == torch.complex64 and data.is_contiguous()
        assert tuple(data.shape[1:]) == (self.M, self.K)
        self._sync_torch()
        call("pty_fft2", self._ctx, _ptr(data), int(data.shape[0]), int(bool(inverse)))
        return data

    def make_exits(self, O, P):
        out = torch.empty((self.N, self.M, self.K), dtype=torch.complex64, device=self.device)
        self._sync_torch()
        call("pty_make_exits", self._ctx, _ptr(O), _ptr(P), _ptr(out))
        return out

    def pmod1(self, exits, alpha=1e-10):
        self._sync_torch()
        e = ctypes.c_double()
        call("pty_pmod1", self._ctx, _ptr(exits), float(alpha), ctypes.byref(e))
        return exits, e.value

    def psup_O(self, exits, P, H=None, alpha=1e-10):
        O = torch.empty((self.U, self.V), dtype=torch.complex64, device=self.device)
        fresh = H is None
        if H is None:
            H = torch.empty((self.U, self.V), dtype=torch.float64, device=self.device)
        self._sync_torch()
        call("pty_psup_O", self._ctx, _ptr(exits), _ptr(P), _ptr(O), _ptr(H), int(fresh), float(alpha))
        return O, H

    def hio_modulus(self, F, m, O, P, w, beta, alpha=1e-10):
        """One HIO modulus step on (N, M, K) exits f = F + w make_exits(O, P):
        m = Pmod(f), then in the same pass m <- (1 + 1/beta) m - f / beta and
        F <- f - beta m.  Returns the un-normalised eMod sum of f."""
        for t in (F, m):
            assert t.dtype == torch.complex64 and t.is_contiguous()
            assert tuple(t.shape) == (self.N, self.M, self.K)
        assert F.data_ptr() != m.data_ptr()
        assert tuple(O.shape) == (self.U, self.V) and tuple(P.shape) == (self.M, self.K)
        self._sync_torch()
        e = ctypes.c_double()
        call("pty_hio_modulus", self._ctx, _ptr(F), _ptr(m), _ptr(O), _ptr(P), float(w), float(beta),
             float(alpha), ctypes.byref(e))
        return e.value

    def add_exits(self, f, O, P, scale):
        """f += scale * make_exits(O, P) in place."""
        assert f.dtype == torch.complex64 and f.is_contiguous()
        assert tuple(f.shape) == (self.N, self.M, self.K)
        assert tuple(O.shape) == (self.U, self.V) and tuple(P.shape) == (self.M, self.K)
        self._sync_torch()
        call("pty_add_exits", self._ctx, _ptr(O), _ptr(P), float(scale), _ptr(f))
        return f

    def back_projection(self, P, alpha=1e-10, exits=False):
        """Back-projection object from the bound data: (O, H, exits or None)."""
        O = torch.empty((self.U, self.V), dtype=torch.complex64, device=self.device)
        H = torch.empty((self.U, self.V), dtype=torch.float64, device=self.device)
        ex = torch.empty((self.N, self.M, self.K), dtype=torch.complex64, device=self.device) \
            if exits else None
        self._sync_torch()
        call("pty_back_projection", self._ctx, _ptr(P), _ptr(ex), _ptr(O), _ptr(H), float(alpha))
        return O, H, ex

    def psup_P(self, exits, O, G=None, alpha=1e-10):
        P = torch.empty((self.M, self.K), dtype=torch.complex64, device=self.device)
        fresh = G is None
        if G is None:
            G = torch.empty((self.M, self.K), dtype=torch.float64, device=self.device)
        self._sync_torch()
        call("pty_psup_P", self._ctx, _ptr(exits), _ptr(O), _ptr(P), _ptr(G), int(fresh), float(alpha))
        return P, G
