"""Benchmark: ERA iterations/s at 4096 positions x 256^2 frames on one MI355X.

Workload ("era-4096x256"): the production-shape scan -- 64 x 64 raster of
4096 positions, 40 px step with +-3 px jitter, 256 x 256 frames, object
2782 x 2782 (smoothed random amplitude/phase), pupil radius 64 px with defocus,
smooth background of 1 % of the mean intensity, Poisson noise at 1e8 photons
per frame, central beamstop masked.  Synthetic data, generated on the GPU by
the library's own frame-formation / FFT kernels plus torch.poisson.

A step is one ERA iteration (method 1: object update with the known probe)
over all 4096 frames, with every input resident in HBM.  ``value`` is the
median of --repeats timed passes of --steps iterations (the per-iteration
FP64 metric readback included); the per-stage kernel durations come from one
more pass of the same iterations with HIP events around every stage.  The same run also
times the difference map with background refinement (method 4) at the same
shape, reported under "extra", and the second half of the metric: wall time
to eMod <= 1e-3 of ERA method 3 on the held-out 32 x 32-position scan of
64^2 noise-free frames from the learned start (committed weights, trained by
``python -m ptychography_amd.train_init``) against the flat start, under
"learned_init".

"multi_scan" reports the batched path: the 16 seeded scans of
``sim.c4_specs()`` (frames 64-192^2, 240-1980 positions, 334 M frame pixels
per iteration), ERA method 3, as sequential single-scan calls and as one batch
with one hipGraph per iteration, inputs resident: scans of one frame size
packed into one engine (the default, "speedup") and unpacked
("speedup_unpacked").

``--copy-calibration`` only copies 1 GiB three times (the plain copy kernel
whose FETCH_SIZE / WRITE_SIZE counters scale the kernel traffic figures).

Prints ONE JSON line (rank 0, N=1).
"""

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# compulsory bytes (per frame pixel, per object pixel) of one ERA method-1
# iteration: amp read 4 + psi' write 8 + psi' read 8; object read/write 16 + FP64 heatmap 8
ERA1_BYTES_PX = 20
OBJ_BYTES_PX = 24
# DM methods 1-3: f read 8 + amp 4 + f write 8 + the overlap projection's read of f 8
DM1_BYTES_PX = 28
# DM method 4 (f, b, amp through two modulus steps + two overlap reads + metric amp read)
DM4_BYTES_PX = 76
# HIO (update 'sample'), compulsory per frame pixel: iterate read 8 + amplitude 4, the
# Psup argument written 8 and read by the object update 8, the next iterate written 8
HIO_BYTES_PX = 36
# what the HIO kernels move by design (row forward f -> m 16, columns 20, row inverse with
# both linear combinations 32, object gather 8): a diagnostic, not the denominator
HIO_KERNEL_BYTES_PX = 76

# algorithmic bytes per frame pixel of each stage's launch (ERA method 1)
STAGE_BYTES_PX_STAGED = {
    "rows_fwd": 8,    # write the row spectrum (object window and probe come from cache)
    "cols": 20,       # read spectrum 8 + amplitude 4, write 8
    "rows_inv": 16,   # read 8, write psi' 8
    "gather": 8,      # read psi' 8 (+ object pixel terms below)
}
# staged path with the inverse row FFT fused into the object update (the default
# for ERA method 1 after the first iteration; timed under "gather"): the row
# spectra are read once and psi' never goes back to HBM
STAGE_BYTES_PX_ROWG = {
    "rows_fwd": 8,
    "cols": 20,
    "gather": 8,      # read the row spectra 8 (+ object pixel terms below)
}
# fused per-frame kernel (timed under the "cols" stage): amplitude 4 in, psi' 8 out
STAGE_BYTES_PX_FUSED = {"cols": 12, "gather": 8}
STAGE_KERNEL = {"rows_fwd": "k_rows_fwd", "cols": "k_cols", "rows_inv": "k_rows_inv"}

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE,
# scaled x2 / x1 (what the counters read for a 1 GiB copy, --copy-calibration);
# summary in profiles/era_4096x256_pmc_traffic.md.  Key: (kernel, frames, side).
MEASURED_TRAFFIC = {
    ("k_rows_fwd", 4096, 256): 3.011e9,
    ("k_cols", 4096, 256): 5.371e9,
    ("k_rows_gather", 4096, 256): 2.589e9,
    ("k_rows_inv", 4096, 256): 4.295e9,
    ("k_gather_O", 4096, 256): 2.639e9,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed passes of --steps iterations; the median is reported")
    ap.add_argument("--dm-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=48,
                    help="frames in the bounded CPU-baseline sample")
    ap.add_argument("--grid", type=int, default=64, help="raster side (64 -> 4096 positions)")
    ap.add_argument("--path", choices=["default", "fused", "staged"], default="default")
    ap.add_argument("--frame-threads", type=int, default=0)
    ap.add_argument("--frame-slots", type=int, default=0)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-init timing")
    ap.add_argument("--no-multi", action="store_true", help="skip the multi-scan timing")
    ap.add_argument("--no-widening", action="store_true",
                    help="skip the HIO and back-projection timings")
    ap.add_argument("--multi-iters", type=int, default=150)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--nn", action="store_true",
                    help="time the learned-init network only (inference 1024 frames, "
                         "training steps of 64)")
    return ap.parse_args()


WEIGHTS = os.path.join(ROOT, "ptychography_amd", "weights", "init64.npz")


def copy_calibration():
    import torch
    x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")   # 1 GiB
    y = torch.empty_like(x)
    x.fill_(1.0)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 1 << 30, "copies": 3}))


def nn_bench(dev, reps=10):
    """Throughput of the init network on its own kernels: inference of 1024
    64^2 frames and training steps (forward + loss + backward + Adam) of 64."""
    import torch
    from ptychography_amd.learned import LAYERS, InitNet
    H, flops = 64, 0
    for l in LAYERS:
        if l[0] == "conv":
            flops += 2 * 9 * l[1] * l[2] * H * H
        else:
            H = H // 2 if l[0] == "pool" else H * 2
    net = InitNet(seed=1, device=dev)
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.rand((1024, 64, 64, 1), device=dev, generator=g).to(torch.bfloat16)
    for _ in range(2):
        net.forward(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(x)
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - t0) / reps
    xb = x[:64].contiguous()
    tb = torch.rand((64, 64, 64, 2), device=dev, generator=g)
    wt = torch.ones(64 * 64 * 2, device=dev)

    def step():
        pred = net.forward(xb, keep=True)
        _, dy = net.loss(pred, tb, wt, grad=True, gscale=1024.0, sync=False)
        net.backward(dy)
        net.adam(lr=1e-4, gscale=1024.0)

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / reps
    print(json.dumps({"nn_inference_ms_1024_frames": 1e3 * t_inf,
                      "nn_inference_tflops": 1024 * flops / t_inf / 1e12,
                      "nn_train_ms_per_step_64": 1e3 * t_step,
                      "nn_train_frames_per_s": 64 / t_step,
                      "nn_train_tflops": 3 * 64 * flops / t_step / 1e12,
                      "forward_gflop_per_frame": flops / 1e9}))


def multi_scan(dev, iters):
    """16 mixed scans, ERA method 3: sequential calls vs one batch with graphs."""
    import torch
    from ptychography_amd import ERA, ERA_many, sim
    scans = [sim.make_scan_gpu(sp, dev) for sp in sim.c4_specs()]
    px = sum(int(np.prod(sc["I"].shape)) for sc in scans)

    def job(sc):
        U, V = sc["O"].shape
        return dict(I=sc["I"], R=sc["R"].copy(), P=sc["P"].copy(),
                    O=np.ones((U, V), np.complex64), mask=sc["mask"], method=3)

    def sequential():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for sc in scans:
            j = job(sc)
            ERA(j.pop("I"), j.pop("R"), j.pop("P"), j.pop("O"), iters, device=dev,
                full_output=False, **j)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def batched(pack=True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ERA_many([job(sc) for sc in scans], iters, device=dev, full_output=False, graphs=True,
                 pack=pack)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    sequential(), batched(), batched(False)
    t_seq = min(sequential() for _ in range(2))
    t_bat = min(batched() for _ in range(2))
    t_unp = min(batched(False) for _ in range(2))
    return {"workload": "c4-16-scans-method3", "scans": len(scans), "frame_px_per_iter": px,
            "iters": iters, "sequential_s": t_seq, "batched_graphs_s": t_bat,
            "sequential_gpx_iter_per_s": px * iters / t_seq / 1e9,
            "batched_gpx_iter_per_s": px * iters / t_bat / 1e9, "speedup": t_seq / t_bat,
            "batched_unpacked_graphs_s": t_unp, "speedup_unpacked": t_seq / t_unp}


def learned_init(dev, target=1e-3, max_iters=1500, reps=5):
    """Wall time to eMod <= target: ERA method 3 from the learned start vs the
    flat start, held-out noise-free 32 x 32 scan of 64^2 frames, I resident;
    with the true probe, and under "approximate_probe" with the true probe
    times (1 + 0.3 x Gaussian noise) as the start both runs refine."""
    import torch
    from ptychography_amd import ERA, sim
    from ptychography_amd.learned import load_weights
    net = load_weights(WEIGHTS, device=dev)
    s = sim.make_scan("c5", noise=False)
    I_d = torch.from_numpy(s.I.astype(np.float32)).to(dev)
    U, V = s.O.shape
    ones = np.ones((U, V), np.complex64)

    def era(iters, init, full, P0):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = ERA(I_d, s.R.copy(), P0.copy(), ones.copy(), iters, mask=1, method=3, device=dev,
                init=init, init_net=net if init else None, full_output=full)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    def compare(P0):
        out = {}
        for name, init in (("plain", None), ("learned", "learned")):
            _, (O, P, info) = era(max_iters, init, True, P0)
            e = np.asarray(info["eMod"])
            hit = np.nonzero(e <= target)[0]
            k = int(hit[0]) + 1 if len(hit) else None
            out[f"{name}_iters"] = k
            out[f"{name}_start_eMod"] = float(e[0])
            if name == "learned":
                out["init"] = info.get("init")
            if k is not None:
                era(k, init, False, P0)
                out[f"{name}_wall_s"] = float(np.median([era(k, init, False, P0)[0]
                                                          for _ in range(reps)]))
        if out.get("plain_wall_s") and out.get("learned_wall_s"):
            out["speedup"] = out["plain_wall_s"] / out["learned_wall_s"]
        return out

    out = {"workload": "c5-heldout-1024x64x64-method3", "target_eMod": target}
    out.update(compare(s.P))
    P_apx = s.P * (1.0 + 0.3 * np.random.default_rng(5).standard_normal(s.P.shape))
    out["approximate_probe"] = compare(P_apx)
    return out


def make_scan_gpu(grid, dev):
    """The C3 recipe at full size (``sim.make_scan_gpu``, the generator the
    full-size tests use), intensities formed on the GPU."""
    from ptychography_amd import sim
    spec = sim.RECIPES["c3"].with_(grid=(grid, grid))
    sc = sim.make_scan_gpu(spec, dev)
    N, M = sc["I"].shape[:2]
    U, V = sc["O"].shape
    mask = sc["mask"]
    if np.ndim(mask) == 0:
        mask = np.ones((M, M), bool)
    sc.update(R=sc["R"].astype(np.int32), mask=mask, N=N, M=M, U=U, V=V)
    return sc


def cpu_model():
    """(model name, logical CPUs) of this host."""
    name = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def widening(scan, dev, iters=6):
    """HIO (update 'sample', beta 0.9, flat start) and back projection on the same
    scan.  HIO's time per iteration is the difference of two calls (iters and
    2 iters, the fastest of three each), which cancels the set-up and the copy
    of O and P back to the host;
    back projection is timed as a whole call (set-up, one pass, O to the host)."""
    import torch
    import ptychography_amd as pty
    N, M, U, V = scan["N"], scan["M"], scan["U"], scan["V"]
    O0 = np.ones((U, V), np.complex64)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def hio(k):
        return timed(lambda: pty.HIO(scan["I"], scan["R"].copy(), scan["P"], O0, k, beta=0.9,
                                     mask=scan["mask"], full_output=False, device=dev))

    hio(1)
    t1 = min(hio(iters) for _ in range(3))
    t2 = min(hio(2 * iters) for _ in range(3))
    per = (t2 - t1) / iters
    bp = [timed(lambda: pty.Back_projection(scan["I"], scan["R"].copy(), scan["P"], O0,
                                            mask=scan["mask"], full_output=False, device=dev))
          for _ in range(2)]
    px = N * M * M
    return {"hio_workload": f"hio-sample-beta0.9-{N}x{M}x{M}-object{U}x{V}",
            "hio_ms_per_iter": 1e3 * per, "hio_iters_per_s": 1.0 / per,
            "hio_roofline_frac": (HIO_BYTES_PX * px + OBJ_BYTES_PX * U * V) / per / (HBM_PEAK_GBS * 1e9),
            "hio_kernel_bytes_frac": (HIO_KERNEL_BYTES_PX * px + OBJ_BYTES_PX * U * V) / per
                                     / (HBM_PEAK_GBS * 1e9),
            "back_projection_call_s": bp[-1]}


def bind(eng, scan, dev, O0, P0):
    import torch
    mask_c = torch.from_numpy(scan["mask"].astype(np.uint8)).to(dev)
    mask_s = torch.from_numpy(np.fft.ifftshift(scan["mask"]).astype(np.uint8)).to(dev)
    amp, I_norm = eng.prepare_amp(scan["I"], mask_c)
    eng.set_data(amp, mask_s, I_norm)
    O_t = torch.from_numpy(O0.astype(np.complex64)).to(dev)
    P_t = torch.from_numpy(P0.astype(np.complex64)).to(dev)
    eng.set_state(O_t, P_t)
    return O_t, P_t, amp


def cpu_baseline(scan, nframes):
    """Oracle ERA (method 1, complex128, reference-faithful make_exits) on the first
    `nframes` frames with the full-size object; per-iteration time scaled to all frames."""
    import oracle
    R = scan["R"][:nframes].astype(np.int64)

    def timed(real, cplx):
        # the oracle follows the reference's dtype rule: float64 I -> complex128
        # arithmetic, float32 I -> complex64
        I = scan["I"][:nframes].cpu().numpy().astype(real)
        O0 = np.ones((scan["U"], scan["V"]), cplx)
        stamps = []
        t0 = time.perf_counter()
        oracle.ERA(I, R.copy(), scan["P"].astype(cplx), O0, 2, method=1,
                   mask=scan["mask"], full_output=False,
                   callback=lambda i, O, P: stamps.append(time.perf_counter()))
        return stamps[-1] - stamps[-2], time.perf_counter() - t0

    per_iter, total = timed(np.float64, np.complex128)
    per_iter64, total64 = timed(np.float32, np.complex64)
    scale = scan["N"] / nframes
    return {"value": 1.0 / (per_iter * scale), "unit": "iters/s", "cores": 1, "kind": "port",
            "value_complex64": 1.0 / (per_iter64 * scale),
            "sample": (f"oracle.ERA method 1, complex128 (value) and complex64 (value_complex64), "
                       f"{nframes} of {scan['N']} frames (256^2) with the full "
                       f"{scan['U']}x{scan['V']} object, 2 iterations each ({total:.1f} s and "
                       f"{total64:.1f} s wall); second iteration {per_iter:.3f} s / "
                       f"{per_iter64:.3f} s scaled by {scan['N']}/{nframes}; host "
                       f"{cpu_model()[0]}, {cpu_model()[1]} logical CPUs, one used (NumPy's FFT "
                       f"is single-threaded)")}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; multi-GPU sharding "
                                   "is a design note only (--gpus must be 1)"}))
        sys.exit(2)
    import torch
    from ptychography_amd.engine import Engine
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.nn:
        nn_bench(dev)
        return

    t_gen = time.perf_counter()
    scan = make_scan_gpu(args.grid, dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    N, M, U, V = scan["N"], scan["M"], scan["U"], scan["V"]
    px = N * M * M

    # ---- ERA method 1 (headline) ----
    eng = Engine(N, M, M, U, V, scan["R"], device=dev)
    if args.path != "default" or args.frame_threads or args.frame_slots:
        eng.set_path(args.path != "staged", args.frame_threads, args.frame_slots)
    O_t, P_t, amp = bind(eng, scan, dev, np.ones((U, V), np.complex64), scan["P"])
    eng.run(0, 1, args.warmup, iter0=0, begin=True)
    it0, passes = args.warmup, []
    for _ in range(max(1, args.repeats)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        metrics = eng.run(0, 1, args.steps, iter0=it0)
        torch.cuda.synchronize()
        passes.append(time.perf_counter() - t0)
        it0 += args.steps
    dt = float(np.median(passes))
    # kernel durations: one more pass of the same iterations with a HIP event
    # pair around every stage on the engine's stream
    eng.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(0, 1, args.steps, iter0=it0)
    torch.cuda.synchronize()
    dt_prof = time.perf_counter() - t0
    prof = eng.profile_read()
    eng.profile(False)
    eng.close()
    del amp, O_t, P_t

    its = args.steps / dt
    fused = prof["rows_fwd"][1] == 0
    rowg = not fused and prof["rows_inv"][1] == 0
    STAGE_BYTES_PX = (STAGE_BYTES_PX_FUSED if fused else
                      STAGE_BYTES_PX_ROWG if rowg else STAGE_BYTES_PX_STAGED)
    stage_avg = {k: (ms / n if n else 0.0) for k, (ms, n) in prof.items()}
    dom = max((k for k in prof if k in STAGE_BYTES_PX), key=lambda k: prof[k][0])
    dom_bytes = STAGE_BYTES_PX[dom] * px + (OBJ_BYTES_PX * U * V if dom == "gather" else 0)
    dom_gbs = dom_bytes / (stage_avg[dom] * 1e-3) / 1e9
    dom_kernel = ("k_frame (fused rows+columns+modulus)" if fused and dom == "cols" else
                  STAGE_KERNEL.get(dom, "k_rows_gather" if rowg else "k_gather_O"))
    iter_bytes = ERA1_BYTES_PX * px + OBJ_BYTES_PX * U * V
    # 5 n log2 n per 2-D transform of n = M^2 points, one forward and one inverse per frame
    fft_flops = 2 * N * 5.0 * M * M * np.log2(M * M)

    # ---- difference map at the same shape: with background refinement
    # (method 4, the workload's own algorithm), object-only (method 1) and
    # object + probe (method 3) with OP_iters (1, 1) and (1, 3) ----
    def dm_timing(method, OP_iters=(1, 1)):
        eng = Engine(N, M, M, U, V, scan["R"], device=dev)
        if args.path != "default" or args.frame_threads or args.frame_slots:
            eng.set_path(args.path != "staged", args.frame_threads, args.frame_slots)
        O_t, P_t, amp = bind(eng, scan, dev, np.ones((U, V), np.complex64), scan["P"])
        if method >= 4:
            b0 = np.sqrt(np.fft.ifftshift(scan["B"])).astype(np.float32)
            eng.set_background(torch.from_numpy(b0), per_frame=False)
        eng.run(1, method, 2, iter0=0, begin=True, OP_iters=OP_iters)
        it, times = 2, []
        for _ in range(max(1, args.repeats)):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            m = eng.run(1, method, args.dm_steps, iter0=it, OP_iters=OP_iters)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t1)
            it += args.dm_steps
        eng.close()
        del amp, O_t, P_t
        per = float(np.median(times)) / args.dm_steps
        return per, float(m[-1, 0])

    # ERA method 3 (object + probe, the algorithm of the 128^2 and mixed-scan
    # workloads) at the same shape, from the true probe
    era3 = None
    if args.dm_steps > 0:
        eng = Engine(N, M, M, U, V, scan["R"], device=dev)
        O_t, P_t, amp = bind(eng, scan, dev, np.ones((U, V), np.complex64), scan["P"])
        eng.run(0, 3, 2, iter0=0, begin=True)
        it, times = 2, []
        for _ in range(max(1, args.repeats)):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            eng.run(0, 3, args.dm_steps, iter0=it)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t1)
            it += args.dm_steps
        eng.close()
        del amp, O_t, P_t
        per = float(np.median(times)) / args.dm_steps
        era3 = {"era3_iters_per_s": 1.0 / per, "era3_ms_per_step": 1e3 * per}

    dm = None
    if args.dm_steps > 0:
        per4, e4 = dm_timing(4)
        per1, e1 = dm_timing(1)
        per3, e3 = dm_timing(3, (1, 3))
        per31, e31 = dm_timing(3, (1, 1))
        dm = {"dm_bg_iters_per_s": 1.0 / per4, "dm_bg_ms_per_step": 1e3 * per4,
              "dm_bg_roofline_frac": (DM4_BYTES_PX * px + OBJ_BYTES_PX * U * V) / per4 / (HBM_PEAK_GBS * 1e9),
              "dm_bg_final_eMod": e4,
              "dm1_iters_per_s": 1.0 / per1, "dm1_ms_per_step": 1e3 * per1,
              "dm1_roofline_frac": (DM1_BYTES_PX * px + OBJ_BYTES_PX * U * V) / per1 / (HBM_PEAK_GBS * 1e9),
              "dm3_op13_iters_per_s": 1.0 / per3, "dm3_op13_ms_per_step": 1e3 * per3,
              "dm3_iters_per_s": 1.0 / per31, "dm3_ms_per_step": 1e3 * per31,
              "repeats": max(1, args.repeats), "timing": "median of the repeats"}
        dm.update(era3 or {})

    learned = None
    if not args.no_learned and os.path.exists(WEIGHTS):
        learned = learned_init(dev)
    multi = None
    if not args.no_multi:
        multi = multi_scan(dev, args.multi_iters)

    widen = None
    if not args.no_widening:
        widen = widening(scan, dev)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scan, args.cpu_frames)

    out = {
        "metric": "ERA/DM iters/s (4096x256^2 frames); time-to-eMod 1e-3, learned init vs plain",
        "value": its,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded smoothed-random object, defocused pupil probe, Poisson noise)",
        "config": {"workload": f"era-method1-{N}x{M}x{M}-object{U}x{V}", "frames": N,
                   "frame": [M, M], "object": [U, V], "method": 1, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS,
                     "traffic": None if fused else MEASURED_TRAFFIC.get((dom_kernel, N, M)),
                     "kernel": dom_kernel,
                     "kernel_ms": stage_avg[dom],
                     "algorithmic_bytes_per_launch": dom_bytes},
        "iteration_roofline_frac": iter_bytes * its / (HBM_PEAK_GBS * 1e9),
        "repeats": max(1, args.repeats),
        "ms_per_step_passes": [1e3 * t / args.steps for t in passes],
        "ms_per_step_profiled_pass": 1e3 * dt_prof / args.steps,
        "fft_gflops": fft_flops * its / 1e9,
        "fft_flop_per_iter": fft_flops,
        "stage_ms_per_launch": stage_avg,
        "object_update": "k_rows_gather (inverse row FFT fused)" if rowg else "k_gather_O",
        "final_eMod": float(metrics[-1, 0]),
        "gen_s": t_gen,
        "cpu_baseline": cpu,
        "extra": dm,
        "learned_init": learned,
        "multi_scan": multi,
        "widening": widen,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
