"""Packed scans: several scans of one frame size in one engine, every stage of
an iteration one launch over all of their frames (``Engine.set_scans``,
``pty_set_scans``; ``ERA_many`` packs by default).

Per scan the packed run must give the object and probe of that scan run
alone bit for bit (the kernels sum the same values in the same order: the
object gather per pixel over ascending frames, the probe over the chunks a
context of that scan alone would cut), ``info['I']`` and the heatmap bit for
bit, and eMod / eCon / eSup to FP64 rounding (partials summed in another
order).  Scans that cannot be packed run unpacked in the same batch."""

import numpy as np
import pytest
import torch

import oracle
from ptychography_amd import ERA, ERA_many
from ptychography_amd._abi import PtyError
from ptychography_amd.batch import pack_key
from ptychography_amd.engine import Engine
from scans import broadband_start, rel_l2, small_scan

pytestmark = pytest.mark.gpu

# four scans of 64^2 frames: different position counts, objects, probes and
# masks (one without), so that each scan's probe, mask and I_norm matter
SAME = [dict(frame=64, grid=(4, 5), step=16, seed=31, beamstop_radius=3.0, dead_frac=0.05),
        dict(frame=64, grid=(3, 3), step=20, seed=32),
        dict(frame=64, grid=(5, 4), step=14, seed=33, beamstop_radius=4.0, dead_frac=0.08),
        dict(frame=64, grid=(2, 6), step=18, seed=34, dead_frac=0.03)]


def _jobs(method, specs=SAME, **extra):
    out = []
    for spec in specs:
        s = small_scan(**spec)
        O0 = broadband_start(s.O.shape, seed=spec["seed"])
        P0 = s.P * (1.0 + 0.05 * broadband_start(s.P.shape, seed=spec["seed"] + 100))
        kw = dict(I=s.I, R=s.R.copy(), P=P0, O=O0, method=method)
        if "beamstop_radius" in spec or "dead_frac" in spec:
            kw["mask"] = s.mask
        kw.update(extra)
        out.append((s, kw))
    return out


def _alone(kw, iters, dev):
    kw = dict(kw)
    return ERA(kw.pop("I"), kw.pop("R").copy(), kw.pop("P"), kw.pop("O"), iters, device=dev, **kw)


def _metrics(info):
    return np.stack([info["eMod"], info["eCon"], info["eSup"]], axis=1)


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("method", [1, 2, 3])
def test_packed_scans_equal_single_scans(method, graphs, gpu_device):
    jobs = _jobs(method)
    keys = {pack_key(0, kw) for _, kw in jobs}
    assert len(keys) == 1 and None not in keys     # one pack of four
    iters = 10
    print(f"packed: 4 scans of 64^2, ERA method {method}, graphs={graphs}", flush=True)
    res = ERA_many([dict(kw, R=kw["R"].copy()) for _, kw in jobs], iters, graphs=graphs,
                   device=gpu_device)
    for k, ((s, kw), r) in enumerate(zip(jobs, res)):
        single = _alone(kw, iters, gpu_device)
        assert len(r) == len(single)
        for a, b in zip(r[:-1], single[:-1]):
            assert a.shape == b.shape and np.array_equal(a, b), f"scan {k}: iterate differs"
        info, ref = r[-1], single[-1]
        assert np.array_equal(info["I"], ref["I"]), f"scan {k}: info['I'] differs"
        assert np.array_equal(info["heatmap"], ref["heatmap"]), f"scan {k}: heatmap differs"
        m, mr = _metrics(info), _metrics(ref)
        err = np.abs(m - mr) / np.abs(mr)
        print(f"  scan {k}: metrics max rel diff {err.max():.1e}", flush=True)
        assert err.max() < 1e-12, f"scan {k}: metrics differ by {err.max():.2e}"


def test_packed_scans_match_the_oracle(gpu_device):
    jobs = _jobs(3)
    iters = 6
    res = ERA_many([dict(kw, R=kw["R"].copy()) for _, kw in jobs], iters, device=gpu_device)
    for k, ((s, kw), (O_g, P_g, info)) in enumerate(zip(jobs, res)):
        O_r, P_r, info_r = oracle.ERA(kw["I"], kw["R"].copy(), kw["P"], kw["O"], iters, method=3,
                                      mask=kw.get("mask", 1))
        e, er = np.asarray(info["eMod"]), np.asarray(info_r["eMod"])
        H = info_r["heatmap"]
        cov = H > 1e-3 * H.max()
        print(f"scan {k}: eMod {np.max(np.abs(e - er) / er):.1e}, O {rel_l2(O_g[cov], O_r[cov]):.1e}, "
              f"P {rel_l2(P_g, P_r):.1e}", flush=True)
        assert np.max(np.abs(e - er) / er) < 1e-4
        assert rel_l2(O_g[cov], O_r[cov]) < 1e-4
        assert rel_l2(P_g, P_r) < 1e-4


def test_unpackable_scans_run_beside_the_pack(gpu_device):
    # scan 1 asks for probe centring: it runs in an engine of its own, the
    # other three of the same frame size are packed; everyone equals its run alone
    jobs = _jobs(3)
    jobs[1][1]["probe_centering"] = True
    assert pack_key(0, jobs[1][1]) is None
    iters = 5
    res = ERA_many([dict(kw, R=kw["R"].copy()) for _, kw in jobs], iters, graphs=True,
                   device=gpu_device)
    for (s, kw), r in zip(jobs, res):
        single = _alone(kw, iters, gpu_device)
        for a, b in zip(r[:-1], single[:-1]):
            assert np.array_equal(a, b)


def test_pack_then_unpack_shift_r_in_place(gpu_device):
    jobs = _jobs(1)
    Rs = [kw["R"] for _, kw in jobs]
    before = [R.copy() for R in Rs]
    ERA_many([kw for _, kw in jobs], 1, device=gpu_device)
    for R, R0 in zip(Rs, before):
        assert np.array_equal(R, R0 - R0.max(axis=0))


def test_set_scans_rejects_bad_layouts(gpu_device):
    s = small_scan(frame=32, grid=(3, 3), step=8, seed=35)
    R = s.R - s.R.max(axis=0)
    N, M, K = s.I.shape
    U, V = s.O.shape
    Ua = -(-U // 16) * 16
    # two copies of the scan, the second one band lower
    Rp = np.concatenate([R, R - np.array([[Ua, 0]])]).astype(np.int32)
    e = Engine(2 * N, M, K, 2 * Ua, V, Rp, device=gpu_device)
    try:
        with pytest.raises(PtyError):   # a band that does not start on a tile row
            e.set_scans([0, N, 2 * N], [0, Ua - 1 if Ua > 16 else 8, 2 * Ua])
        with pytest.raises(PtyError):   # frames of scan 0 outside its band
            e.set_scans([0, N + 1, 2 * N], [0, Ua, 2 * Ua])
        with pytest.raises(PtyError):   # bands must end at U
            e.set_scans([0, N, 2 * N], [0, Ua, 2 * Ua - 16])
        e.set_scans([0, N, 2 * N], [0, Ua, 2 * Ua])
        # runs before the data are bound (set_scans unbinds them) are refused
        with pytest.raises(PtyError):
            e.run(0, 1, 1, begin=True)
        amp, norms = e.prepare_amp(torch.from_numpy(np.concatenate([s.I, s.I])).to(gpu_device))
        assert len(norms) == 2 and norms[0] == norms[1]
        e.set_data(amp, None, 0.0)
        O = torch.ones((2 * Ua, V), dtype=torch.complex64, device=gpu_device)
        P = torch.from_numpy(np.stack([s.P, s.P]).astype(np.complex64)).to(gpu_device)
        e.set_state(O, P)
        for bad in (dict(method=4), dict(method=3, OP_iters=(1, 2)), dict(method=1, probe_centering=True),
                    dict(method=1, Pmod_probe_until=2)):
            m = bad.pop("method")
            with pytest.raises(PtyError):
                e.run(0, m, 1, begin=True, **bad)
        with pytest.raises(PtyError):
            e.run(1, 1, 1, begin=True)          # DM is not packed
        m = e.run(0, 1, 3, begin=True)
        assert m.shape == (3, 2, 3)
        # the two identical scans give identical metrics and object bands
        assert np.array_equal(m[:, 0], m[:, 1])
        Oc = O.cpu().numpy()
        assert np.array_equal(Oc[:U], Oc[Ua:Ua + U])
    finally:
        e.close()
