"""Served form (continuous batching across jobs): every job's result is bit-identical whether it
runs alone (one-shot ``totem_stem_run``, or alone on a server) or in a mix with jobs of other
grid sizes that arrive while it is on the GPU, including when a small wave arena forces partial
admission and page recycling between grid sizes.  One served job is also checked against the
FP64 oracle (<= 1e-4 relative per pixel)."""
import time

import numpy as np
import pytest

from oracle import multislice as oms
from totem_amd import optics, serve, specimen, stem

pytestmark = pytest.mark.gpu

MAX_APER, MAX_MID = 4, 2


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _cfg(n, np_side, S, scan, M=1, mid=(), flag="l", seed=0, step=5, begin=(1, 1)):
    a = 3.905
    cells = max(1, int(round(np_side * 0.15 / a)))
    x = specimen.perovskite(cells, cells, max(1, S // 2), a=a)
    rng = np.random.default_rng(seed)
    x[:, 1:3] += rng.normal(0, 0.05, (len(x), 2))
    x[:, 1:3] %= cells * a
    dz = a / 2
    ab = optics.Aberrations(C1=-5.0, Cs_um=1.0) if M > 1 else None
    return stem.StemConfig(atoms=x, np_side=np_side, dx=cells * a / np_side, n=n, kv=200.0, alpha_mrad=20.0,
                           slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag, aberrations=ab,
                           focal_members=M, focal_rms_nm=3.0, detectors_mrad=((30.0, 90.0), (0.0, 10.0)),
                           width_red=scan[0], hight_red=scan[1], step=step, begin_row=begin[0],
                           begin_col=begin[1], mid_slices=mid)


JOBS = [
    dict(n=128, np_side=256, S=6, scan=(3, 2), mid=(2,), seed=1),
    dict(n=256, np_side=256, S=8, scan=(4, 4), M=2, flag="p", seed=2, step=7),
    dict(n=64, np_side=128, S=10, scan=(5, 5), seed=3, step=3),
    dict(n=512, np_side=512, S=4, scan=(2, 2), flag="p", seed=4, step=11),
    dict(n=128, np_side=256, S=12, scan=(4, 3), mid=(5, 9), seed=5, step=4),
    dict(n=256, np_side=256, S=5, scan=(12, 12), seed=6, step=2),
]


def _device_jobs():
    return [serve.DeviceJob(stem.prepare(_cfg(**k))) for k in JOBS]


def _one_shot(dj):
    torch = _torch()
    from totem_amd import engine
    dj.det.zero_()
    engine.stem_run(dj.job, dj.M * dj.D, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return dj.det.cpu().numpy().copy()


def _run_on_server(server, jobs, schedule, limit_s=120.0):
    """Submit jobs per the schedule {step: [job, ...]} and step the server until all are done."""
    torch = _torch()
    for dj in jobs:
        dj.det.zero_()
    torch.cuda.synchronize()
    pending = {k: list(v) for k, v in schedule.items()}
    ids, done = set(), set()
    steps = 0
    t0 = time.perf_counter()
    while pending or not ids <= done:
        for dj in pending.pop(steps, []):
            ids.add(server.submit(dj.job))
        server.step()
        done.update(server.poll())
        steps += 1
        assert time.perf_counter() - t0 < limit_s, "server did not finish"
    return steps


def test_served_jobs_equal_alone_and_in_mix(engine):
    torch = _torch()
    print("case server alone vs mix", flush=True)
    jobs = _device_jobs()
    alone = [_one_shot(dj) for dj in jobs]
    srv = engine.Server(4 << 30, MAX_APER, MAX_MID)
    try:
        # each job alone on the server
        for dj, ref in zip(jobs, alone):
            _run_on_server(srv, [dj], {0: [dj]})
            np.testing.assert_array_equal(dj.det.cpu().numpy(), ref)
        # all of them in a staggered mix of grid sizes
        _run_on_server(srv, jobs, {0: [jobs[0]], 3: [jobs[1], jobs[2]], 4: [jobs[3]], 7: [jobs[4], jobs[5]]})
        for k, (dj, ref) in enumerate(zip(jobs, alone)):
            got = dj.det.cpu().numpy()
            print("  job %d (N=%d): in-mix == alone: %s" % (k, dj.N, np.array_equal(got, ref)))
            np.testing.assert_array_equal(got, ref)
    finally:
        srv.close()
    # one served job against the FP64 oracle on the same transmission
    dj = jobs[0]
    inp, c = dj.inp, dj.inp.cfg
    t_host = dj.t.cpu().numpy().astype(np.complex128)
    res_o, mid_o = oms.stem_cpu(t_host, inp.probe, inp.gfsf, inp.psf, c.begin_row, c.begin_col, c.width_red,
                                c.hight_red, c.step, inp.aper, tuple(sorted(c.mid_slices)))
    myresul, mid = dj.images()
    rel = max(np.max(np.abs(myresul - res_o) / np.abs(res_o)), np.max(np.abs(mid - mid_o) / np.abs(mid_o)))
    print("  served job vs FP64 oracle: max rel per pixel %.3e" % rel)
    assert rel <= 1e-4
    del torch


def test_small_arena_partial_admission_and_page_recycling(engine):
    """An arena of two pages (63 slots of 256^2 each): the 144-probe 256^2 job is admitted in two
    parts; the 512^2, 128^2 and second 256^2 jobs wait for pages to drain and reuse them."""
    print("case server small arena", flush=True)
    jobs = _device_jobs()
    sel = [jobs[5], jobs[3], jobs[4], jobs[1]]
    alone = [_one_shot(dj) for dj in sel]
    srv = engine.Server(2 * (2048 * 2048 * 8 + 256 * 1024), MAX_APER, MAX_MID)   # two pages
    try:
        _run_on_server(srv, sel, {0: sel[:2], 2: sel[2:]})
        for dj, ref in zip(sel, alone):
            np.testing.assert_array_equal(dj.det.cpu().numpy(), ref)
    finally:
        srv.close()


def test_serving_loop_poisson_mix(engine):
    """A short Poisson load of mixed grid sizes through serve.serve: every request finishes and equals
    its one-shot result."""
    torch = _torch()
    print("case serving loop", flush=True)
    rng = np.random.default_rng(5)
    reqs = []
    t = 0.0
    for k in range(8):
        cfg, kind = serve.mixed_request(rng, grids=((128, 0.4), (256, 0.4), (512, 0.2)), slices=(4, 12),
                                        scan=(2, 6), cells=2)
        t += rng.exponential(0.02)
        reqs.append(serve.Request(rid=k, arrival=t, cfg=cfg, kind=kind))
    srv = engine.Server(2 << 30, MAX_APER, MAX_MID)
    try:
        wall, peak = serve.serve(reqs, srv, hbm_budget=8 << 30)
    finally:
        srv.close()
    lat = np.array([r.latency for r in reqs])
    print("  %d requests in %.2f s, p50 %.3f s, p99 %.3f s, peak resident %.1f MiB"
          % (len(reqs), wall, np.percentile(lat, 50), np.percentile(lat, 99), peak / 2**20))
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    for r in reqs:
        dj = serve.DeviceJob(r.inp)
        _one_shot(dj)
        ref, _ = dj.images()
        np.testing.assert_array_equal(r.result[0], ref)
    del torch
