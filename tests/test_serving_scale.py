"""Served form at the shapes of the uneven-batch and serving workloads.

* Uneven batch: the seeded job sequence of the uneven-batch workload (seed 1004: SrTiO3,
  256^2-1024^2 grids including 768^2, 10-60 slices, defocus in [-10, 10] nm, 8x8-32x32 scans),
  drawn until it holds two 768^2 and two 1024^2 jobs (15 jobs, 204 000 probe.slices).  Every job
  is bit-identical run alone (one-shot) and in the served mix through a cache-sized wave arena;
  one 768^2 and one 1024^2 job agree with the FP64 oracle (potential built from the same atoms,
  propagation at three scan positions) within 1e-4 relative per pixel.
* Serving load: Poisson arrivals of the mixed request generator (seed 1005) with its real grid
  mix plus one forced 2048^2 request of >= 60 slices, through the request loop with an HBM
  budget that makes requests wait for others to finish and a two-page wave arena that admits
  probes in parts.  Every request's images equal its one-shot result bit for bit.

The reference runs one myProjectfun call per simulation (cuda_source_lobato.cu:1018-1562); its
memory-sized batching (:1249-1260) is what the served form generalises across jobs.
"""
import os
import subprocess

import numpy as np
import pytest

from totem_amd import serve, stem

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _one_shot(engine, dj):
    torch = _torch()
    dj.det.zero_()
    engine.stem_run(dj.job, dj.M * dj.D, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return dj.det.cpu().numpy().copy()


def _uneven_jobs():
    rng = np.random.default_rng(1004)
    cfgs = []
    while sum(c.n == 768 for c in cfgs) < 2 or sum(c.n == 1024 for c in cfgs) < 2:
        cfgs.append(serve.uneven_request(rng))
    return cfgs


def _oracle_images(inp, positions, threads):
    """FP64 C++ oracle from the job's atoms: potential of every slice, then the listed positions."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    from oracle import cpu_stem
    c = inp.cfg
    kw = {}
    if c.paraflag == "n":
        kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
    ele_i = inp.ele_n_i if len(inp.ele_n_i) else None
    _, t = cpu_stem.potential_slices(c.np_side, inp.ele_n, None, ele_i, None, inp.series_n,
                                     inp.series_n_i if ele_i is not None else None, inp.s2, inp.gx, inp.gy,
                                     inp.sigma, inp.parameter, inp.aper_true, c.paraflag, threads=threads,
                                     want_v=False, **kw)
    ash = np.fft.ifftshift(inp.aper, axes=(-2, -1))
    wrap = c.n == c.np_side
    r0 = [(c.begin_row - 1 + (d // c.width_red) * c.step) % c.np_side for d in positions]
    c0 = [(c.begin_col - 1 + (d % c.width_red) * c.step) % c.np_side for d in positions]
    img, _ = cpu_stem.stem_positions(t, inp.probe, inp.gfsf, inp.psf, ash, r0, c0, periodic=wrap, threads=threads)
    return img


def test_uneven_batch_shape_alone_equals_mix_and_oracle(engine):
    cfgs = _uneven_jobs()
    sides = [c.n for c in cfgs]
    print("case uneven batch: %d jobs, grids %s, slices %d-%d" %
          (len(cfgs), sorted(set(sides)), min(c.n_slices for c in cfgs), max(c.n_slices for c in cfgs)), flush=True)
    jobs = [serve.DeviceJob(stem.prepare(c)) for c in cfgs]
    alone = [_one_shot(engine, dj) for dj in jobs]
    srv = engine.Server(int(0.25 * 2**30), 8, 16)
    try:
        for dj in jobs:
            dj.det.zero_()
        wall = serve.run_all(srv, jobs)
    finally:
        srv.close()
    work = sum(dj.probe_slices for dj in jobs)
    print("  served %d probe.slices in %.2f s" % (work, wall), flush=True)
    for k, (dj, ref) in enumerate(zip(jobs, alone)):
        np.testing.assert_array_equal(dj.det.cpu().numpy(), ref, err_msg="job %d (N=%d)" % (k, dj.N))
    threads = max(1, min(16, os.cpu_count() or 1))
    for n in (768, 1024):
        k = sides.index(n)
        dj = jobs[k]
        D = dj.D
        positions = [0, D // 2, D - 1]
        ref = _oracle_images(dj.inp, positions, threads)
        got = dj.images()[0][positions]
        rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
        print("  job %d (N=%d, %d slices): served vs FP64 oracle at positions %s: max rel per pixel %.2e"
              % (k, n, dj.S, positions, rel), flush=True)
        assert rel <= 1e-4


def test_serving_load_shape_alone_equals_mix(engine):
    torch = _torch()
    rng = np.random.default_rng(1005)
    reqs, t = [], 0.0
    for k in range(10):
        cfg, kind = serve.mixed_request(rng, scan=(8, 12))
        t += rng.exponential(0.05)
        reqs.append(serve.Request(rid=k, arrival=t, cfg=cfg, kind=kind))
    big, kind = serve.mixed_request(rng, grids=((2048, 1.0),), slices=(60, 120), scan=(8, 8))
    reqs.insert(2, serve.Request(rid=10, arrival=reqs[1].arrival + 0.01, cfg=big, kind=kind))
    assert big.n == 2048 and big.n_slices >= 60
    budget = int(2.5e9)
    assert serve.footprint_bytes(big) > budget / 2           # nothing else fits beside it
    print("case serving load: %d requests, grids %s, 2048^2 request with %d slices, budget %.1f GB" %
          (len(reqs), sorted({r.cfg.n for r in reqs}), big.n_slices, budget / 1e9), flush=True)
    srv = engine.Server(2 * (2048 * 2048 * 8 + 256 * 1024), 8, 16)     # two pages: admission in parts
    try:
        wall, peak = serve.serve(reqs, srv, hbm_budget=budget)
    finally:
        srv.close()
    lat = np.array([r.latency for r in reqs])
    print("  %d requests in %.2f s, p50 %.3f s, p99 %.3f s, peak resident %.2f GB"
          % (len(reqs), wall, np.percentile(lat, 50), np.percentile(lat, 99), peak / 1e9), flush=True)
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    assert any(r.extra.get("budget_wait") for r in reqs), "no request waited for the HBM budget"
    for r in reqs:
        dj = serve.DeviceJob(r.inp)
        _one_shot(engine, dj)
        ref, mid = dj.images()
        np.testing.assert_array_equal(r.result[0], ref, err_msg="request %d (N=%d)" % (r.rid, r.cfg.n))
        np.testing.assert_array_equal(r.result[1], mid)
    del torch
