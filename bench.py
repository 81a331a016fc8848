"""Throughput benchmark of the multislice STEM engine on one MI355X.

Workload (default): frozen-phonon STEM of SrTiO3 20x20x40 cells (80 000 atoms), Peng + correction
factors, 1024^2 grid over 78.1 A, 80 slices, 300 kV, 24 mrad, Cs 1 um, C1 -5 nm, ADF 70-200 and
BF 0-10 mrad, 64x64 probe positions (periodic windows), seeded phonon configurations.  One step is
one phonon configuration: displacement, slicing and transmission build on the GPU (resident
specimen, Philox-seeded) + all 4096 probes through all 80 slices.

Metric: probe.slice/s (one probe wave through one slice: transmit, 2-D FFT, propagate, inverse FFT),
with the specimen's atoms and coefficient rows, probes, propagator and detector masks resident in HBM
when the timed region starts.

The JSON line also carries the plain one-probe-per-launch rate (same kernels, batch = 1, no graph),
the roofline of the dominant kernel from HIP events, and the FP64 oracle timed on the host cores
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_threads():
    """Host threads of the CPU baseline: the process's CPU share (OMP_NUM_THREADS where the launcher
    sets it, as on a one-GPU box), else every core os.cpu_count() reports."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_baseline(cfg, atoms, inp, positions, threads):
    """The reference's multicore CPU path restated in compiled FP64 (oracle/cpu_stem.cpp, OpenMP over
    slices for the potential and over probe positions for the propagation, like its parfor loops) on
    the same phonon configuration: the whole potential build from the displaced atoms, then
    len(positions) probes through every slice.  Returns (potential build s, propagation s, images
    [len(positions), A])."""
    from oracle import cpu_stem
    from totem_amd import stem
    rows = stem.prepare(cfg, atoms=atoms, optics_too=False)
    kw = {}
    if cfg.paraflag == "n":
        kw = dict(ele_n_corr=rows.ele_n_corr, corr_matrix=rows.corr_matrix, series_n_corr=rows.series_n_corr)
    ele_i = rows.ele_n_i if len(rows.ele_n_i) else None
    t0 = time.perf_counter()
    _, t = cpu_stem.potential_slices(cfg.np_side, rows.ele_n, None, ele_i, None, rows.series_n,
                                     rows.series_n_i if ele_i is not None else None, inp.s2, inp.gx, inp.gy,
                                     inp.sigma, inp.parameter, inp.aper_true, cfg.paraflag, threads=threads,
                                     want_v=False, **kw)
    build_s = time.perf_counter() - t0
    ash = np.fft.ifftshift(inp.aper, axes=(-2, -1))
    r0 = [(cfg.begin_row - 1 + (d // cfg.width_red) * cfg.step) % cfg.np_side for d in positions]
    c0 = [(cfg.begin_col - 1 + (d % cfg.width_red) * cfg.step) % cfg.np_side for d in positions]
    t1 = time.perf_counter()
    img, _ = cpu_stem.stem_positions(t, inp.probe, inp.gfsf, inp.psf, ash, r0, c0, periodic=True, threads=threads)
    prop_s = time.perf_counter() - t1
    return build_s, prop_s, img


def measured_traffic(kname, n, workload):
    """HBM bytes per dispatch of the kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (copy-calibrated read + write MiB columns), or None when no summary covers this workload."""
    path = os.path.join(ROOT, "profiles", "stem_passes_%d_%s_traffic.txt" % (n, workload))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(kname + " "):
                f = line[len(kname):].split()
                return (float(f[1]) + float(f[3])) * 2**20
    return None


def measured_valu_issue(kname, n, workload):
    """Fraction of the kernel's duration its VALU instructions take to issue (SQ_INSTS_VALU x 4
    cycles / 1024 SIMDs / 2.4 GHz over the kernel-trace duration), from the committed SQ counter
    summary, or None when no summary covers this workload."""
    path = os.path.join(ROOT, "profiles", "stem_passes_%d_%s_sq_counters.txt" % (n, workload))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        text = fh.read()
    m = re.search(re.escape(kname) + r":[^\n]*\n\s*VALU issue time [\d.]+ us .*?= (\d+)% of", text)
    return int(m.group(1)) / 100.0 if m else None


def run_hrtem_batch(n_images, cpu_images, seed, n=320):
    """The batch HRTEM tool's producer step (cuda_stem64.py:164-704 call shape: draw aberrations,
    transfer function x exit wave, inverse FFT, |.|^2) on synthetic inputs of the tool's shape: a
    320^2 exit wave, 80 kV, every aberration drawn.  Returns images/s through the host API (host
    exit wave in, host images out, as the tool's queue receives them) and the FP64 oracle's rate on
    the host for a few of the same images, with its error."""
    from oracle import hrtem as ohr
    from totem_amd import hrtem
    rng = np.random.default_rng(seed)
    lam, sampling = 0.00417552, 0.01
    f = (np.arange(n) - n // 2) / (n * sampling)
    U, V = np.meshgrid(f, f)
    w = np.exp(1j * rng.uniform(-0.5, 0.5, (n, n)))
    mywave = np.fft.fftshift(np.fft.fft2(w))
    pp2 = np.zeros(49)
    pp2[0], pp2[1], pp2[2], pp2[8] = lam, 10.0, 10.0, 1.0e3
    mrr = np.abs(rng.normal(0.0, 5.0, 21)) + 1.0
    pp1 = {"lambda": lam, "alafa": 3e-4, "delta_yita": 2.0, "gfsf": 1.0, "mtotal": 1}
    ai = rng.random((n_images, 21))
    for _ in range(2):                                          # warm-up: plans, device and page-locked buffers
        img, paras = hrtem.images(pp1, pp2, mrr, U, V, mywave, ai)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        img = paras = None                                      # the consumer has taken the previous batch
        img, paras = hrtem.images(pp1, pp2, mrr, U, V, mywave, ai)
    gpu_wall = (time.perf_counter() - t0) / reps
    k = min(cpu_images, n_images)
    t1 = time.perf_counter()
    ref = ohr.images(mywave, paras[:k], U, V, lam, pp1["alafa"], 1, pp1["delta_yita"], pp1["gfsf"])
    cpu_wall = time.perf_counter() - t1
    err = max(float(np.linalg.norm(img[b] - ref[b]) / np.linalg.norm(ref[b])) for b in range(k))
    return {"grid": n, "images_per_call": n_images, "images_per_s": n_images / gpu_wall,
            "ms_per_call": gpu_wall * 1e3, "includes": "host->device inputs, all kernels, images back to host",
            "cpu_oracle": {"value": k / cpu_wall, "unit": "images/s", "cores": 1, "kind": "port",
                           "sample": "%d images, FP64 numpy oracle (oracle/hrtem.py)" % k},
            "oracle_max_rel_l2": err, "seed": seed}


def run_precision_stress(scan):
    """The precision-stress specimen (diamond-cubic Si [110] ~100 x 100 x 300 A with 50 A amorphous-C
    caps, 270 000 atoms) at 2048^2 and ~210 slices of 1.92 A: transmission built on the device from
    the resident specimen, then a scan x scan probe grid (periodic windows) through every slice."""
    import torch
    from totem_amd import engine, optics, stem
    cfg = stem.precision_c3(scan=scan)
    S, N, NP = cfg.n_slices, cfg.n, cfg.np_side
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    t = torch.empty((S, NP, NP), dtype=torch.complex64, device=dev)
    sp = stem.device_specimen(cfg)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    sp.transmission(0, 0, t.data_ptr(), stream=st)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    sp.close()
    probe, gfsf = optics.probe_stack(N, cfg.dx, cfg.kv, cfg.alpha_mrad, cfg.aberrations, cfg.focal_members,
                                     cfg.focal_rms_nm, cfg.source_mrad, cfg.tilt_mrad, cfg.tilt_phi_deg)
    psf = optics.propagator(N, cfg.dx, cfg.kv, cfg.slice_thickness)
    aper = optics.detector_masks(N, cfg.dx, cfg.kv, cfg.detectors_mrad)
    probes = torch.tensor(probe.astype(np.complex64), device=dev)
    prop = torch.tensor(optics.unit_complex64(psf), device=dev)
    aper_d = torch.tensor(aper, device=dev)
    aper2 = torch.tensor(np.fft.ifftshift(aper, axes=(-2, -1)).copy(), device=dev)
    M, A, D = len(gfsf), aper.shape[0], cfg.depth
    det = torch.zeros((1, A, M * D), dtype=torch.float64, device=dev)
    job = engine.stem_job(N, NP, S, t.data_ptr(), probes.data_ptr(), M, prop.data_ptr(), aper_d.data_ptr(), A, D,
                          cfg.width_red, cfg.step, cfg.begin_row - 1, cfg.begin_col - 1, det.data_ptr(),
                          aper2_ptr=aper2.data_ptr(), wrap=True)
    batch = min(M * D, engine.preferred_batch(N))
    ws_bytes = engine.stem_workspace_bytes(job, batch)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    engine.stem_run(job, batch, workspace_ptr=ws.data_ptr(), workspace_bytes=ws_bytes, stream=st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {"workload": "Si [110] ~100x100x300 A + 2x50 A amorphous C, %d atoms, Lobato, 200 kV 22 mrad, "
                       "ADF 60-200 mrad" % len(cfg.atoms),
           "grid": N, "slices": S, "probes": M * D, "probes_per_launch": batch,
           "probe_slices_per_s": M * D * S / wall, "wall_s": wall, "transmission_build_s": build_s,
           "band_columns": engine.stem_band(optics.unit_complex64(psf),
                                            engine.plan_columns(N, engine.plan(N)[0]))[1]}
    del t, ws, probes, det
    torch.cuda.empty_cache()
    return out


def run_serving(n_requests, rate, seed, arena_gib, budget_gb):
    """Poisson arrivals of mixed-size STEM requests (grids 256^2 50%, 512^2 30%, 1024^2 15%,
    2048^2 5%; 10-120 slices; scans 8x8-64x64; Lobato or Peng; SrTiO3 or amorphous C) through the
    served engine, admitted by footprint against ``budget_gb`` (1e9 bytes).  Each request is
    prepared on the host (slicing, optics) and built on the GPU after it arrives, so latency
    (arrival to host images) includes that preparation.  Returns the latency percentiles with the
    sample count and the aggregate probe.slice/s."""
    from totem_amd import engine, serve
    rng = np.random.default_rng(seed)
    reqs, t, work = [], 0.0, 0
    mix = {}
    for k in range(n_requests):
        cfg, kind = serve.mixed_request(rng)
        t += rng.exponential(1.0 / rate)
        reqs.append(serve.Request(rid=k, arrival=t, cfg=cfg, kind=kind))
        work += serve.focal_count(cfg) * cfg.depth * cfg.n_slices
        mix["%d" % cfg.n] = mix.get("%d" % cfg.n, 0) + 1
    srv = engine.Server(int(arena_gib * 2**30), 8, 16)
    try:
        wall, peak = serve.serve(reqs, srv, hbm_budget=int(budget_gb * 1e9))
    finally:
        srv.close()
    lat = np.array([r.latency for r in reqs])
    prep = np.array([r.extra.get("prepare_s", 0.0) for r in reqs])
    by = {}
    for r in reqs:
        by.setdefault("%d" % r.cfg.n, []).append(r.latency)
    return {"requests": n_requests, "latency_samples": int(len(lat)), "arrival_rate_per_s": rate, "grids": mix,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "max_latency_s": float(lat.max()),
            "p50_latency_s_by_grid": {k: float(np.percentile(v, 50)) for k, v in sorted(by.items())},
            "latency_includes": "arrival -> host preparation (slicing, optics) -> transmission build -> "
                                "served propagation -> images on the host",
            "p50_host_prepare_s": float(np.percentile(prep, 50)), "wall_s": wall, "probe_slices": int(work),
            "throughput_probe_slices_per_s": work / wall, "peak_resident_GB": peak / 1e9,
            "arena_GiB": arena_gib, "hbm_budget_GB": budget_gb, "seed": seed}


def run_erf_build(reps=3):
    """Device build of the frozen-phonon specimen's transmission (80 000 atoms, Peng+corr, 80 slices
    of 1024^2) with erf-weighted slicing at multiceng 0, 0.5 and 1: displacement, slicing, sort and
    potential of one phonon configuration, best of ``reps`` after one untimed build."""
    import torch
    from dataclasses import replace
    from totem_amd import engine, stem
    cfg0 = stem.perovskite_c2(scan=4)
    out = {}
    for m in (0.0, 0.5, 1.0):
        cfg = replace(cfg0, multiceng=m)
        sp = stem.device_specimen(cfg, sigma_by_z=stem.C2_PHONON_SIGMA)
        t = torch.empty((cfg.n_slices, cfg.np_side, cfg.np_side), dtype=torch.complex64, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        sp.transmission(1, 0, t.data_ptr(), stream=st)
        torch.cuda.synchronize()
        best = None
        for k in range(reps):
            t0 = time.perf_counter()
            sp.transmission(1, k + 1, t.data_ptr(), stream=st)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        out["%g" % m] = best * 1e3
        sp.close()
        del t
        torch.cuda.empty_cache()
    return {"workload": "SrTiO3 20x20x40 cells (80 000 atoms), Peng+corr, 80 slices of 1024^2, one phonon "
                        "configuration: displacement + slicing + sort + potential + t",
            "ms_per_config_by_multiceng": out}


def run_drop_in():
    """One ``myProjectfun`` call through the drop-in library libwaveCudaPengAndCorr.so with the mex
    argument list of the frozen-phonon specimen (80 000 atoms, static snapshot, Peng+corr): host
    rows in, 1024^2 potential built, 64x64 probes of 512^2 windows through 80 slices, detector images
    back on the host.  This is the PCIe-inclusive form (the reference's contract); the second call
    is reported, the first includes the library's one-time set-up."""
    from dataclasses import replace
    from totem_amd import stem
    cfg = replace(stem.perovskite_c2(scan=64), n=512)
    t0 = time.perf_counter()
    inp = stem.prepare(cfg)
    prep = time.perf_counter() - t0
    walls, outs = [], []
    for _ in range(2):
        t0 = time.perf_counter()
        outs.append(stem.run(inp, drop_in=True)[0])
        walls.append(time.perf_counter() - t0)
    M, D, S = len(inp.gfsf), cfg.depth, cfg.n_slices
    return {"library": "libwaveCudaPengAndCorr.so", "symbol": "myProjectfun",
            "workload": "SrTiO3 20x20x40 cells static snapshot, Peng+corr, 1024^2 potential, 80 slices, "
                        "64x64 probes of 512^2 windows",
            "s_per_call": walls[1], "first_call_s": walls[0], "host_prepare_s": prep,
            "probe_slices_per_s": M * D * S / walls[1], "repeat_bit_identical": bool(np.array_equal(*outs)),
            "includes": "host rows -> device, potential build, propagation, allocation, images -> host"}


def run_uneven(n_jobs, seed, arena_gib):
    """Uneven batch: n_jobs seeded perovskite jobs of mixed grid sides (256^2-1024^2, 768^2 with a
    radix-3 stage), thickness, defocus and scan size, all in one served run, against the same jobs
    run one at a time through the one-shot entry point (each job fully batched)."""
    import torch
    from totem_amd import engine, serve, stem
    rng = np.random.default_rng(seed)
    cfgs = [serve.uneven_request(rng) for _ in range(n_jobs)]
    jobs = [serve.DeviceJob(stem.prepare(c)) for c in cfgs]
    work = sum(dj.probe_slices for dj in jobs)
    srv = engine.Server(int(arena_gib * 2**30), 8, 16)
    try:
        serve.run_all(srv, jobs[:2])          # warm the per-size plans and streams
        for dj in jobs:
            dj.det.zero_()
        wall = serve.run_all(srv, jobs)
    finally:
        srv.close()
    served = [dj.det.cpu().numpy().copy() for dj in jobs]
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for dj in jobs:
        engine.stem_run(dj.job, dj.M * dj.D, stream=st)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    same = all(np.array_equal(a, dj.det.cpu().numpy()) for a, dj in zip(served, jobs))
    sides = {}
    for c in cfgs:
        sides["%d" % c.n] = sides.get("%d" % c.n, 0) + 1
    return {"jobs": n_jobs, "grids": sides, "probe_slices": int(work), "served_wall_s": wall,
            "served_probe_slices_per_s": work / wall, "one_job_at_a_time_wall_s": seq,
            "one_job_at_a_time_probe_slices_per_s": work / seq, "bit_identical": same, "seed": seed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed phonon configurations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c2", "c1"], default="c2")
    ap.add_argument("--graph", type=int, default=1, help="capture the slice loop in a hipGraph (0: launch directly)")
    ap.add_argument("--scan", type=int, default=64, help="scan side of the 1024^2 workload (64: 64x64 probes)")
    ap.add_argument("--batch", type=int, default=0,
                    help="probes in flight per launch (0: measured by the engine's autotuner)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the cache-fit batch and the current pass plan instead of measuring them")
    ap.add_argument("--plan", type=int, default=-1, help="pass plan of the grid (-1: default / autotuned)")
    ap.add_argument("--host-slicing", action="store_true",
                    help="displace and slice on the host before the timed loop (default: on the device, per step)")
    ap.add_argument("--plain-probes", type=int, default=16)
    ap.add_argument("--cpu-probes", type=int, default=0,
                    help="probe positions of the CPU baseline (0: one per host thread)")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="host threads of the CPU baseline (0: OMP_NUM_THREADS, else os.cpu_count())")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uneven-jobs", type=int, default=48, help="jobs of the uneven-batch leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=120, help="requests of the mixed serving load (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--stress-scan", type=int, default=32,
                    help="scan side of the 2048^2 precision-stress leg (0: skip)")
    ap.add_argument("--hrtem-images", type=int, default=256,
                    help="images per call of the batch HRTEM leg (0: skip)")
    ap.add_argument("--hrtem-only", action="store_true", help="run only the batch HRTEM leg (profiling)")
    ap.add_argument("--erf-build", type=int, default=1, help="time the erf-weighted specimen build (0: skip)")
    ap.add_argument("--drop-in", type=int, default=1,
                    help="time one myProjectfun call through the drop-in library (0: skip)")
    ap.add_argument("--only", choices=["stress", "uneven", "serving", "hrtem", "erf", "dropin"],
                    help="run only this leg and print it alone (profiling)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-traffic kernel for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (multi-GPU sharding is design-only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from totem_amd import engine, optics, stem

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    engine.lib()
    if args.calibrate_copy_mib > 0:
        src = torch.ones(args.calibrate_copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
    if args.hrtem_only:
        args.only = "hrtem"
    if args.only == "hrtem":
        print(json.dumps({"hrtem_batch": run_hrtem_batch(max(args.hrtem_images, 1), 4, seed=1006)}), flush=True)
        return
    if args.only == "stress":
        print(json.dumps({"precision_stress": run_precision_stress(max(args.stress_scan, 1))}), flush=True)
        return
    if args.only == "erf":
        print(json.dumps({"erf_weighted_build": run_erf_build()}), flush=True)
        return
    if args.only == "dropin":
        print(json.dumps({"drop_in_call": run_drop_in()}), flush=True)
        return
    if args.only == "uneven":
        print(json.dumps({"uneven_batch": run_uneven(max(args.uneven_jobs, 1), seed=1004, arena_gib=0.25)}),
              flush=True)
        return
    if args.only == "serving":
        print(json.dumps({"serving": run_serving(max(args.serve_requests, 1), args.serve_rate, seed=1005,
                                                 arena_gib=0.25, budget_gb=288)}), flush=True)
        return
    dev = torch.device("cuda:0")
    if args.workload == "c2":
        cfg = stem.perovskite_c2(scan=args.scan)
        sig = stem.C2_PHONON_SIGMA
        name = ("SrTiO3 20x20x40 cells frozen phonon, Peng+corr, 1024^2, 80 slices, %dx%d probes, "
                "300 kV 24 mrad Cs 1um C1 -5nm, ADF 70-200 + BF 0-10 mrad" % (args.scan, args.scan))
    else:
        cfg = stem.perovskite_c1()
        sig = {}
        name = "SrTiO3 5x5x10 cells static, Lobato, 256^2, 20 slices, 16x16 probes, 200 kV 20 mrad, ADF 60-200 mrad"
    n_cfg = args.warmup + args.steps
    if args.host_slicing:
        snaps = stem.phonon_snapshots(cfg, n_cfg, sig, seed=1002) if sig else [cfg.atoms] * n_cfg
        rows = [stem.prepare_specimen(cfg, x) for x in snaps]
    else:
        rows = list(range(n_cfg))   # configuration numbers of the resident specimen
        spec = stem.device_specimen(cfg, sigma_by_z=sig or None)
    inp = stem.prepare(cfg)
    S, N, NP = cfg.n_slices, cfg.n, cfg.np_side
    M, A, D = len(inp.gfsf), inp.aper.shape[0], cfg.depth

    probes = torch.tensor(inp.probe.astype(np.complex64), device=dev)
    prop = torch.tensor(optics.unit_complex64(inp.psf), device=dev)
    aper = torch.tensor(inp.aper, device=dev)
    aper2 = torch.tensor(inp.aper2, device=dev)
    t = torch.empty((S, NP, NP), dtype=torch.complex64, device=dev)
    det = torch.zeros((1, A, M * D), dtype=torch.float64, device=dev)
    job = engine.stem_job(N, NP, S, t.data_ptr(), probes.data_ptr(), M, prop.data_ptr(), aper.data_ptr(), A, D,
                          cfg.width_red, cfg.step, cfg.begin_row - 1, cfg.begin_col - 1, det.data_ptr(),
                          aper2_ptr=aper2.data_ptr(), wrap=True)
    tuned = None
    if args.plan >= 0:
        engine.set_plan(N, args.plan)
    if args.batch <= 0 and not args.no_autotune:
        tuned = engine.autotune(N)        # plan x probes per launch
    batch = min(M * D, args.batch if args.batch > 0 else engine.preferred_batch(N))
    plan_index, plan_desc = engine.plan(N)
    ws_bytes = engine.stem_workspace_bytes(job, batch)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def build(r):
        if not args.host_slicing:
            spec.transmission(1002, r, t.data_ptr(), stream=stream)
            return
        corr = cfg.paraflag == "n"
        engine.build_transmission(cfg.paraflag, NP, S, r["series_n"], r["ele_n"], inp.gx, inp.gy, inp.s2,
                                  inp.aper_true, inp.sigma, inp.parameter, t.data_ptr(),
                                  series_n_corr=r["series_n_corr"] if corr else None,
                                  ele_n_corr=r["ele_n_corr"] if corr else None,
                                  corr_info_matrix=r["corr_matrix"] if corr else None, stream=stream)

    def step(r, timing=False):
        build(r)
        engine.stem_run(job, batch, graph=bool(args.graph) and not timing, workspace_ptr=ws.data_ptr(),
                        workspace_bytes=ws_bytes, stream=stream, timing=timing)

    for k in range(args.warmup):
        step(rows[k])
    torch.cuda.synchronize()
    engine.kernel_times(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(rows[args.warmup + k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    probe_slices = args.steps * M * D * S
    value = probe_slices / wall

    # per-kernel durations: the last configuration once more with HIP events around every pass
    # (on the launch stream; the events synchronise per batch, so this run is not the timed one)
    engine.stem_run(job, batch, workspace_ptr=ws.data_ptr(), workspace_bytes=ws_bytes, stream=stream, timing=True)
    torch.cuda.synchronize()
    rows_ms, rows_n, cols_ms, cols_n = engine.kernel_times(reset=True)

    # potential build alone (per phonon configuration)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    build(rows[-1])
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    engine.stem_run(job, batch, workspace_ptr=ws.data_ptr(), workspace_bytes=ws_bytes, stream=stream)
    torch.cuda.synchronize()
    det_ref = det.cpu().numpy()[0].copy()

    # plain form: same kernels, one probe per launch sequence, no graph
    pp = min(args.plain_probes, D)
    plain_value = plain_match = None
    if pp > 0:
        det_p = torch.zeros((1, A, M * pp), dtype=torch.float64, device=dev)
        job_p = engine.stem_job(N, NP, S, t.data_ptr(), probes.data_ptr(), M, prop.data_ptr(), aper.data_ptr(), A,
                                pp, cfg.width_red, cfg.step, cfg.begin_row - 1, cfg.begin_col - 1, det_p.data_ptr(),
                                aper2_ptr=aper2.data_ptr(), wrap=True)
        engine.stem_run(job_p, 1, workspace_ptr=ws.data_ptr(), workspace_bytes=ws_bytes, stream=stream)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        engine.stem_run(job_p, 1, workspace_ptr=ws.data_ptr(), workspace_bytes=ws_bytes, stream=stream)
        torch.cuda.synchronize()
        plain_wall = time.perf_counter() - tp
        plain_value = M * pp * S / plain_wall
        plain_match = bool(np.array_equal(det_p.cpu().numpy()[0], det_ref[:, :M * pp]))

    # roofline of the dominant kernel (HIP events on the launch stream).  The passes move the
    # wave's band columns only (the columns where the propagator has a nonzero entry, widened
    # to whole column groups): f = band / N of each probe's wave is read and written.
    n2 = float(N) * N
    _, band = engine.stem_band(optics.unit_complex64(inp.psf), engine.plan_columns(N, engine.plan(N)[0]))
    f = band / float(N)
    bytes_rows = batch * 16.0 * f * n2 + 8.0 * NP * NP      # band of psi read+write + the shared t slice
    bytes_cols = batch * 16.0 * f * n2 + 8.0 * f * n2       # band of psi read+write + its propagator columns
    rows_avg = rows_ms / max(rows_n, 1)
    cols_avg = cols_ms / max(cols_n, 1)
    if rows_avg >= cols_avg:
        kname, kavg, kbytes = "k_pass_rows<%d>" % N, rows_avg, bytes_rows
    else:
        kname, kavg, kbytes = "k_pass_cols<%d>" % N, cols_avg, bytes_cols
    achieved = kbytes / (kavg * 1e-3) / 1e9 if kavg > 0 else None
    traffic = measured_traffic(kname, N, args.workload)
    # the other pass of the slice, and both together (per probe.slice: 32 f N^2 bytes + shared slices)
    oname, oavg, obytes = (("k_pass_cols<%d>" % N, cols_avg, bytes_cols) if kname.startswith("k_pass_rows")
                           else ("k_pass_rows<%d>" % N, rows_avg, bytes_rows))
    oach = obytes / (oavg * 1e-3) / 1e9 if oavg > 0 else None
    both = (bytes_rows + bytes_cols) / ((rows_avg + cols_avg) * 1e-3) / 1e9 if rows_avg + cols_avg > 0 else None

    out = {
        "metric": "probe.slice/s at 1024^2 FP32 (batched vs plain one-probe-per-launch form)",
        "value": value,
        "unit": "probe.slice/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded specimen; phonon configurations drawn on the device)",
        "config": {"workload": name, "grid": N, "slices": S, "probes_per_step": M * D, "probes_per_launch": batch,
                   "batch_autotuned": tuned is not None,
                   "plan": {"index": plan_index, "of": engine.plan_count(N), "schedule": plan_desc,
                            "autotuned": tuned is not None},
                   "phonon_configs_timed": args.steps, "paraflag": cfg.paraflag, "graph": bool(args.graph),
                   "specimen": ("host-sliced rows prepared before the timed loop" if args.host_slicing else
                                "displaced, sliced and built on the device inside every step")},
        "plain": {"value": plain_value, "probes": pp,
                  "speedup_batched_vs_plain": value / plain_value if plain_value else None,
                  "bit_identical_to_batched": plain_match},
        "potential_build_s_per_config": build_s,
        "potential_build_includes": "rows -> t" if args.host_slicing else "displacement + slicing + rows -> t",
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_ms": kavg,
                     "band_columns": band, "band_fraction": f,
                     "valu_issue_frac": measured_valu_issue(kname, N, args.workload),
                     # the full-grid transform's bytes for the same launch over its measured time
                     # (the rate at which the pass advances probe.slices, in those bytes)
                     "full_grid_equivalent_frac":
                         ((batch * 16.0 * n2 + 8.0 * (NP * NP if kname.startswith("k_pass_rows") else n2))
                          / (kavg * 1e-3) / 1e9 / HBM_PEAK_GBS) if kavg > 0 else None,
                     "rows_avg_ms": rows_avg, "cols_avg_ms": cols_avg},
        "roofline_other_pass": {"kernel": oname, "achieved": oach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": (oach / HBM_PEAK_GBS) if oach else None,
                                "traffic": measured_traffic(oname, N, args.workload), "avg_launch_ms": oavg,
                                "valu_issue_frac": measured_valu_issue(oname, N, args.workload)},
        "roofline_both_passes": {"achieved": both, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": (both / HBM_PEAK_GBS) if both else None},
    }
    if not args.no_cpu:
        threads = args.cpu_cores if args.cpu_cores > 0 else cpu_threads()
        positions = list(range(min(args.cpu_probes if args.cpu_probes > 0 else threads, D)))
        if args.host_slicing:
            atoms = snaps[-1] if sig else cfg.atoms
        else:
            atoms = np.array(cfg.atoms, dtype=np.float64)
            atoms[:, 1:4] = spec.positions(1002, rows[-1])
            if sig:
                atoms[:, 6] = 0.0
        build_cpu, prop_cpu, cres = cpu_baseline(cfg, atoms, inp, positions, threads)
        # the whole step on the host: one configuration's potential plus D probes, the probes
        # D / len(positions) times the measured batch of one probe per thread
        step_cpu = build_cpu + prop_cpu * D * M / (len(positions) * M)
        g = det_ref.reshape(A, M, D)[:, :, positions]
        gpu = np.einsum("amd,m->da", g, inp.gfsf)
        rel = float(np.max(np.abs(gpu - cres) / np.abs(cres)))
        out["cpu_baseline"] = {
            "value": M * D * S / step_cpu, "unit": "probe.slice/s", "cores": threads, "kind": "port",
            "sample": "phonon configuration %d: potential of all %d slices (%.1f s) and %d probe positions "
                      "x %d slices (%.1f s), FP64 C++ restatement of the reference CPU path "
                      "(oracle/cpu_stem.cpp, OpenMP, %d threads; os.cpu_count() %d); value = %d probes x %d "
                      "slices over the build plus the propagation time scaled to all %d probes"
                      % (rows[-1] if not args.host_slicing else n_cfg - 1, S, build_cpu, len(positions), S,
                         prop_cpu, threads, os.cpu_count() or 0, M * D, S, M * D),
            "potential_build_s": build_cpu,
            "propagation_probe_slices_per_s": len(positions) * M * S / prop_cpu}
        out["oracle_max_rel_err"] = rel
        out["oracle_error_covers"] = "GPU potential build + propagation vs the FP64 C++ oracle from the same atoms"
    if (args.uneven_jobs > 0 or args.serve_requests > 0 or args.hrtem_images > 0 or args.stress_scan > 0 or
            args.erf_build or args.drop_in):
        if not args.host_slicing:
            spec.close()
        del ws, t, probes, det
        torch.cuda.empty_cache()
    if args.hrtem_images > 0:
        out["hrtem_batch"] = run_hrtem_batch(args.hrtem_images, 8, seed=1006)
    if args.stress_scan > 0:
        out["precision_stress"] = run_precision_stress(args.stress_scan)
    if args.erf_build:
        out["erf_weighted_build"] = run_erf_build()
    if args.drop_in:
        out["drop_in_call"] = run_drop_in()
    if args.uneven_jobs > 0:
        out["uneven_batch"] = run_uneven(args.uneven_jobs, seed=1004, arena_gib=0.25)
    if args.serve_requests > 0:
        out["serving"] = run_serving(args.serve_requests, args.serve_rate, seed=1005, arena_gib=0.25, budget_gb=288)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
