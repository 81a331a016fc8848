"""Serving: STEM requests of mixed grid sizes on one GPU with continuous batching.

A request is one STEM simulation (specimen, optics, scan).  The loop
(:func:`serve`) admits arrived requests while their device footprint fits the
HBM budget and hands them to a builder thread, which prepares each request on
the host (specimen slicing, optics) and builds its transmission stack on a
stream of its own; meanwhile the loop keeps stepping the engine's served form
(:class:`engine.Server`, which advances probes of all in-flight jobs together,
one slice per step, and admits new probes as wave slots free up) and submits
each request as soon as its build has finished.  Latency runs from arrival to
results on the host, host preparation included.

This is the request loop around the reference's one-call-per-simulation
``myProjectfun`` (``cuda_source_lobato.cu:1018-1562``): the arithmetic of each
job is the same as a call of its own.
"""
import queue
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import engine, optics, specimen, stem

GiB = 1 << 30


class DeviceJob:
    """Device-resident inputs and outputs of one STEM job (torch tensors on cuda)."""

    def __init__(self, inp, dev="cuda", stream=None, wrap=None, exit_wave=False):
        import torch
        c = inp.cfg
        self.inp = inp
        self.S, self.N, self.NP = c.n_slices, c.n, c.np_side
        self.M, self.A, self.D, self.L = len(inp.gfsf), inp.aper.shape[0], c.depth, len(c.mid_slices)
        wrap = (c.n == c.np_side) if wrap is None else wrap
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        self.t = torch.empty((self.S, self.NP, self.NP), dtype=torch.complex64, device=dev)
        corr = c.paraflag == "n" and inp.corr_matrix is not None and len(inp.corr_matrix) > 0
        engine.build_transmission(
            c.paraflag, self.NP, self.S, inp.series_n, inp.ele_n, inp.gx, inp.gy, inp.s2, inp.aper_true,
            inp.sigma, inp.parameter, self.t.data_ptr(),
            atom_slice_i=inp.series_n_i if len(inp.series_n_i) else None,
            proj_coff_mat_i=inp.ele_n_i if len(inp.ele_n_i) else None,
            series_n_corr=inp.series_n_corr if corr else None, ele_n_corr=inp.ele_n_corr if corr else None,
            corr_info_matrix=inp.corr_matrix if corr else None, stream=st)
        self.probes = torch.tensor(inp.probe.astype(np.complex64), device=dev)
        self.prop = torch.tensor(optics.unit_complex64(inp.psf), device=dev)
        self.aper = torch.tensor(inp.aper, device=dev)
        self.aper2 = torch.tensor(inp.aper2, device=dev)
        self.det = torch.zeros((self.L + 1, self.A, self.M * self.D), dtype=torch.float64, device=dev)
        self.exit = (torch.zeros((self.M * self.D, self.N, self.N), dtype=torch.complex64, device=dev)
                     if exit_wave else None)
        self.job = engine.stem_job(self.N, self.NP, self.S, self.t.data_ptr(), self.probes.data_ptr(), self.M,
                                   self.prop.data_ptr(), self.aper.data_ptr(), self.A, self.D, c.width_red, c.step,
                                   c.begin_row - 1, c.begin_col - 1, self.det.data_ptr(),
                                   aper2_ptr=self.aper2.data_ptr(), mid=sorted(c.mid_slices), wrap=wrap,
                                   exit_ptr=None if self.exit is None else self.exit.data_ptr())
        # inputs ready before another stream reads them (the copies above ran on the current
        # stream, the build on `st`)
        torch.cuda.current_stream().synchronize()
        if stream is not None:
            torch.cuda.ExternalStream(stream).synchronize()

    @property
    def probe_slices(self):
        return self.M * self.D * self.S

    @property
    def footprint(self):
        return footprint_bytes(self.inp.cfg, self.M, self.A)

    def images(self):
        """(myresul [D, A], mid [D*A, L]) with the focal-spread weighting, as the mex returns them."""
        det = self.det.cpu().numpy()
        g = self.inp.gfsf
        w = det.reshape(self.L + 1, self.A, self.M, self.D)
        img = np.einsum("lamd,m->lad", w, g)
        myresul = img[self.L].T.copy()
        mid = np.zeros((self.D * self.A, self.L))
        for a in range(self.A):
            mid[a * self.D:(a + 1) * self.D, :] = img[:self.L, a, :].T
        return myresul, mid


def footprint_bytes(cfg, M=None, A=None):
    """Device bytes a request holds while in service: transmission stack, probes, propagator,
    masks and outputs, plus the potential builder's phase tables for its build (about 16 bytes per
    term and pixel row of the largest slice, counted with the specimen's rows per slice; the
    builder releases large tables after a build).  Wave slots come from the server's arena."""
    n2, np2 = cfg.n * cfg.n, cfg.np_side * cfg.np_side
    L = len(cfg.mid_slices)
    M = focal_count(cfg) if M is None else M
    A = len(cfg.detectors_mrad) if A is None else A
    rows = 2 * len(cfg.atoms) / max(cfg.n_slices, 1) + 64     # corrections double the rows
    build = int(16 * rows * cfg.np_side)
    return int(8 * cfg.n_slices * np2 + 8 * M * n2 + 8 * n2 + 16 * A * n2 + 8 * (L + 1) * A * M * cfg.depth) + build


def focal_count(cfg):
    """Focal-spread members the probe stack of ``cfg`` will hold (``Gaussian_focal``: one member
    unless (mtotal - 1) // 2 > 0)."""
    return cfg.focal_members if (cfg.focal_members - 1) // 2 > 0 else 1


# ----------------------------------------------------------------- request mix

@dataclass
class Request:
    rid: int
    arrival: float
    cfg: object
    kind: str = ""
    inp: object = None          # prepared inputs (prepared by the serving loop's builder when None)
    admitted: float = None
    built: float = None
    done: float = None
    job: DeviceJob = None
    jid: int = None
    result: tuple = None
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


GRID_MIX = ((256, 0.50), (512, 0.30), (1024, 0.15), (2048, 0.05))


def mixed_request(rng, grids=GRID_MIX, slices=(10, 120), scan=(8, 64), cells=5):
    """One seeded request of the mixed serving load: grid side from ``grids`` (side, probability),
    thickness uniform in ``slices``, square scan of side uniform in ``scan``, Lobato or Peng 50/50,
    SrTiO3 (``cells`` x ``cells`` cells laterally) or amorphous carbon of the same box, 200 kV,
    20 mrad, ADF 60-200 mrad, periodic windows over one unit cell."""
    sides, p = zip(*grids)
    n = int(rng.choice(sides, p=np.asarray(p) / np.sum(p)))
    S = int(rng.integers(slices[0], slices[1] + 1))
    side = int(rng.integers(scan[0], scan[1] + 1))
    flag = "l" if rng.random() < 0.5 else "p"
    a = 3.905
    dz = a / 2
    box = cells * a
    if rng.random() < 0.5:
        kind = "perovskite"
        atoms = specimen.perovskite(cells, cells, max(1, int(np.ceil(S / 2))), a=a)
    else:
        kind = "amorphous-C"
        atoms = specimen.amorphous(box, box, 0.0, S * dz, Z=6, seed=int(rng.integers(1 << 30)))
    dx = box / n
    step = max(1, int(round(a / dx / side)))
    cfg = stem.StemConfig(atoms=atoms, np_side=n, dx=dx, n=n, kv=200.0, alpha_mrad=20.0, slice_thickness=dz,
                          n_slices=S, slice_origin=-dz / 4, paraflag=flag, detectors_mrad=((60.0, 200.0),),
                          width_red=side, hight_red=side, step=step, begin_row=1, begin_col=1)
    return cfg, kind


def uneven_request(rng):
    """One job of the uneven batch: SrTiO3 with a seeded supercell of 3-8 cells laterally, grid side
    from {256, 512, 768, 1024}, 10-60 slices, defocus uniform in [-10, 10] nm, square scan of side
    8-32 over one unit cell, 200 kV, 20 mrad, ADF 60-200 mrad, Lobato."""
    n = int(rng.choice((256, 512, 768, 1024)))
    S = int(rng.integers(10, 61))
    side = int(rng.integers(8, 33))
    cells = int(rng.integers(3, 9))
    a = 3.905
    dz = a / 2
    atoms = specimen.perovskite(cells, cells, max(1, int(np.ceil(S / 2))), a=a)
    dx = cells * a / n
    step = max(1, int(round(a / dx / side)))
    ab = optics.Aberrations(C1=float(rng.uniform(-10.0, 10.0)))
    return stem.StemConfig(atoms=atoms, np_side=n, dx=dx, n=n, kv=200.0, alpha_mrad=20.0, slice_thickness=dz,
                           n_slices=S, slice_origin=-dz / 4, paraflag="l", aberrations=ab,
                           detectors_mrad=((60.0, 200.0),), width_red=side, hight_red=side, step=step,
                           begin_row=1, begin_col=1)


def run_all(server, jobs, max_wall=600.0):
    """Submit every DeviceJob at once and step the server until all are done; returns wall seconds."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ids = {server.submit(dj.job) for dj in jobs}
    done = set()
    while not ids <= done:
        server.step()
        done.update(server.poll())
        if time.perf_counter() - t0 > max_wall:
            raise RuntimeError("server run exceeded %.0f s" % max_wall)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


# ----------------------------------------------------------------- serving loop

def serve(requests, server, hbm_budget, max_wall=600.0, on_done=None):
    """Run the requests (sorted by ``arrival``, seconds after the start) through ``server``.

    A request is admitted when it has arrived and its footprint fits ``hbm_budget`` next to the
    requests in service or being built (first come, first served; a request larger than the budget
    is admitted alone).  A builder thread prepares admitted requests (``stem.prepare`` when
    ``inp`` is None) and builds their transmission stacks on its own stream; the loop submits each
    one when its build is complete and keeps stepping the server in between.  Returns (wall
    seconds, peak resident bytes); each request gets ``admitted``, ``built``, ``done`` (seconds),
    ``extra['prepare_s']`` and ``result`` (mex-shaped images)."""
    import torch
    waiting = deque(sorted(requests, key=lambda r: r.arrival))
    arrived = deque()
    in_service = {}
    todo, ready = queue.Queue(), queue.Queue()
    errors = []
    dev = torch.cuda.current_device()
    t0 = time.perf_counter()

    def builder():
        torch.cuda.set_device(dev)
        st = torch.cuda.Stream(device=dev)
        while True:
            r = todo.get()
            if r is None:
                return
            try:
                tp = time.perf_counter()
                if r.inp is None:
                    r.inp = stem.prepare(r.cfg)
                r.extra["prepare_s"] = time.perf_counter() - tp
                with torch.cuda.stream(st):
                    r.job = DeviceJob(r.inp, stream=st.cuda_stream)
                r.built = time.perf_counter() - t0
            except BaseException as e:   # reported on the loop thread
                errors.append(e)
            ready.put(r)

    th = threading.Thread(target=builder, name="totem-serve-builder", daemon=True)
    th.start()
    building = 0
    resident = peak = 0
    try:
        while waiting or arrived or in_service or building:
            now = time.perf_counter() - t0
            if now > max_wall:
                raise RuntimeError("serving loop exceeded %.0f s" % max_wall)
            while waiting and waiting[0].arrival <= now:
                arrived.append(waiting.popleft())
            while arrived:
                r = arrived[0]
                fp = footprint_bytes(r.cfg)
                if (in_service or building) and resident + fp > hbm_budget:
                    r.extra["budget_wait"] = True
                    break
                arrived.popleft()
                r.admitted = time.perf_counter() - t0
                resident += fp
                peak = max(peak, resident)
                building += 1
                todo.put(r)
            while True:
                try:
                    r = ready.get_nowait()
                except queue.Empty:
                    break
                building -= 1
                if errors:
                    raise errors[0]
                r.jid = server.submit(r.job.job)
                in_service[r.jid] = r
            if in_service:
                server.step()
                for jid in server.poll():
                    r = in_service.pop(jid)
                    r.result = r.job.images()
                    r.done = time.perf_counter() - t0
                    resident -= footprint_bytes(r.cfg)
                    if on_done:
                        on_done(r)
                    r.job = None
            else:
                nxt = waiting[0].arrival - (time.perf_counter() - t0) if waiting else 0.001
                time.sleep(max(0.0, min(0.002, nxt)))
    finally:
        todo.put(None)
        th.join()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, peak
