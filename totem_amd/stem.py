"""STEM simulation driver: configuration -> myProjectfun inputs -> engine.

Mirrors the reference's GUI path ``STEMsimulation`` (``ToTEM_submit.m:1476-1616``)
and its marshalling script ``cuda_stem.m:1-107``: it builds the specimen rows
(:mod:`totem_amd.specimen`), the optics (:mod:`totem_amd.optics`) and then the
41 mex arguments, and calls the engine variant chosen by ``paraflag``
('l' -> waveCudalobato, 'p'/'n' -> waveCudaPengAndCorr).
"""
from dataclasses import dataclass, field, replace

import numpy as np

from . import optics, specimen


@dataclass
class StemConfig:
    atoms: np.ndarray                    # [n, 8] = [Z x y z valence ionicity B occ], x/y inside the potential box
    np_side: int                         # potential grid side Np (P_Height = P_Width)
    dx: float                            # sampling, Angstrom/pixel
    n: int                               # probe window side N (Height = Width)
    kv: float                            # accelerating voltage
    alpha_mrad: float                    # probe convergence semi-angle
    slice_thickness: float               # eachthick, Angstrom
    n_slices: int
    slice_origin: float = 0.0            # slicethick[0]
    paraflag: str = "l"                  # 'l' Lobato, 'p' Peng, 'n' Peng + correction
    aberrations: optics.Aberrations = None
    focal_members: int = 1               # mtotal
    focal_rms_nm: float = 0.0
    source_mrad: float = 0.0             # spatial-coherence semi-angle (alafa)
    detectors_mrad: tuple = ((60.0, 200.0),)
    width_red: int = 1                   # scan columns
    hight_red: int = 1                   # scan rows
    step: int = 1                        # scan step, pixels
    begin_row: int = 1                   # 1-based window origin of the first position (probe_ingreenNrow)
    begin_col: int = 1
    mid_slices: tuple = ()               # 1-based mid-slice outputs
    multiceng: int = 0
    dw_selected: bool = False            # DW attenuation (radiobutton14/4)
    tilt_mrad: float = 0.0
    tilt_phi_deg: float = 0.0

    @property
    def depth(self):
        return self.width_red * self.hight_red


@dataclass
class StemInputs:
    cfg: StemConfig
    ele_n: np.ndarray
    ele_n_i: np.ndarray
    series_n: np.ndarray
    series_n_i: np.ndarray
    ele_n_corr: np.ndarray
    series_n_corr: np.ndarray
    corr_info: np.ndarray
    corr_matrix: np.ndarray              # [cim_len, Np, Np] or None
    gx: np.ndarray
    gy: np.ndarray
    s2: np.ndarray
    aper_true: np.ndarray
    parameter: float
    sigma: float
    probe: np.ndarray                    # [M, N, N] complex128
    gfsf: np.ndarray
    psf: np.ndarray                      # [N, N] complex128, ifftshifted
    aper: np.ndarray                     # [A, N, N] centred
    aper2: np.ndarray                    # [A, N, N] ifftshifted
    extra: dict = field(default_factory=dict)


def slice_boundaries(cfg):
    return cfg.slice_origin + cfg.slice_thickness * np.arange(cfg.n_slices)


def prepare_specimen(cfg, atoms=None):
    """Per-slice coefficient rows of one specimen snapshot: dict with ele_n, ele_n_i, series_n,
    series_n_i, ele_n_corr, series_n_corr, corr_info, corr_matrix."""
    inp = prepare(cfg, atoms, optics_too=False)
    return {k: getattr(inp, k) for k in ("ele_n", "ele_n_i", "series_n", "series_n_i", "ele_n_corr",
                                         "series_n_corr", "corr_info", "corr_matrix")}


def phonon_snapshots(cfg, count, sigma_by_z, seed=0):
    """Frozen-phonon atom matrices: every atom displaced by an isotropic Gaussian of rms sigma_by_z[Z]
    (B = 8 pi^2 sigma^2, ``randshiftpos.m:4``), redrawn to stay inside the slab.  Seeded."""
    x = cfg.atoms.copy()
    for z, sg in sigma_by_z.items():
        x[x[:, 0] == z, 6] = 8 * np.pi ** 2 * sg ** 2
    st = slice_boundaries(cfg)
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        xs = specimen.frozen_phonon(x, cfg.slice_thickness, st, tuple(cfg.mid_slices), rng, bcol=6)
        xs[:, 6] = 0.0  # frozen phonon only: no Debye-Waller attenuation on top of the displacements
        out.append(xs)
    return out


def specimen_rows(cfg, atoms=None):
    """Every atom's coefficient rows as ONE slice, for slicing on the device.

    Returns (rows, term_atom, corr_matrix): rows holds series_n, ele_n, series_n_i, ele_n_i,
    series_n_corr, ele_n_corr exactly as :func:`prepare` builds them for a single slice that holds
    the whole specimen (multiceng = 0, so no erf weighting); term_atom gives the atom (row of
    ``atoms``) of each row in the order ele_n, ele_n_i, ele_n_corr -- the order
    ``getpotentialpara_lotabo_peng.m`` and ``getpotentialpara_corr_peng.m`` emit them in (neutral,
    then fully ionic, then mixed; corrections for neutral atoms)."""
    x = np.asarray(cfg.atoms if atoms is None else atoms, dtype=np.float64)
    if len(x) == 0:
        raise ValueError("specimen has no atoms")
    z = x[:, 3]
    one = replace(cfg, atoms=x, n_slices=1, slice_origin=float(z.min()) - 1.0,
                  slice_thickness=float(z.max() - z.min()) + 2.0, multiceng=0, mid_slices=())
    inp = prepare(one, x, optics_too=False)
    ion = x[:, 5]
    ta = [np.nonzero(ion == 0)[0], np.nonzero(ion == 1)[0], np.nonzero((ion > 0) & (ion < 1))[0]]
    if cfg.paraflag == "n":
        ta.append(np.nonzero(ion == 0)[0])
    flag = cfg.paraflag
    rows = dict(series_n=np.asarray(inp.series_n[:1], dtype=np.float64), ele_n=inp.ele_n,
                series_n_i=np.full(1, float(len(inp.ele_n_i))),
                ele_n_i=inp.ele_n_i if len(inp.ele_n_i) else np.zeros((0, 26 if flag == "l" else 25)))
    if flag == "n":
        rows["series_n_corr"] = np.asarray(inp.series_n_corr[:1], dtype=np.float64)
        rows["ele_n_corr"] = inp.ele_n_corr
    return rows, np.concatenate(ta).astype(np.int32), inp.corr_matrix


def device_specimen(cfg, sigma_by_z=None, rms=None, atoms=None):
    """The specimen resident on the GPU (:class:`totem_amd.engine.DeviceSpecimen`), displaced and
    sliced there per frozen-phonon configuration.  Displacement rms per atom from ``sigma_by_z``
    ({Z: rms}, Angstrom; the Debye-Waller column is then cleared as in :func:`phonon_snapshots`) or
    ``rms`` directly; neither gives a static specimen.  ``cfg.multiceng`` > 0 (Peng) spreads atoms
    near faces over neighbouring slices as :func:`prepare` does."""
    from . import engine
    x = np.array(cfg.atoms if atoms is None else atoms, dtype=np.float64)
    if sigma_by_z is not None:
        rms = np.array([sigma_by_z.get(float(zz), sigma_by_z.get(int(zz), 0.0)) for zz in x[:, 0]])
        x[:, 6] = 0.0
    rows, term_atom, cim = specimen_rows(cfg, x)
    gx, gy, s2 = optics.reciprocal_grids(cfg.np_side, cfg.dx)
    return engine.DeviceSpecimen(cfg.paraflag, cfg.np_side, rows, term_atom, x[:, 1:4], rms, gx, gy, s2,
                                 np.ones_like(s2), optics.interaction_sigma(cfg.kv),
                                 optics.potential_parameter(cfg.np_side, cfg.dx), cfg.n_slices, cfg.slice_origin,
                                 cfg.slice_thickness, tuple(cfg.mid_slices), cim, cfg.multiceng)


def prepare(cfg, atoms=None, optics_too=True):
    """All engine inputs for one configuration (one frozen-phonon snapshot if ``atoms`` is given)."""
    x = cfg.atoms if atoms is None else atoms
    flag = cfg.paraflag
    st = slice_boundaries(cfg)
    all_n, all_n_i = specimen.atom_parameters(x, flag)
    all_n, all_n_i = specimen.fold_debye_waller(all_n, all_n_i, flag, cfg.dw_selected)
    multiceng = 0 if flag == "l" else cfg.multiceng  # ToTEM_submit.m:874-877
    dbmode = 1.0 if cfg.dw_selected else 0.0
    ele_n, ele_n_i, series_n, series_n_i = specimen.slice_atoms(all_n, all_n_i, cfg.slice_thickness, st,
                                                               multiceng, flag, dbmode)
    gx, gy, s2 = optics.reciprocal_grids(cfg.np_side, cfg.dx)
    ele_n_corr = np.zeros((0, 5))
    series_n_corr = np.zeros(cfg.n_slices)
    corr_info = np.zeros((0, 20))
    corr_matrix = None
    if flag == "n":
        cn = specimen.corr_parameters(x, cfg.dw_selected)
        ele_n_corr, corr_info, series_n_corr = specimen.slice_corr(cn, cfg.slice_thickness, st, multiceng)
        corr_matrix = specimen.corr_info_matrix(corr_info, s2)
    if optics_too:
        probe, gfsf = optics.probe_stack(cfg.n, cfg.dx, cfg.kv, cfg.alpha_mrad, cfg.aberrations, cfg.focal_members,
                                         cfg.focal_rms_nm, cfg.source_mrad, cfg.tilt_mrad, cfg.tilt_phi_deg)
        psf = optics.propagator(cfg.n, cfg.dx, cfg.kv, cfg.slice_thickness, cfg.tilt_mrad, cfg.tilt_phi_deg)
        aper = optics.detector_masks(cfg.n, cfg.dx, cfg.kv, cfg.detectors_mrad)
        aper2 = np.fft.ifftshift(aper, axes=(-2, -1)).copy()
    else:
        probe = gfsf = psf = aper = aper2 = None
    if len(series_n) < cfg.n_slices:
        series_n = np.concatenate([series_n, np.zeros(cfg.n_slices - len(series_n))])
    return StemInputs(cfg=cfg, ele_n=ele_n, ele_n_i=ele_n_i, series_n=series_n, series_n_i=series_n_i,
                      ele_n_corr=ele_n_corr, series_n_corr=series_n_corr, corr_info=corr_info,
                      corr_matrix=corr_matrix, gx=gx, gy=gy, s2=s2, aper_true=np.ones_like(s2),
                      parameter=optics.potential_parameter(cfg.np_side, cfg.dx),
                      sigma=optics.interaction_sigma(cfg.kv), probe=probe, gfsf=gfsf, psf=psf, aper=aper,
                      aper2=aper2)


def mex_args(inp):
    """The 41 mex arguments exactly as ``cuda_stem.m:22-97`` assembles them (row-major numpy views)."""
    c = inp.cfg
    S = c.n_slices
    series_n_i = inp.series_n_i if len(inp.series_n_i) else np.zeros(S)
    mid = np.asarray(sorted(c.mid_slices), dtype=np.int32)
    flag = c.paraflag
    corr = flag == "n" and inp.corr_matrix is not None and len(inp.corr_matrix) > 0
    return [
        np.ascontiguousarray(inp.probe.real), np.ascontiguousarray(inp.probe.imag), inp.aper,
        inp.series_n, None, series_n_i, None,
        inp.ele_n, inp.ele_n_i if len(inp.ele_n_i) else np.zeros((0, 26 if flag == "l" else 25)),
        inp.gx, inp.gy, inp.s2,
        np.ascontiguousarray(inp.psf.real), np.ascontiguousarray(inp.psf.imag), inp.gfsf,
        c.n, c.n, c.depth, c.np_side, c.np_side, S, c.step, c.begin_row, c.begin_col, c.width_red,
        inp.sigma, len(inp.gfsf), inp.aper.shape[0], inp.parameter, inp.aper_true,
        0, 0, mid, len(mid), inp.aper2,
        inp.series_n_corr if corr else np.zeros(S), np.zeros(S),
        inp.ele_n_corr if corr else np.zeros((0, 5)), np.zeros((0, 8)),
        inp.corr_matrix if corr else None, len(inp.corr_matrix) if corr else 0,
    ]


def mex_name(paraflag, potential_only=False):
    """``cuda_stem.m:82-95``: 'l' -> lobato, 'p'/'n' -> PengAndCorr; 'P' suffix for the potential build."""
    base = "waveCudalobato" if paraflag == "l" else "waveCudaPengAndCorr"
    return base + ("P" if potential_only else "")


def run(inp, potential_only=False, drop_in=False):
    """Run the engine through the mex-shaped entry point; returns the four mex outputs.  With
    ``drop_in`` the call goes through ``myProjectfun`` of the matching drop-in library."""
    from . import engine
    return engine.wave_cuda(mex_name(inp.cfg.paraflag, potential_only), *mex_args(inp), drop_in=drop_in)


# ------------------------------------------------------------------ workloads

def perovskite_c1(n=256, np_side=256, scan=16, step=3):
    """The small Lobato workload: SrTiO3 5x5x10 cells (1250 atoms), 256^2 over 19.525 A, 20 slices of
    1.9525 A, 200 kV, 20 mrad, ADF 60-200 mrad, 16x16 scan at 3 px steps (periodic windows)."""
    a = 3.905
    x = specimen.perovskite(5, 5, 10, a=a)
    dz = a / 2
    return StemConfig(atoms=x, np_side=np_side, dx=5 * a / np_side, n=n, kv=200.0, alpha_mrad=20.0,
                      slice_thickness=dz, n_slices=20, slice_origin=-dz / 4, paraflag="l",
                      detectors_mrad=((60.0, 200.0),), width_red=scan, hight_red=scan, step=step,
                      begin_row=1, begin_col=1)


def perovskite_c2(scan=64, step=3):
    """The frozen-phonon workload: SrTiO3 20x20x40 cells (80 000 atoms), 1024^2 over 78.1 A, 80 slices of
    1.9525 A, Peng + correction, 300 kV, 24 mrad, Cs = 1 um, C1 = -5 nm, ADF 70-200 and BF 0-10 mrad,
    64x64 scan at 3 px (periodic windows).  Phonon rms 0.08 A (Sr, Ti), 0.10 A (O)."""
    a = 3.905
    x = specimen.perovskite(20, 20, 40, a=a)
    dz = a / 2
    return StemConfig(atoms=x, np_side=1024, dx=20 * a / 1024, n=1024, kv=300.0, alpha_mrad=24.0,
                      slice_thickness=dz, n_slices=80, slice_origin=-dz / 4, paraflag="n",
                      aberrations=optics.Aberrations(C1=-5.0, Cs_um=1.0),
                      detectors_mrad=((70.0, 200.0), (0.0, 10.0)), width_red=scan, hight_red=scan, step=step,
                      begin_row=1, begin_col=1)


C2_PHONON_SIGMA = {38: 0.08, 22: 0.08, 8: 0.10}


def precision_c3(scan=32, side=100.0, thick=300.0, cap=50.0, n=2048):
    """The precision-stress workload: diamond-cubic Si (a = 5.431 A) viewed along [110], about
    ``side`` x ``side`` x ``thick`` A (whole repeats), capped top and bottom by ``cap`` A of amorphous
    carbon at 2.0 g/cm^3 (0.1003 atoms/A^3) with 1.4 A minimum separation (seeded), on an n^2 grid,
    slices of 1.92 A, 200 kV, 22 mrad, Lobato, 32 x 32 scan over one [110] repeat (periodic windows)."""
    a = 5.431
    ax, ay = a / np.sqrt(2), a
    nx = int(np.ceil(side / ax))
    ny = int(round(nx * ax / ay))
    box = nx * ax
    si = specimen.diamond_110(nx * ax, ny * ay, thick, a=a, origin=(0.0, 0.0, cap))
    si = si[(si[:, 1] < box) & (si[:, 2] < box)]
    z_si = si[:, 3].max() + 0.5
    top = specimen.amorphous(box, box, 0.0, cap, Z=6, seed=1003)
    bottom = specimen.amorphous(box, box, z_si, z_si + cap, Z=6, seed=2003)
    atoms = np.concatenate([top, si, bottom])
    dz = 1.92
    S = int(np.ceil((z_si + cap) / dz))
    dx = box / n
    step = max(1, int(round(ax / dx / scan)))
    return StemConfig(atoms=atoms, np_side=n, dx=dx, n=n, kv=200.0, alpha_mrad=22.0, slice_thickness=dz,
                      n_slices=S, slice_origin=-dz / 4, paraflag="l", detectors_mrad=((60.0, 200.0),),
                      width_red=scan, hight_red=scan, step=step, begin_row=1, begin_col=1)
