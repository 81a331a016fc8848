"""Writes the golden fixtures of the STEM path: ``python tests/golden/make_golden.py``.

Inputs are the reference's own specimen file ``allresul.mat`` (a 791-atom SrTiO3
model: ``x_new`` = [Z x y z valence ionicity B occ], ``slicethick`` = 41 slice
faces 0.823175 A apart, ``eachthick``), read as data with ``scipy.io.loadmat``
from the reference checkout and stored here as ``allresul_specimen.npz``.  The
expected outputs are those of the FP64 oracle (``oracle/``) on the cases of
``CASES`` (Lobato at 200 kV; Peng + correction at 300 kV with a 3-member focal
series), both on a 256^2 potential over 36 A with a 128^2 probe window:

* potential slices: a 64 x 64 crop (rows/columns 96..159) of V at slices 0, 20
  and 40 (complex64) plus, for every slice, sum V and sum |V|^2 (FP64);
* the mex outputs of a 4 x 4 scan: ``myresul`` [D, A] and ``mid_ceng_mat`` [D*A, L].

A second set comes from the reference's Python-tool input file ``pydata.mat`` (read as data): its
parameter structs, U/V grids and exit wave (320^2) go to ``pydata_inputs.npz``; the FP64 oracle's
batch HRTEM images for seeded draws (``HRTEM_CASES``) go to ``pydata_golden.npz`` as the crops the
tool's consumer cuts (``position`` from ``othervar``) plus per-image sums.

A third set, ``allresul_phonon.npz``, holds frozen-phonon displacements of the stored specimen: the
oracle's positions (``oracle/phonon.py``) for ``PHONON`` (seed, configurations, rms per element,
mid-slice faces) and the per-slice atom counts of each configuration.  It needs only the stored
specimen: ``python tests/golden/make_golden.py phonon``.

Only this script reads the reference checkout; the tests read the ``.npz`` files.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/allresul.mat"

CROP = (slice(96, 160), slice(96, 160))
V_SLICES = (0, 20, 40)
CASES = {
    "lobato": dict(paraflag="l", kv=200.0, alpha_mrad=20.0, detectors_mrad=((30.0, 59.0), (0.0, 10.0)),
                   mid_slices=(10, 25), focal_members=1, focal_rms_nm=0.0, aberrations=None),
    "pengcorr": dict(paraflag="n", kv=300.0, alpha_mrad=20.0, detectors_mrad=((25.0, 46.0), (0.0, 10.0)),
                     mid_slices=(20,), focal_members=3, focal_rms_nm=3.0, aberrations=(-5.0, 1.0)),
}


def config(spec, name):
    """StemConfig of golden case ``name`` on the stored specimen (no reference access)."""
    sys.path.insert(0, ROOT)
    from totem_amd import optics, stem
    c = dict(CASES[name])
    ab = c.pop("aberrations")
    ab = optics.Aberrations(C1=ab[0], Cs_um=ab[1]) if ab else None
    dz = float(spec["eachthick"])
    faces = np.asarray(spec["slicethick"], dtype=np.float64)
    cfg = stem.StemConfig(atoms=np.asarray(spec["atoms"], dtype=np.float64), np_side=256, dx=36.0 / 256, n=128,
                          slice_thickness=dz, n_slices=len(faces), slice_origin=float(faces[0]), aberrations=ab,
                          width_red=4, hight_red=4, step=6, begin_row=40, begin_col=44, **c)
    assert np.allclose(stem.slice_boundaries(cfg), faces, rtol=0, atol=1e-9)
    return cfg


def oracle_outputs(cfg):
    """(V [S, Np, Np] complex128, myresul [D, A], mid [D*A, L]) from the FP64 oracle."""
    sys.path.insert(0, ROOT)
    from oracle import multislice as oms
    from oracle import potential as opot
    from totem_amd import stem
    inp = stem.prepare(cfg)
    kw = {}
    if cfg.paraflag == "n":
        kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
    V, t = opot.potential_slices(cfg.np_side, inp.ele_n, None, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                 inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, cfg.paraflag, **kw)
    res, mid = oms.stem_cpu(t, inp.probe, inp.gfsf, inp.psf, cfg.begin_row, cfg.begin_col, cfg.width_red,
                            cfg.hight_red, cfg.step, inp.aper, tuple(sorted(cfg.mid_slices)))
    return V, res, mid


def summarise(V, res, mid):
    return {"v_crop": V[(slice(None),) + CROP][list(V_SLICES)].astype(np.complex64),
            "v_sum": V.sum(axis=(1, 2)), "v_norm2": (np.abs(V) ** 2).sum(axis=(1, 2)),
            "myresul": res, "mid": mid}


REF_PY = "/root/reference/pydata.mat"
HRTEM_CASES = {"pydata": dict(seed=1006, batches=3),
               "focal3": dict(seed=1007, batches=2, mtotal=3, delta_yita=4.0, gfsf=0.6)}


def hrtem_case(inputs, name):
    """(para_part1, ai [B, 21]) of golden HRTEM case ``name`` on the stored pydata inputs."""
    c = dict(HRTEM_CASES[name])
    rng = np.random.Generator(np.random.PCG64(c.pop("seed")))
    ai = rng.random((c.pop("batches"), 21))
    pp1 = {k: float(inputs["pp1_" + k]) for k in ("sampling", "mtotal", "alafa", "lambda", "gmax", "gfsf",
                                                   "delta_yita")}
    pp1.update({k: float(v) for k, v in c.items()})
    return pp1, ai


def crop(inputs, img):
    ov = inputs["othervar"]
    top, left, width, height = int(ov[3]), int(ov[4]), int(ov[5]), int(ov[6])
    return img[..., left:left + width, top:top + height]


def hrtem_golden(inputs):
    sys.path.insert(0, ROOT)
    from oracle import hrtem as oh
    out = {}
    for name in HRTEM_CASES:
        pp1, ai = hrtem_case(inputs, name)
        paras = oh.new_pp2(inputs["pp2"], inputs["mrr"], ai)
        img = oh.images(inputs["mywave"], paras, inputs["U"], inputs["V"], pp1["lambda"], pp1["alafa"],
                        int(pp1["mtotal"]), pp1["delta_yita"], pp1["gfsf"])
        out[name + "_crop"] = crop(inputs, img)
        out[name + "_sum"] = img.sum(axis=(1, 2))
        out[name + "_paras"] = paras
    return out


PHONON = dict(seed=20241202, configs=(0, 1, 5), rms_by_z={38: 0.08, 22: 0.08, 8: 0.10}, mid=(10, 25))


def phonon_golden():
    sys.path.insert(0, ROOT)
    from oracle import phonon as oph
    spec = np.load(os.path.join(HERE, "allresul_specimen.npz"))
    atoms = spec["atoms"]
    faces = spec["slicethick"]
    dz = float(spec["eachthick"])
    S = len(faces)
    rms = np.array([PHONON["rms_by_z"][int(z)] for z in atoms[:, 0]])
    lo, hi, mid_faces = oph.slab(float(faces[0]), dz, S, PHONON["mid"])
    out = {"seed": np.uint64(PHONON["seed"]), "configs": np.asarray(PHONON["configs"], dtype=np.int32),
           "rms": rms, "mid": np.asarray(PHONON["mid"], dtype=np.int32)}
    pos, counts = [], []
    for c in PHONON["configs"]:
        p = oph.displace(atoms[:, 1:4], rms, PHONON["seed"], c, lo, hi, mid_faces)
        pos.append(p)
        counts.append([len(m) for m in oph.slice_members(p[:, 2], float(faces[0]), dz, S)])
    out["positions"] = np.stack(pos)
    out["slice_counts"] = np.asarray(counts, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "allresul_phonon.npz"), **out)
    print("phonon", out["positions"].shape, out["slice_counts"].sum(axis=1), flush=True)


def main():
    import scipy.io as sio
    m = sio.loadmat(REF_PY)
    p1 = m["para_part1"][0][0]
    names = ("sampling", "mtotal", "alafa", "lambda", "gmax", "gfsf", "delta_yita", "tilt", "phitilt", "tiltx",
             "tilty")
    inputs = {"pp1_" + k: np.float64(np.asarray(p1[i]).ravel()[0]) for i, k in enumerate(names)}
    inputs["pp2"] = np.array([float(np.asarray(x).ravel()[0]) for x in m["para_part2"][0][0]])
    inputs["mrr"] = np.array([float(np.asarray(x).ravel()[0]) for x in m["myrandrange"][0][0]])
    inputs["U"] = np.asarray(m["U"], dtype=np.float64)
    inputs["V"] = np.asarray(m["V"], dtype=np.float64)
    inputs["mywave"] = np.asarray(m["mywave"], dtype=np.complex128)
    inputs["othervar"] = np.asarray(m["othervar"], dtype=np.float64).ravel()
    inputs["commpara"] = np.asarray(m["commpara"], dtype=np.float64).ravel()
    np.savez_compressed(os.path.join(HERE, "pydata_inputs.npz"), **inputs)
    np.savez_compressed(os.path.join(HERE, "pydata_golden.npz"), **hrtem_golden(inputs))
    d = sio.loadmat(REF)["allresul"][0, 0]
    atoms = np.asarray(d["x_new"], dtype=np.float64)
    faces = np.asarray(d["slicethick"], dtype=np.float64).ravel()
    dz = float(np.asarray(d["eachthick"]).ravel()[0])
    gap = np.min(np.abs(atoms[:, 3][:, None] - faces[None, :]))
    assert gap > 1e-6, "an atom sits on a slice face"
    spec = {"atoms": atoms, "slicethick": faces, "eachthick": np.float64(dz)}
    np.savez_compressed(os.path.join(HERE, "allresul_specimen.npz"), **spec)
    out = {}
    for name in CASES:
        V, res, mid = oracle_outputs(config(spec, name))
        for k, v in summarise(V, res, mid).items():
            out["%s_%s" % (name, k)] = v
        print(name, "myresul", res.ravel()[:4], "mid", mid.shape, flush=True)
    np.savez_compressed(os.path.join(HERE, "allresul_golden.npz"), **out)
    phonon_golden()


if __name__ == "__main__":
    if sys.argv[1:] == ["phonon"]:
        phonon_golden()
    else:
        main()
