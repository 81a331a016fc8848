"""Frozen-phonon displacement and specimen slicing on the device (randshiftpos.m:1-44,
CalAllEquilent20210223_lobato_peng.m:23-57,213-226, CalAllEquilent20210305_peng_corr.m:25-70).

CPU: the oracle's Philox4x32-10 against the Random123 known-answer vectors; the normals' moments;
accepted positions obey the slab and mid-slice faces; slicing at and next to faces; correction rows
near faces shared half and half; the one-slice rows the device specimen starts from sit at their atoms.
GPU: displaced positions equal the oracle's (same generator, <= 1e-12 Angstrom); the transmission
built from the device-sliced specimen is bit-identical to totem_build_transmission on the rows the
host slices from the same positions, and within 1e-6 relative L2 of the FP64 oracle potential; with
multiceng > 0 (erf-weighted rows, shared corrections) within 1e-6 of the oracle and of the host-sliced
build.
"""
import numpy as np
import pytest

from oracle import phonon as oph
from totem_amd import specimen, stem


def test_philox_known_answer_vectors():
    # Random123 kat_vectors, philox4x32_10: zero, all-ones and pi-digit counters and keys
    assert oph.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    f = 0xFFFFFFFF
    assert oph.philox4x32_10((f, f, f, f), (f, f)) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert oph.philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_normals_have_unit_variance_and_are_independent():
    g = np.array([oph.normals3(12345, a, 3, 0) for a in range(20000)])
    assert np.all(np.abs(g.mean(axis=0)) < 0.03)
    assert np.all(np.abs(g.var(axis=0) - 1.0) < 0.04)
    c = np.corrcoef(g.T)
    assert np.all(np.abs(c - np.eye(3)) < 0.03)
    k = ((g - g.mean(0)) ** 4).mean(0) / g.var(0) ** 2
    assert np.all(np.abs(k - 3.0) < 0.15)
    # other configurations and seeds give other draws
    assert oph.normals3(12345, 0, 3, 0) != oph.normals3(12345, 0, 4, 0) != oph.normals3(12346, 0, 3, 0)


def test_displacement_obeys_slab_and_faces():
    rng = np.random.default_rng(0)
    S, dz, origin = 6, 1.5, -0.4
    lo, hi, faces = oph.slab(origin, dz, S, mid=(2, 4))
    xyz = np.column_stack([rng.uniform(0, 10, 400), rng.uniform(0, 10, 400), rng.uniform(lo, hi, 400)])
    rms = np.full(400, 0.4)
    rms[::10] = 0.0
    out = oph.displace(xyz, rms, 7, 1, lo, hi, faces)
    np.testing.assert_array_equal(out[::10], xyz[::10])
    assert np.all((out[:, 2] >= lo) & (out[:, 2] < hi))
    for f in faces:
        np.testing.assert_array_equal(out[:, 2] < f, xyz[:, 2] < f)
    moved = out - xyz
    assert np.count_nonzero(np.all(moved == 0, axis=1)) == 40     # only the rms = 0 atoms stay
    assert 0.3 < moved[rms > 0].std() < 0.45
    # an atom that no attempt can keep inside keeps its position
    far = oph.displace(np.array([[1.0, 2.0, 0.0]]), np.array([1e6]), 1, 0, lo, hi)
    np.testing.assert_array_equal(far, [[1.0, 2.0, 0.0]])


def test_slicing_at_faces_goes_up():
    origin, dz, S = -0.25, 0.5, 4
    st = origin + dz * np.arange(S)
    z = np.concatenate([st, np.nextafter(st, -np.inf), [st[-1] + dz, origin - 1e-9]])
    mem = oph.slice_members(z, origin, dz, S)
    for i in range(S):
        assert i in mem[i]                 # exactly on st[i]: slice i
        assert (S + i) not in mem[i]       # just below st[i]: slice i - 1
        if i > 0:
            assert (S + i) in mem[i - 1]
    assert all(2 * S not in m and 2 * S + 1 not in m for m in mem)   # above the top, below the bottom
    # the host slicer agrees with brute force
    atoms = np.zeros((len(z), 8))
    atoms[:, 0] = 8
    atoms[:, 3] = z
    atoms[:, 7] = 1.0
    all_n, all_n_i = specimen.atom_parameters(atoms, "l")
    _, _, series, _ = specimen.slice_atoms(all_n, all_n_i, dz, st, 0, "l", 0.0)
    np.testing.assert_array_equal(series, [len(m) for m in mem])


def test_correction_rows_near_faces_are_shared_half_and_half():
    dz, S = 1.5, 4
    st = -0.2 + dz * np.arange(S)
    z = [st[2] + 0.005, st[1] + dz - 0.005, st[0] + 0.005, st[3] + dz - 0.005, st[2] + 0.7]
    x = np.zeros((len(z), 8))
    x[:, 0], x[:, 1], x[:, 2], x[:, 3], x[:, 7] = 8, np.arange(len(z)), 0.5, z, 0.8
    cn = specimen.corr_parameters(x)
    e, _, series = specimen.slice_corr(cn, dz, st, 1)
    # [x y B occ*w cls]: atoms 0 and 1 split over slices (1, 2), the outer faces and atom 4 do not
    got = sorted((int(r[0]), float(r[3])) for r in e)
    assert got == [(0, 0.4), (0, 0.4), (1, 0.4), (1, 0.4), (2, 0.8), (3, 0.8), (4, 0.8)]
    np.testing.assert_array_equal(series, [1, 2, 3, 1])
    e0, _, series0 = specimen.slice_corr(cn, dz, st, 0)     # multiceng = 0: no sharing
    assert sorted(float(r[3]) for r in e0) == [0.8] * 5


def _ionic_perovskite(S, nx=2, seed=3):
    a = 3.905
    x = specimen.perovskite(nx, nx, S // 2, a=a)
    rng = np.random.default_rng(seed)
    x[:, 1:3] += rng.normal(0, 0.05, (len(x), 2))
    o = np.nonzero(x[:, 0] == 8)[0][::3]
    x[o, 4], x[o, 5] = -2, 0.5                      # mixed atomic/ionic oxygen
    ti = np.nonzero(x[:, 0] == 22)[0][::2]
    x[ti, 4], x[ti, 5] = 4, 1.0                     # fully ionic titanium
    x[::5, 7] = 0.9
    return x


@pytest.mark.parametrize("flag", ["l", "n"])
def test_one_slice_rows_sit_at_their_atoms(flag):
    x = _ionic_perovskite(4)
    cfg = stem.StemConfig(atoms=x, np_side=64, dx=0.12, n=32, kv=200.0, alpha_mrad=20.0, slice_thickness=1.95,
                          n_slices=4, slice_origin=-0.5, paraflag=flag)
    rows, ta, cim = stem.specimen_rows(cfg)
    n0, ni = len(rows["ele_n"]), len(rows["ele_n_i"])
    assert rows["series_n"].tolist() == [n0] and rows["series_n_i"].tolist() == [ni]
    nc = len(rows["ele_n_corr"]) if flag == "n" else 0
    assert len(ta) == n0 + ni + nc
    assert ni > 0 and np.count_nonzero(x[:, 5] == 1) > 0
    np.testing.assert_array_equal(rows["ele_n"][:, :2], x[ta[:n0], 1:3])
    np.testing.assert_array_equal(rows["ele_n_i"][:, :2], x[ta[n0:n0 + ni], 1:3])
    if flag == "n":
        assert nc > 0 and cim is not None
        np.testing.assert_array_equal(rows["ele_n_corr"][:, :2], x[ta[n0 + ni:], 1:3])
    # every atom contributes its rows once: the multi-slice host rows are a regrouping of these
    inp = stem.prepare(cfg, optics_too=False)
    assert sorted(map(tuple, inp.ele_n)) == sorted(map(tuple, rows["ele_n"]))


# --------------------------------------------------------------------------- GPU

def _case(flag, S=6, np_side=128, nx=2, rms_scale=1.0):
    a = 3.905
    x = _ionic_perovskite(S, nx)
    dz = a / 2
    cfg = stem.StemConfig(atoms=x, np_side=np_side, dx=nx * a / np_side, n=64, kv=300.0, alpha_mrad=24.0,
                          slice_thickness=dz, n_slices=S, slice_origin=-dz / 4, paraflag=flag, mid_slices=(2, 4))
    sig = {38: 0.25 * rms_scale, 22: 0.2 * rms_scale, 8: 0.3 * rms_scale}
    return cfg, sig


def _host_transmission(engine, cfg, atoms):
    import torch
    inp = stem.prepare(cfg, atoms=atoms, optics_too=False)
    S, npx = cfg.n_slices, cfg.np_side
    t = torch.empty((S, npx, npx), dtype=torch.complex64, device="cuda")
    corr = cfg.paraflag == "n"
    engine.build_transmission(
        cfg.paraflag, npx, S, inp.series_n, inp.ele_n, inp.gx, inp.gy, inp.s2, inp.aper_true, inp.sigma,
        inp.parameter, t.data_ptr(), atom_slice_i=inp.series_n_i if len(inp.series_n_i) else None,
        proj_coff_mat_i=inp.ele_n_i if len(inp.ele_n_i) else None,
        series_n_corr=inp.series_n_corr if corr else None, ele_n_corr=inp.ele_n_corr if corr else None,
        corr_info_matrix=inp.corr_matrix if corr else None, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return t.cpu().numpy(), inp


@pytest.mark.gpu
def test_device_displacement_matches_oracle(engine):
    print("case specimen positions", flush=True)
    cfg, sig = _case("l", rms_scale=1.5)
    x = cfg.atoms
    rms = np.array([sig[int(z)] for z in x[:, 0]])
    lo, hi, faces = oph.slab(cfg.slice_origin, cfg.slice_thickness, cfg.n_slices, cfg.mid_slices)

    def first_draw_rejected(seed, config, a):
        z = x[a, 3] + rms[a] * oph.normals3(seed, a, config, 0)[2]
        return not (lo <= z < hi and all((x[a, 3] < f) == (z < f) for f in faces))

    redrawn = 0
    with stem.device_specimen(cfg, sigma_by_z=sig) as sp:
        for seed, config in [(1, 0), (1, 1), (2 ** 40 + 5, 7)]:
            p = sp.positions(seed, config)
            ref = oph.displace(x[:, 1:4], rms, seed, config, lo, hi, faces)
            err = np.abs(p - ref).max()
            r = sum(first_draw_rejected(seed, config, a) for a in range(len(x)))
            redrawn += r
            print("  seed %d config %d: max |device - oracle| %.2e A, %d atoms redrawn" % (seed, config, err, r))
            assert err <= 1e-12
            np.testing.assert_array_equal(p, sp.positions(seed, config))
    assert redrawn > 0


@pytest.mark.gpu
@pytest.mark.parametrize("flag", ["l", "n"])
def test_device_sliced_transmission_is_host_sliced_transmission(engine, flag):
    import torch
    from oracle import potential as opot
    print("case specimen transmission flag=%s" % flag, flush=True)
    cfg, sig = _case(flag)
    S, npx = cfg.n_slices, cfg.np_side
    t = torch.empty((S, npx, npx), dtype=torch.complex64, device="cuda")
    with stem.device_specimen(cfg, sigma_by_z=sig) as sp:
        for config in range(3):
            sp.transmission(99, config, t.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            td = t.cpu().numpy()
            xd = cfg.atoms.copy()
            xd[:, 1:4] = sp.positions(99, config)
            xd[:, 6] = 0.0
            th, inp = _host_transmission(engine, cfg, xd)
            same = np.array_equal(td, th)
            kw = {}
            if flag == "n":
                kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
            _, to = opot.potential_slices(npx, inp.ele_n, None, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                          inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, flag, **kw)
            err = max(float(np.linalg.norm(td[s] - to[s]) / np.linalg.norm(to[s])) for s in range(S))
            print("  config %d: bit-identical to host slicing: %s; max t rel L2 vs oracle %.3e; atoms per slice %s"
                  % (config, same, err, inp.series_n.astype(int).tolist()))
            assert same
            assert err <= 1e-6


@pytest.mark.gpu
def test_frozen_phonon_configuration_at_full_size(engine):
    """One configuration of the 80 000-atom SrTiO3 specimen (80 slices of 1024^2, Peng + correction):
    device displacement + slicing + transmission against host slicing of the same positions."""
    import time

    import torch
    print("case specimen 80000 atoms 80 x 1024^2", flush=True)
    cfg = stem.perovskite_c2()
    S, npx = cfg.n_slices, cfg.np_side
    t = torch.empty((S, npx, npx), dtype=torch.complex64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    sp = stem.device_specimen(cfg, sigma_by_z=stem.C2_PHONON_SIGMA)
    t1 = time.perf_counter()
    try:
        sp.transmission(5, 0, t.data_ptr(), stream=st)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for c in range(1, 4):
            sp.transmission(5, c, t.data_ptr(), stream=st)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        td = t.cpu().numpy()
        xd = cfg.atoms.copy()
        xd[:, 1:4] = sp.positions(5, 3)
        xd[:, 6] = 0.0
    finally:
        sp.close()
    t4 = time.perf_counter()
    stem.prepare(cfg, atoms=xd, optics_too=False)
    t5 = time.perf_counter()
    th, inp = _host_transmission(engine, cfg, xd)
    t6 = time.perf_counter()
    print("  resident specimen setup %.2f s; device configuration (displace + slice + build) %.1f ms (first %.1f ms);"
          " host slicing %.2f s + build from host rows %.1f ms" % (t1 - t0, (t3 - t2) / 3 * 1e3, (t2 - t1) * 1e3,
                                                                  t5 - t4, (t6 - t5 - (t5 - t4)) * 1e3))
    assert np.array_equal(td, th)


def _near_face_case(flag, multiceng):
    """A perovskite whose atoms partly sit within 0.01 A of slice faces (corrections shared) and within
    multiceng * dz of them (erf-weighted rows); mixed and fully ionic atoms included."""
    cfg, sig = _case(flag)
    x = cfg.atoms.copy()
    st = stem.slice_boundaries(cfg)
    dz = cfg.slice_thickness
    rng = np.random.default_rng(5)
    pick = rng.choice(len(x), 12, replace=False)
    for j, a in enumerate(pick):
        i = 1 + j % (cfg.n_slices - 2)
        x[a, 3] = st[i] + 0.004 if j % 2 == 0 else st[i] + dz - 0.004
    from dataclasses import replace
    return replace(cfg, atoms=x, multiceng=multiceng), sig


@pytest.mark.gpu
@pytest.mark.parametrize("flag,multiceng", [("p", 1.0), ("n", 1.0), ("p", 0.5), ("n", 0.5)])
def test_device_erf_weighted_slicing_matches_oracle(engine, flag, multiceng):
    import torch
    from oracle import potential as opot
    print("case specimen slicing flag=%s multiceng=%g" % (flag, multiceng), flush=True)
    cfg, sig = _near_face_case(flag, multiceng)
    S, npx = cfg.n_slices, cfg.np_side
    t = torch.empty((S, npx, npx), dtype=torch.complex64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for label, sigma_by_z in (("static", None), ("displaced", sig)):
        with stem.device_specimen(cfg, sigma_by_z=sigma_by_z) as sp:
            sp.transmission(42, 1, t.data_ptr(), stream=st)
            torch.cuda.synchronize()
            td = t.cpu().numpy().astype(np.complex128)
            xd = cfg.atoms.copy()
            xd[:, 1:4] = sp.positions(42, 1)
        if sigma_by_z is not None:
            xd[:, 6] = 0.0
        th, inp = _host_transmission(engine, cfg, xd)
        kw = {}
        if flag == "n":
            kw = dict(ele_n_corr=inp.ele_n_corr, corr_matrix=inp.corr_matrix, series_n_corr=inp.series_n_corr)
        _, to = opot.potential_slices(npx, inp.ele_n, None, inp.ele_n_i, None, inp.series_n, inp.series_n_i,
                                      inp.s2, inp.gx, inp.gy, inp.sigma, inp.parameter, inp.aper_true, flag, **kw)
        e_o = max(float(np.linalg.norm(td[s] - to[s]) / np.linalg.norm(to[s])) for s in range(S))
        e_h = max(float(np.linalg.norm(td[s] - th[s]) / np.linalg.norm(th[s])) for s in range(S))
        n_rows = len(inp.ele_n) + len(inp.ele_n_i)
        print("  %s: host rows %d (atoms %d), corr rows %s; max t rel L2 vs oracle %.3e, vs host-sliced build %.3e"
              % (label, n_rows, len(xd), len(inp.ele_n_corr) if flag == "n" else "-", e_o, e_h))
        assert n_rows > len(xd)
        assert e_o <= 1e-6 and e_h <= 1e-6


def _small_desc_args(flag="n"):
    x = _ionic_perovskite(4)
    cfg = stem.StemConfig(atoms=x, np_side=64, dx=0.12, n=32, kv=200.0, alpha_mrad=20.0, slice_thickness=1.95,
                          n_slices=4, slice_origin=-0.5, paraflag=flag, mid_slices=(2,))
    rows, ta, cim = stem.specimen_rows(cfg)
    from totem_amd import optics
    gx, gy, s2 = optics.reciprocal_grids(cfg.np_side, cfg.dx)
    return dict(paraflag=flag, np_side=64, rows=rows, term_atom=ta, xyz=x[:, 1:4], rms=np.full(len(x), 0.1), kx=gx,
                ky=gy, s_2=s2, aper_true=np.ones_like(s2), sigma=0.7, parameter=1.0, slices=4, slice_origin=-0.5,
                slice_thickness=1.95, mid=(2,), corr_info_matrix=cim)


@pytest.mark.parametrize("change,match", [
    (dict(slices=0), "slices must be positive"),
    (dict(slice_thickness=0.0), "slice_thickness must be positive"),
    (dict(multiceng=-1.0), "multiceng must be finite"),
    (dict(mid=(5,)), "mid slice out of range"),
    (dict(rms="negative"), "rms must be finite"),
    (dict(term_atom="short"), "n_terms"),
    (dict(term_atom="swapped"), "does not sit at its atom"),
])
def test_device_specimen_rejects_bad_descriptions(change, match):
    """Argument checks run before any device call, so they are testable without a GPU."""
    from totem_amd import engine
    kw = _small_desc_args()
    for k, v in change.items():
        if v == "negative":
            kw["rms"] = kw["rms"].copy()
            kw["rms"][3] = -0.1
        elif v == "short":
            kw["term_atom"] = kw["term_atom"][:-1]
        elif v == "swapped":
            ta = kw["term_atom"].copy()
            ta[0], ta[1] = ta[1], ta[0]
            kw["term_atom"] = ta
        else:
            kw[k] = v
    with pytest.raises(engine.TotemError, match=match):
        engine.DeviceSpecimen(**kw)
