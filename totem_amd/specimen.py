"""Specimen preparation on the host (FP64): atoms -> per-slice coefficient rows.

Restates the reference's host-side preparation that feeds the engine:

* atom matrix ``x = [Z x y z valence ionicity B occ]`` (``ToTEM_submit.m:107-112``);
* coefficient assembly ``getpotentialpara_lotabo_peng.m:20-86`` and
  ``getpotentialpara_corr_peng.m:16-24`` plus the Debye-Waller handling of
  ``CommonPart_forsimulation_lobato_peng.m:88-121``;
* slice assignment with erf weights ``CalAllEquilent20210223_lobato_peng.m:23-93,213-226``
  and the correction-term slicing ``CalAllEquilent20210305_peng_corr.m:25-70``;
* the correction table Lobato(s) - Peng(s) per class ``ToTEM_submit.m:1597-1610``;
* frozen-phonon displacement with redraw ``randshiftpos.m:1-44`` (seeded here;
  the reference seeds with ``rng('shuffle')``).

Also builds the benchmark specimens (perovskite, diamond-cubic [110], amorphous
carbon) used by the engine's workloads.
"""
import numpy as np
from scipy.special import erf

from . import tables

# ---------------------------------------------------------------------------- specimens


def perovskite(nx, ny, nz, a=3.905, zA=38, zB=22, zO=8, B=(0.0, 0.0, 0.0), origin=(0.0, 0.0, 0.0)):
    """Cubic ABO3 supercell, A at (0,0,0), B at (1/2,1/2,1/2), O at the face centres.

    Returns the reference's atom matrix [n, 8] = [Z x y z valence ionicity B occ];
    B = (B_A, B_B, B_O) Debye-Waller factors in Angstrom^2.
    """
    basis = [
        (zA, 0.0, 0.0, 0.0, B[0]),
        (zB, 0.5, 0.5, 0.5, B[1]),
        (zO, 0.5, 0.5, 0.0, B[2]),
        (zO, 0.5, 0.0, 0.5, B[2]),
        (zO, 0.0, 0.5, 0.5, B[2]),
    ]
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    rows = []
    for z, fx, fy, fz, b in basis:
        pos = (cells + np.array([fx, fy, fz])) * a + np.asarray(origin)
        blk = np.zeros((len(cells), 8))
        blk[:, 0] = z
        blk[:, 1:4] = pos
        blk[:, 6] = b
        blk[:, 7] = 1.0
        rows.append(blk)
    x = np.concatenate(rows)
    order = np.lexsort((x[:, 1], x[:, 2], x[:, 3]))
    return x[order]


def diamond_110(lx, ly, lz, a=5.431, Z=14, B=0.0, origin=(0.0, 0.0, 0.0)):
    """Diamond-cubic crystal viewed along [110]: beam (z) || [110], x || [-110], y || [001].

    Orthorhombic repeat a/sqrt2 x a x a/sqrt2 with 4 atoms; tiles a box of at
    least lx x ly x lz Angstrom (rounded up to whole repeats).
    """
    ax, ay, az = a / np.sqrt(2), a, a / np.sqrt(2)
    # diamond basis in cubic fractional coords
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    dia = np.concatenate([fcc, fcc + 0.25]) % 1.0
    # rotate: new x = (-1,1,0)/sqrt2, y = (0,0,1), z = (1,1,0)/sqrt2, in units of a
    cart = dia * a
    ex = np.array([-1, 1, 0]) / np.sqrt(2)
    ey = np.array([0, 0, 1.0])
    ez = np.array([1, 1, 0]) / np.sqrt(2)
    # expand the cubic cell so the rotated orthorhombic cell is covered, then fold
    sh = np.array([[i, j, k] for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)])
    pts = (cart[None, :, :] + sh[:, None, :] * a).reshape(-1, 3)
    r = np.stack([pts @ ex, pts @ ey, pts @ ez], axis=1)
    per = np.array([ax, ay, az])
    r = np.round(r % per, 9)
    r[np.abs(r - per) < 1e-6] = 0.0  # a coordinate that rounds onto the period is the same site as 0
    r = np.unique(r, axis=0)
    assert len(r) == 4, "diamond [110] repeat must hold 4 sites"
    nx, ny, nz = (int(np.ceil(l / p - 1e-9)) for l, p in ((lx, ax), (ly, ay), (lz, az)))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) * np.array([ax, ay, az])
    pos = (cells[:, None, :] + r[None, :, :]).reshape(-1, 3) + np.asarray(origin)
    x = np.zeros((len(pos), 8))
    x[:, 0] = Z
    x[:, 1:4] = pos
    x[:, 6] = B
    x[:, 7] = 1.0
    return x


def amorphous(lx, ly, z0, z1, density_per_A3=0.1003, dmin=1.4, Z=6, seed=0, B=0.0, batch=8192):
    """Uniform random atoms in [0,lx)x[0,ly)x[z0,z1) at the given number density with a minimum
    separation dmin (periodic in x and y), by random sequential addition: candidates are drawn in
    batches, dropped if closer than dmin to an accepted atom, and accepted in draw order within a
    batch (a candidate loses to an earlier accepted one).  Seeded (numpy PCG64)."""
    from scipy.spatial import cKDTree
    rng = np.random.Generator(np.random.PCG64(seed))
    lz = z1 - z0
    target = int(round(density_per_A3 * lx * ly * lz))
    box = np.array([lx, ly, 2.0 * lz + 4.0 * dmin])  # z padded: no wrap-around within dmin
    scale = np.array([lx, ly, lz])
    pts = np.zeros((0, 3))
    rounds = 0
    while len(pts) < target:
        rounds += 1
        if rounds > 10000:
            raise RuntimeError("amorphous: density too high for dmin")
        cand = rng.random((batch, 3)) * scale
        if len(pts):
            d, _ = cKDTree(pts, boxsize=box).query(cand, k=1, distance_upper_bound=dmin)
            cand = cand[~(d < dmin)]
        if len(cand) == 0:
            continue
        pairs = cKDTree(cand, boxsize=box).query_pairs(dmin, output_type="ndarray")
        keep = np.ones(len(cand), dtype=bool)
        if len(pairs):
            pairs = np.sort(pairs, axis=1)
            pairs = pairs[np.lexsort((pairs[:, 0], pairs[:, 1]))]
            for i, j in pairs:
                if keep[i]:
                    keep[j] = False
        pts = np.concatenate([pts, cand[keep][:target - len(pts)]])
    pos = pts + np.array([0.0, 0.0, z0])
    x = np.zeros((len(pos), 8))
    x[:, 0] = Z
    x[:, 1:4] = pos
    x[:, 6] = B
    x[:, 7] = 1.0
    return x


# ------------------------------------------------------------------ coefficient assembly


def atom_parameters(x, flag):
    """``getpotentialpara_lotabo_peng.m:1-87``.

    Returns (all_nuclear [n, 16], all_nuc_ion [m, 28]):
      all_nuclear = [Z x y z B occ*(1-ion) a1 b1 ... a5 b5]       (ionicity 0 or 1)
      all_nuc_ion = [Z x y z B occ ion  ionpar(10)  1-ion  atompar(10)]  (0 < ionicity < 1)
    flag: 'p'/'n' Peng tables, 'l' Lobato tables (ions always from Peng's ion table).
    """
    x = np.asarray(x, dtype=np.float64)
    neutral = tables.lobato_table() if flag == "l" else tables.peng_table()
    ion_tab = tables.peng_ion_table()

    def atom_par(zs):
        return tables.interleaved(neutral[zs.astype(int) - 1])

    def ion_par(z, val):
        k = np.nonzero((ion_tab[:, 0] == z) & (ion_tab[:, 1] == val))[0]
        if len(k) == 0:
            raise ValueError("no ion parameters for Z=%g valence=%g" % (z, val))
        return tables.interleaved(ion_tab[k[0], 2:])

    rows = []
    b = np.nonzero(x[:, 5] == 0)[0]
    if len(b):
        blk = np.zeros((len(b), 16))
        blk[:, 0] = x[b, 0]
        blk[:, 1:4] = x[b, 1:4]
        blk[:, 4] = x[b, 6]
        blk[:, 5] = x[b, 7] * (1 - x[b, 5])
        blk[:, 6:16] = atom_par(x[b, 0])
        rows.append(blk)
    b = np.nonzero(x[:, 5] == 1)[0]
    if len(b):
        blk = np.zeros((len(b), 16))
        blk[:, 0] = x[b, 0]
        blk[:, 1:4] = x[b, 1:4]
        blk[:, 4] = x[b, 6]
        blk[:, 5] = x[b, 7] * x[b, 5]
        blk[:, 6:16] = np.stack([ion_par(x[i, 0], x[i, 4]) for i in b])
        rows.append(blk)
    all_nuclear = np.concatenate(rows) if rows else np.zeros((0, 16))
    b = np.nonzero((x[:, 5] > 0) & (x[:, 5] < 1))[0]
    all_nuc_ion = np.zeros((len(b), 28))
    if len(b):
        all_nuc_ion[:, 0] = x[b, 0]
        all_nuc_ion[:, 1:4] = x[b, 1:4]
        all_nuc_ion[:, 4] = x[b, 6]
        all_nuc_ion[:, 5] = x[b, 7]
        all_nuc_ion[:, 6] = x[b, 5]
        all_nuc_ion[:, 7:17] = np.stack([ion_par(x[i, 0], x[i, 4]) for i in b])
        all_nuc_ion[:, 17] = 1 - x[b, 5]
        all_nuc_ion[:, 18:28] = atom_par(x[b, 0])
    return all_nuclear, all_nuc_ion


def fold_debye_waller(all_nuclear, all_nuc_ion, flag, dw_selected):
    """Peng only: add B to every b_i when DW or absorption is selected
    (``CommonPart_forsimulation_lobato_peng.m:88-116``).  Lobato keeps B for the
    per-atom exp(-B s^2) factor instead."""
    all_nuclear = all_nuclear.copy()
    all_nuc_ion = all_nuc_ion.copy()
    if flag in ("p", "n") and dw_selected:
        if len(all_nuclear):
            all_nuclear[:, 7:16:2] += all_nuclear[:, 4:5]
        if len(all_nuc_ion):
            all_nuc_ion[:, 8:17:2] += all_nuc_ion[:, 4:5]
            all_nuc_ion[:, 19:28:2] += all_nuc_ion[:, 4:5]
    return all_nuclear, all_nuc_ion


def corr_parameters(x, dw_selected=False):
    """``getpotentialpara_corr_peng.m:15-24`` (neutral atoms) with the DW column
    zeroed when neither DW nor absorption is selected
    (``CommonPart_forsimulation_lobato_peng.m:26-35``).

    Returns corr_nuclear [n, 26] = [Z x y z B occ*(1-ion) peng(a1 b1..a5 b5) lobato(a1 b1..a5 b5)].
    """
    x = np.asarray(x, dtype=np.float64)
    b = np.nonzero(x[:, 5] == 0)[0]
    out = np.zeros((len(b), 26))
    if len(b):
        z = x[b, 0].astype(int)
        out[:, 0] = x[b, 0]
        out[:, 1:4] = x[b, 1:4]
        out[:, 4] = x[b, 6]
        out[:, 5] = x[b, 7] * (1 - x[b, 5])
        out[:, 6:16] = tables.interleaved(tables.peng_table()[z - 1])
        out[:, 16:26] = tables.interleaved(tables.lobato_table()[z - 1])
        if not dw_selected:
            out[:, 4] = 0.0
    return out


# ------------------------------------------------------------------------------ slicing


def _contribution(b, z1, z2):
    """``project_poten_contribute`` (``CalAllEquilent20210223_lobato_peng.m:213-226``)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        sb = np.sqrt(b)
        t2 = 2 * np.pi * z2 / sb
        t1 = 2 * np.pi * z1 / sb
    return np.abs(0.5 * erf(t2) - 0.5 * erf(t1))


def slice_atoms(all_n, all_n_i, eachthick, slicethick, multiceng, flag, dbmode):
    """``CalAllEquilent20210223_lobato_peng.m:1-93,122-203``.

    Returns (ele_n, ele_n_i, series_n, series_n_i):
      ele_n   [sum n, 13] = [x y occ a1 b1 ... a5 b5]          (Peng)
              [sum n, 14] = [x y occ a1 b1 ... a5 b5 B*dbmode] (Lobato)
      ele_n_i [sum n_i, 25 | 26] = [x y occ ion ionpar(10) 1-ion atompar(10) (B*dbmode)]
    The a_i of atoms near a slice face are weighted by erf contributions when
    multiceng > 0 (the reference forces multiceng = 0 for Lobato).
    """
    slicethick = np.asarray(slicethick, dtype=np.float64)
    S = len(slicethick)

    def run(all_x, zc, acols, bcols):
        blocks, counts = [], []
        for i in range(S):
            lo = slicethick[i] - multiceng * eachthick
            hi = slicethick[i] + multiceng * eachthick + eachthick
            jj = np.nonzero((all_x[:, 3] >= lo) & (all_x[:, 3] < hi))[0]
            counts.append(len(jj))
            blk = all_x[jj].copy()
            z1 = blk[:, 3] - (slicethick[i] + eachthick)
            z2 = blk[:, 3] - slicethick[i]
            if i == 0:
                z2[:] = np.inf
            if i == S - 1:
                z1[:] = -np.inf
            z2[np.abs(z2) >= multiceng * eachthick] = np.inf
            z1[np.abs(z1) >= multiceng * eachthick] = -np.inf
            for ac, bc in zip(acols, bcols):
                blk[:, ac] = blk[:, ac] * _contribution(blk[:, bc], z1, z2)
            blocks.append(blk)
        arr = np.concatenate(blocks) if blocks else np.zeros((0, all_x.shape[1]))
        return arr, np.asarray(counts, dtype=np.float64)

    ele_n = np.zeros((0, 14 if flag == "l" else 13))
    series_n = np.zeros(0)
    if len(all_n):
        arr, series_n = run(all_n, 3, [6, 8, 10, 12, 14], [7, 9, 11, 13, 15])
        temp = arr[:, 4].copy()
        ele_n = np.delete(arr, [0, 3, 4], axis=1)
        if flag == "l":
            ele_n = np.concatenate([ele_n, (temp * dbmode)[:, None]], axis=1)
    ele_n_i = np.zeros((0, 26 if flag == "l" else 25))
    series_n_i = np.zeros(0)
    if len(all_n_i):
        acols = [7, 9, 11, 13, 15, 18, 20, 22, 24, 26]
        bcols = [8, 10, 12, 14, 16, 19, 21, 23, 25, 27]
        arr, series_n_i = run(all_n_i, 3, acols, bcols)
        temp = arr[:, 4].copy()
        ele_n_i = np.delete(arr, [0, 3, 4], axis=1)
        if flag == "l":
            ele_n_i = np.concatenate([ele_n_i, (temp * dbmode)[:, None]], axis=1)
    return ele_n, ele_n_i, series_n, series_n_i


def _unique_rows_sorted(a):
    """MATLAB ``unique(a,'rows')``: sorted unique rows and 1-based class index per row."""
    uniq, inv = np.unique(a, axis=0, return_inverse=True)
    return uniq, inv.reshape(-1) + 1


def slice_corr(corr_nuclear, eachthick, slicethick, multiceng, discon=0.015):
    """``CalAllEquilent20210305_peng_corr.m:25-70``.

    Returns (ele_n_corr [n, 5] = [x y B occ*w cls], corr_info [k, 20], series_n_corr [S]).
    """
    slicethick = np.asarray(slicethick, dtype=np.float64)
    S = len(slicethick)
    if len(corr_nuclear) == 0:
        return np.zeros((0, 5)), np.zeros((0, 20)), np.zeros(S)
    corr_info, cls = _unique_rows_sorted(corr_nuclear[:, 6:26])
    c7 = np.concatenate([corr_nuclear[:, :6], cls[:, None].astype(np.float64)], axis=1)
    rows = []
    for i in range(S):
        jj = np.nonzero((c7[:, 3] >= slicethick[i]) & (c7[:, 3] < slicethick[i] + eachthick))[0]
        n = len(jj)
        blk = np.zeros((n, 11))
        blk[:, :7] = c7[jj]
        blk[:, 7] = 1.0
        blk[:, 8] = i + 1
        blk[:, 9] = c7[jj, 3] - slicethick[i]
        blk[:, 10] = c7[jj, 3] - slicethick[i] - eachthick
        rows.append(blk)
    e = np.concatenate(rows)
    if eachthick > 2 * discon and multiceng >= 1:
        jj = np.nonzero((e[:, 9] < discon) & (e[:, 8] > 1))[0]
        if len(jj):
            e[jj, 7] = 0.5
            cp = e[jj].copy()
            cp[:, 8] -= 1
            e = np.concatenate([e, cp])
        jjjj = np.nonzero((e[:, 10] > -discon) & (e[:, 8] < S))[0]
        if len(jjjj):
            e[jjjj, 7] = 0.5
            cp = e[jjjj].copy()
            cp[:, 8] += 1
            e = np.concatenate([e, cp])
    e = e[:, :9]
    order = np.argsort(e[:, 8], kind="stable")
    e = e[order]
    series = np.asarray([np.count_nonzero(e[:, 8] == i + 1) for i in range(S)], dtype=np.float64)
    e[:, 5] = e[:, 5] * e[:, 7]
    e = np.delete(e, [8, 7, 3, 0], axis=1)
    return e, corr_info, series


def corr_info_matrix(corr_info, s2):
    """Per-class correction Lobato(s) - Peng(s) on the s^2 grid, [k, Np, Np]
    (``ToTEM_submit.m:1597-1610``; corr_info columns 1..10 Peng a1 b1.., 11..20 Lobato a1 b1..)."""
    s2 = np.asarray(s2, dtype=np.float64)
    g2 = 4 * s2
    out = np.zeros((len(corr_info),) + s2.shape)
    for i, ci in enumerate(corr_info):
        lob = sum(ci[10 + 2 * q] * (2 + ci[11 + 2 * q] * g2) / (1 + ci[11 + 2 * q] * g2) ** 2 for q in range(5))
        pen = sum(ci[2 * q] * np.exp(-s2 * ci[2 * q + 1]) for q in range(5))
        out[i] = lob - pen
    return out


# ------------------------------------------------------------------------ frozen phonon


def frozen_phonon(pos, eachthick, slicethick, mid_slice_num=(), rng=None, bcol=4):
    """``randshiftpos.m:1-44``: Gaussian displacement with sigma = sqrt(B/(8 pi^2)) per axis,
    redrawn until every atom stays inside [slicethick[0], slicethick[-1]+eachthick) and
    atoms do not cross the listed mid-slice faces.  ``pos`` columns 1..3 are x, y, z and
    ``bcol`` is the Debye-Waller column (4 in the all_nuclear layout)."""
    rng = rng if rng is not None else np.random.Generator(np.random.PCG64(0))
    slicethick = np.asarray(slicethick, dtype=np.float64)
    out = pos.copy()
    sig = np.sqrt(pos[:, bcol] / 8 / np.pi / np.pi)
    out[:, 1:4] = pos[:, 1:4] + rng.standard_normal((len(pos), 3)) * sig[:, None]
    lo, hi = slicethick[0], slicethick[-1] + eachthick
    for _ in range(10000):
        bad = np.nonzero((out[:, 3] < lo) | (out[:, 3] >= hi))[0]
        if len(bad) == 0:
            break
        out[bad, 1:4] = pos[bad, 1:4] + rng.standard_normal((len(bad), 3)) * sig[bad, None]
    for m in mid_slice_num:
        face = slicethick[m - 1] + eachthick
        k = pos[:, 3] < face
        for _ in range(10000):
            kk = out[:, 3] < face
            diff = np.nonzero(k != kk)[0]
            if len(diff) == 0:
                break
            out[diff, 1:4] = pos[diff, 1:4] + rng.standard_normal((len(diff), 3)) * sig[diff, None]
    return out
