"""STEM experiment set-up on the host (FP64): the inputs the multislice engine consumes.

This restates the reference's GUI-side set-up so that callers of the engine can
build the same input arrays without MATLAB:

* constants and electron optics: ``ToTEM_submit.m:1479-1489`` (PARAMETER, sigma),
  ``ToTEM_submit.m:782-783`` (wavelength);
* reciprocal grids and s^2: ``ToTEM_submit.m:1503-1507``;
* Fresnel propagator with the 2/3 band limit: ``ToTEM_submit.m:1520-1533``;
* annular detector masks: ``ToTEM_submit.m:1538-1551``;
* probe: ``ToTEM_submit.m:788-806`` with ``CommonPara_TEM:2122-2155``,
  ``WholeTCC_residual_newTEM:2232-2268``, ``WholeTCC2D_newTEM:2274-2375``,
  aperture ``myaperture:715-736``;
* focal-spread series weights: ``Gaussian_focal/gaussian_focal/energyspread:598-636``.

Units follow the reference: wavelengths in nm inside the probe code, Angstrom
elsewhere; voltages in kV at the interface.
"""
import numpy as np

# Physical constants exactly as the reference writes them (ToTEM_submit.m:1479-1483);
# they are not CODATA values, and PARAMETER/sigma/lambda depend on them.
H_PLANCK = 6.6255916e-34
HBAR = H_PLANCK / 2.0 / np.pi
E_CHARGE = 1.602102e-19
M_ELECTRON = 9.1093807e-31
C_LIGHT = 2.9979251e8


def wavelength_nm(kv):
    """Relativistic electron wavelength in nm (``ToTEM_submit.m:782``)."""
    v = kv * 1000.0
    return 1.0e9 * H_PLANCK * C_LIGHT / np.sqrt(E_CHARGE * v * (2 * M_ELECTRON * C_LIGHT ** 2 + E_CHARGE * v))


def wavelength_A(kv):
    return 10.0 * wavelength_nm(kv)


def interaction_sigma(kv):
    """Interaction constant sigma in 1/(kV*Angstrom) (``ToTEM_submit.m:1489``).

    The engine applies exp(i * sigma/1000 * V) with V in V*Angstrom.
    """
    v = kv * 1000.0
    mc2 = M_ELECTRON * C_LIGHT ** 2
    return 2 * np.pi / (wavelength_nm(kv) * 10.0 * v / 1000.0) * (mc2 + E_CHARGE * v) / (2 * mc2 + E_CHARGE * v)


def potential_parameter(np_side, dx):
    """PARAMETER = 2 pi hbar^2 / (e m_e) / (Np^2 dx^2) in Angstrom units (``ToTEM_submit.m:1488``)."""
    return 2 * np.pi * HBAR * HBAR / (E_CHARGE * M_ELECTRON * 1.0e-10 * 1.0e-10) / (np_side * np_side * dx * dx)


def centred_freq(n, dx):
    """(-n/2 .. n/2-1)/(n dx): the centred reciprocal grid of an even n."""
    return np.arange(-n // 2, n // 2, dtype=np.float64) / (n * dx)


def reciprocal_grids(np_side, dx):
    """gx (varies along columns), gy (varies along rows), s2 = (gx^2+gy^2)/4, all [Np, Np].

    ``ToTEM_submit.m:1503-1507``: meshgrid of the centred frequencies, s = g/2.
    """
    f = centred_freq(np_side, dx)
    gx, gy = np.meshgrid(f, f)
    s2 = (gx / 2) ** 2 + (gy / 2) ** 2
    return gx, gy, s2


def propagator(n, dx, kv, dz, tilt_mrad=0.0, tilt_phi_deg=0.0, cut=9.0):
    """Fresnel propagator H(k) = exp(-i pi lambda dz |k+tilt|^2), zero where |k+tilt|^2 > 1/(cut dx^2),
    returned ifftshifted (zero frequency at [0, 0]), complex128 [n, n].

    STEM and CBED cut at 2/3 of the Nyquist radius (cut 9, ``ToTEM_submit.m:1520-1533,1371-1374``);
    the HRTEM plane-wave path at 1/2 (cut 16, ``:1311-1314``).
    """
    lam_nm = wavelength_nm(kv)
    tilt = tilt_mrad * 1e-3
    phi = -tilt_phi_deg
    tiltx = tilt / lam_nm * np.cos(phi / 180 * np.pi)
    tilty = tilt / lam_nm * np.sin(phi / 180 * np.pi)
    f = centred_freq(n, dx)
    px, py = np.meshgrid(f, f)
    p2 = (px + tiltx * 0.1) ** 2 + (py + tilty * 0.1) ** 2
    h = np.exp(-1j * np.pi * lam_nm * 10 * dz * p2)
    h[p2 > 1.0 / (cut * dx * dx)] = 0
    return np.fft.ifftshift(h)


def detector_masks(n, dx, kv, ranges_mrad):
    """Centred annular masks [A, n, n] (float64 0/1): min(r)^2 <= k^2 <= max(r)^2,
    r = mrad*1e-3/lambda[Angstrom] (``ToTEM_submit.m:1538-1551``)."""
    lam_A = wavelength_A(kv)
    f = centred_freq(n, dx)
    px, py = np.meshgrid(f, f)
    k2 = px ** 2 + py ** 2
    out = []
    for r in np.atleast_2d(np.asarray(ranges_mrad, dtype=np.float64)):
        rr = r * 1e-3 / lam_A
        out.append(((k2 >= rr.min() ** 2) & (k2 <= rr.max() ** 2)).astype(np.float64))
    return np.stack(out)


# ----------------------------------------------------------------------------- probe

def _mround(x):
    """MATLAB round (half away from zero) for non-negative x."""
    return int(np.floor(x + 0.5))


def _gaussian_focal_weights(mtotal, rms, delta_yita):
    """``gaussian_focal:608-617``."""
    m = (mtotal - 1) // 2
    if m == 0:
        return np.ones(1)
    a = np.arange(-m, m + 1) * delta_yita
    return delta_yita * np.exp(-a * a / (2 * rms * rms)) / (np.sqrt(2 * np.pi) * rms)


def _energy_spread_step(lam_nm, mtotal, rms, gmax):
    """``energyspread:620-634``: the focal step that best matches the Gaussian envelope."""
    m = (mtotal - 1) // 2
    pre = 0.1
    uu = np.arange(0, gmax + 1e-12, 0.01)
    fixvalue = np.exp(-0.5 * (np.pi * rms * lam_nm) ** 2 * uu ** 4)
    # MATLAB colon fieldmin*rms:pre:fieldmax*rms
    lo, hi = 0.5 * rms, 2.0 * rms
    count = int(np.floor((hi - lo) / pre + 1e-10)) + 1
    temp = lo + pre * np.arange(count)
    err = []
    ms = np.arange(-m, m + 1)[:, None]
    for d in temp:
        g = _gaussian_focal_weights(mtotal, rms, d)
        val = np.real(g @ np.exp(-1j * np.pi * lam_nm * ms * d * uu[None, :] ** 2))
        err.append(np.max(np.abs(fixvalue - val)))
    err = np.asarray(err)
    return temp[err == err.min()][0]


def focal_spread(mtotal, rms_nm, lam_nm, gmax):
    """(delta_yita, gfsf) from ``Gaussian_focal:598-605``; mtotal = 1 gives (0, [1])."""
    if (mtotal - 1) // 2 == 0:
        return 0.0, np.ones(1)
    d = _energy_spread_step(lam_nm, mtotal, rms_nm, gmax)
    return d, _gaussian_focal_weights(mtotal, rms_nm, d)


class Aberrations:
    """Aberration coefficients in the reference's GUI units, converted as
    ``readtccfromhandles_newTEM:2159-2224`` does (Cs, A3, S3, A4, D4, B4 in um -> nm;
    A5, C5, S5, D5 in mm -> nm; C1, A1, A2, B2 in nm; angles in degrees, negated)."""

    def __init__(self, C1=0.0, A1=(0.0, 0.0), A2=(0.0, 0.0), B2=(0.0, 0.0), Cs_um=0.0,
                 A3=(0.0, 0.0), S3=(0.0, 0.0), A4=(0.0, 0.0), D4=(0.0, 0.0), B4=(0.0, 0.0),
                 A5=(0.0, 0.0), C5_mm=0.0, S5=(0.0, 0.0), D5=(0.0, 0.0)):
        def pol(v, scale):
            mag, ang = v
            phi = -ang
            return mag * scale * np.cos(phi / 180 * np.pi), mag * scale * np.sin(phi / 180 * np.pi)

        self.C1 = float(C1)
        self.A1x, self.A1y = pol(A1, 1.0)
        self.A2x, self.A2y = pol(A2, 1.0)
        self.B2x, self.B2y = pol(B2, 1.0)
        self.C3 = float(Cs_um) * 1e3
        self.A3x, self.A3y = pol(A3, 1e3)
        self.S3x, self.S3y = pol(S3, 1e3)
        self.A4x, self.A4y = pol(A4, 1e3)
        self.D4x, self.D4y = pol(D4, 1e3)
        self.B4x, self.B4y = pol(B4, 1e3)
        self.A5x, self.A5y = pol(A5, 1e6)
        self.C5 = float(C5_mm) * 1e6
        self.S5x, self.S5y = pol(S5, 1e6)
        self.D5x, self.D5y = pol(D5, 1e6)


def residual_chi(ab, u, v, lam_nm):
    """Residual aberration function W (nm) beyond A1/C1: ``WholeTCC_residual_newTEM:2232-2268``."""
    wu, wv = u * lam_nm, v * lam_nm
    w = wu + 1j * wv
    wc = wu - 1j * wv
    W = (np.real((ab.A2x / 3 + 1j * ab.A2y / 3) * wc ** 3)
         + np.real((ab.B2x + 1j * ab.B2y) * w ** 2 * wc)
         + np.real((ab.A3x / 4 + 1j * ab.A3y / 4) * wc ** 4)
         + np.real(wc * w ** 3 * (ab.S3x + 1j * ab.S3y))
         + np.real(ab.C3 * w ** 2 * wc ** 2) / 4
         + np.real((ab.A4x / 5 + 1j * ab.A4y / 5) * wc ** 5)
         + np.real((ab.D4x + 1j * ab.D4y) * w ** 4 * wc)
         + np.real((ab.B4x + 1j * ab.B4y) * w ** 3 * wc ** 2)
         + np.real((ab.A5x / 6 + 1j * ab.A5y / 6) * wc ** 6)
         + np.real(w ** 4 * wc ** 2 * (ab.S5x + 1j * ab.S5y))
         + np.real(ab.C5 * w ** 3 * wc ** 3) / 6
         + np.real((ab.D5x + 1j * ab.D5y) * w ** 5 * wc))
    return W


def _chi_gradient(ab, u, v, lam):
    """Gradient of W used by the spatial-coherence envelope (``WholeTCC2D_newTEM:2302-2343``,
    the A1, B2, A2, C3 and C5 terms plus the defocus term added per series member)."""
    gu = (lam * (ab.A1x * u + ab.A1y * v)
          + ab.A2x * lam ** 2 / 3 * (3 * u * u - 3 * v * v) + ab.A2y * lam ** 2 / 3 * 6 * u * v
          + ab.B2x * (3 * u ** 2 + v ** 2) * lam ** 2 - ab.B2y * lam ** 2 * 2 * u * v
          + ab.C3 * lam ** 3 * (u ** 2 + v ** 2) * u
          + 1 / 6 * lam ** 5 * ab.C5 * 3 * (u ** 2 + v ** 2) ** 2 * 2 * u)
    gv = (lam * (ab.A1y * u - ab.A1x * v)
          + ab.B2x * lam * lam * 2 * u * v - ab.B2y * lam * lam * (u ** 2 + 3 * v ** 2)
          + ab.C3 * lam ** 3 * (u ** 2 + v ** 2) * v
          + 1 / 6 * lam ** 5 * ab.C5 * 3 * (u ** 2 + v ** 2) ** 2 * 2 * v)
    return gu, gv


def tcc_stack(n, dx, kv, alpha_mrad, ab=None, mtotal=1, focal_rms_nm=0.0, source_mrad=0.0,
              tilt_mrad=0.0, tilt_phi_deg=0.0):
    """Transfer functions of the focal series, centred: (tcc [M, n, n] complex128, gfsf [M],
    aperture [n, n]) with tcc_m = E_s * E_f,m * exp(-i 2 pi W / lambda) (``WholeTCC2D_newTEM``,
    ``ToTEM_submit.m:2274-2375``) and the aperture |k| <= alpha/lambda (``myaperture`` with flag 1,
    ``:715-736``).  Sampling in nm = dx*0.1 (``CommonPara_TEM:2128``)."""
    ab = ab or Aberrations()
    lam = wavelength_nm(kv)
    sampling = dx * 0.1
    gmax = alpha_mrad * 0.1  # edit_Ape * 0.1 (CommonPara_TEM:2137)
    delta_yita, gfsf = focal_spread(mtotal, focal_rms_nm, lam, gmax)
    tilt = tilt_mrad * 1e-3
    phi = -tilt_phi_deg
    tiltx = tilt / lam * np.cos(phi / 180 * np.pi)
    tilty = tilt / lam * np.sin(phi / 180 * np.pi)
    f = np.arange(-(n // 2), n - n // 2, dtype=np.float64) / (sampling * n)
    U, V = np.meshgrid(f, f)
    U = U + tiltx
    V = V + tilty
    wu, wv = U * lam, V * lam
    W = (np.real((ab.A1x + 1j * ab.A1y) * (wu - 1j * wv) ** 2) / 2
         + np.real(ab.C1 * (wu + 1j * wv) * (wu - 1j * wv)) / 2
         + residual_chi(ab, U, V, lam))
    gu0, gv0 = _chi_gradient(ab, U, V, lam)
    alafa = source_mrad * 1e-3
    # aperture (myaperture with flag 1): keep k1 <= |k| <= k2, grid without the tilt
    fk = np.arange(-_mround((n - 1) / 2), _mround(n / 2), dtype=np.float64) * (1.0 / (sampling * n))
    uu, vv = np.meshgrid(fk, fk)
    kk = np.sqrt(uu * uu + vv * vv)
    ap = np.ones((n, n))
    ap[(kk < 0) | (kk > gmax * 0.01 / lam)] = 0
    M = len(gfsf)
    tcc = np.zeros((M, n, n), dtype=np.complex128)
    for m in range(1, M + 1):
        gu = gu0 + ab.C1 * lam * U
        gv = gv0 + ab.C1 * lam * V
        Es = np.exp(-alafa * alafa * np.pi * np.pi / lam / lam * (gu ** 2 + gv ** 2))
        Ef = np.exp(-1j * np.pi * (m - (M + 1) / 2) * delta_yita * (U ** 2 + V ** 2) * lam)
        tcc[m - 1] = Es * Ef * np.exp(-1j * 2 * np.pi * W / lam)
    return tcc, np.asarray(gfsf, dtype=np.float64), ap


def probe_stack(n, dx, kv, alpha_mrad, ab=None, mtotal=1, focal_rms_nm=0.0, source_mrad=0.0,
                tilt_mrad=0.0, tilt_phi_deg=0.0):
    """Probe wavefunctions [M, n, n] complex128 and focal-series weights gfsf [M]:
    probe_m = fftshift(ifft2(ifftshift(ap * tcc_m))) / sqrt(sum |ifft2(...)|^2)
    (``ToTEM_submit.m:788-806``), tcc_m and ap from :func:`tcc_stack`."""
    tcc, gfsf, ap = tcc_stack(n, dx, kv, alpha_mrad, ab, mtotal, focal_rms_nm, source_mrad, tilt_mrad,
                              tilt_phi_deg)
    probes = []
    for m in range(len(gfsf)):
        w = np.fft.ifft2(np.fft.ifftshift(ap * tcc[m]))
        norm = np.sqrt(np.sum(np.abs(w) ** 2))
        probes.append(np.fft.fftshift(w) / norm)
    return np.stack(probes), gfsf


def dpc_masks(n, dx, kv, rings):
    """Segmented (4-quadrant) detector masks, centred, [4 * len(rings), n, n] (``makeAPER``,
    ``ToTEM_submit.m:1139-1183``).  rings: (inner mrad, outer mrad, rotation degrees) per ring;
    quadrants q = 1..4 of ring i are masks 4i..4i+3, each the ring times the angular sector
    [phi0 + 90(q-1), phi0 + 90q) of atan2(ky, kx), phi0 derived from the rotation as the reference
    does (rotations above 180 reduced by 180, then mapped to [-180, -90))."""
    lam_A = wavelength_A(kv)
    f = centred_freq(n, dx)
    px, py = np.meshgrid(f, f)
    p2 = px * px + py * py
    theta = np.arctan2(py, px)
    out = []
    for k1, k2, rot in rings:
        r1, r2 = k1 * 1e-3 / lam_A, k2 * 1e-3 / lam_A
        ring = ((p2 >= r1 * r1) & (p2 <= r2 * r2)).astype(np.float64)
        a = rot - 180 if rot > 180 else rot
        if -180 <= a < -90:
            d = a
        elif -90 <= a < 0:
            d = a - 90
        elif 0 <= a < 90:
            d = a - 180
        elif 90 <= a < 180:
            d = a - 270
        else:
            raise ValueError("detector rotation out of range: %r" % (rot,))
        q = [((theta >= (d + 90 * j) * np.pi / 180) & (theta < (d + 90 * (j + 1)) * np.pi / 180)).astype(np.float64)
             for j in range(3)]
        q.append(1 - q[0] - q[1] - q[2])
        out.extend(ring * m for m in q)
    return np.stack(out)


def unit_complex64(h):
    """complex128 -> complex64 keeping unit-modulus entries as close to |h| = 1 as FP32 allows.

    Each component is taken from its nearest float or one of that float's two
    neighbours, choosing the pair whose squared modulus is nearest 1 (ties: the
    closer value); entries that are not unit-modulus (e.g. the zeros outside the
    band limit) are rounded normally.  The engine's C ABI applies the same rule to
    the propagator it receives in double precision.
    """
    h = np.asarray(h, dtype=np.complex128)
    c, s = h.real, h.imag
    c0, s0 = c.astype(np.float32), s.astype(np.float32)
    best_c, best_s = c0.copy(), s0.copy()
    best = np.full(h.shape, np.inf)
    for i in (-1, 0, 1):
        cc = c0 if i == 0 else np.nextafter(c0, np.float32(2.0 * i))
        for j in (-1, 0, 1):
            ss = s0 if j == 0 else np.nextafter(s0, np.float32(2.0 * j))
            dn = np.abs(cc.astype(np.float64) ** 2 + ss.astype(np.float64) ** 2 - 1.0)
            dv = np.abs(cc - c) + np.abs(ss - s)
            score = dn * 1e6 + dv
            better = score < best
            best = np.where(better, score, best)
            best_c = np.where(better, cc, best_c)
            best_s = np.where(better, ss, best_s)
    unit = np.abs(np.abs(h) ** 2 - 1.0) < 1e-12
    out = h.astype(np.complex64)
    out[unit] = (best_c + 1j * best_s).astype(np.complex64)[unit]
    return out
