"""FP64 restatement of the batch HRTEM imaging of ``cuda_stem64.py`` (oracle; test infrastructure only).

One batch element b draws a set of aberrations around the nominal ones and images the exit wave
through them (``producer``, ``cuda_stem64.py:164-704``):

    para_b = createNewPP2(pp2, mrr, ai_b)                                          (:176-240)
    omega = lambda*(U + iV)
    chi   = Re(A2/3 w*^3) + Re(B2 w^2 w*) + Re(A3/4 w*^4) + Re(S3 w* w^3) + C3/4 |w|^4
            + Re(A4/5 w*^5) + Re(D4 w^4 w*) + Re(B4 w^3 w*^2) + Re(A5/6 w*^6)
            + Re(S5 w^4 w*^2) + C5/6 |w|^6 + Re(D5 w^5 w*)                           (getKai :258-311)
    W     = Re(A1 w*^2)/2 + C1 |w|^2/2 + chi                                         (getW :314-336)
    g_u, g_v = the kernels' gradient expressions, term by term                       (:338-519)
    E_s   = expf(-alpha^2 pi^2 / lambda^2 (g_u^2 + g_v^2)), single precision,
            then transposed per batch element                                        (:522-533, :142)
    t     = exp(-2 pi i W / lambda)                                                  (getT :535-547)
    tl_m  = exp(-i pi (m+1-(M+1)/2) dEta |k|^2 lambda) t E_s                          (getTlittleAll :549-579)
    I_b   = gfsf * |ifft2(ifftshift(mywave * tl_0))|^2                               (:681-700)

The last line is what the reference returns for any mtotal M: the quadrant swap covers only the
first ``batches`` of the M*batches products (myIfftShift with P_Slices = batches, :682-687) and
``accumulate_result`` reads element b of them with one weight (gfsf_num = 1, :692-700), so only
the m = 0 focal member reaches the output.

Pins (``tests/test_hrtem.py``): no aberrations and no spread give |ifft2(ifftshift(mywave))|^2;
chi equals the MATLAB residual aberration function (an independent restatement); the gradient
terms of A1, B2, C3, C5 and C1 match a numerical derivative of W; the envelope transpose and the
parameter draw are checked on constructed cases.
"""
import numpy as np

# createNewPP2 (:183-240): indices 1..21 of pp2 are drawn as pp2[i] + (ai[i-1] - 0.5) * 2 * mrr[i-1];
# 26..48 are the x/y components of the (drawn) magnitudes and angles.
_XY = ((26, 2, 3), (28, 4, 5), (30, 6, 7), (32, 9, 10), (34, 11, 12), (37, 13, 14), (39, 15, 16),
       (41, 17, 18), (43, 19, 20), (45, 22, 23), (47, 24, 25))


def new_pp2(pp2, mrr, ai):
    """Drawn parameter rows [B, 49] from the nominal pp2 [49], ranges mrr [21] and uniform draws
    ai [B, 21] (``createNewPP2``, ``cuda_stem64.py:176-240``)."""
    ai = np.atleast_2d(np.asarray(ai, dtype=np.float64))
    B = ai.shape[0]
    r = np.zeros((B, 49))
    r[:, 0] = pp2[0]
    r[:, 1:22] = pp2[1:22][None, :] + (ai[:, :21] - 0.5) * 2 * np.asarray(mrr[:21])[None, :]
    r[:, 22:26] = pp2[22:26][None, :]
    for dst, mag, ang in _XY:
        r[:, dst] = r[:, mag] * np.cos(r[:, ang] / 180 * np.pi)
        r[:, dst + 1] = r[:, mag] * np.sin(r[:, ang] / 180 * np.pi)
    r[:, 36] = r[:, 8]
    return r


def chi(p, u, v, lam):
    """getKai (:258-311) for one parameter row p [49]."""
    w = lam * u + 1j * lam * v
    wc = np.conj(w)
    c = lambda x, y: x + 1j * y
    return (np.real(c(p[28] / 3, p[29] / 3) * wc ** 3) + np.real(c(p[30], p[31]) * w ** 2 * wc)
            + np.real(c(p[32] / 4, p[33] / 4) * wc ** 4) + np.real(c(p[34], p[35]) * wc * w ** 3)
            + p[8] * np.real(w ** 2 * wc ** 2) / 4 + np.real(c(p[37] / 5, p[38] / 5) * wc ** 5)
            + np.real(c(p[39], p[40]) * w ** 4 * wc) + np.real(c(p[41], p[42]) * w ** 3 * wc ** 2)
            + np.real(c(p[43] / 6, p[44] / 6) * wc ** 6) + np.real(c(p[45], p[46]) * w ** 4 * wc ** 2)
            + p[21] * np.real(w ** 3 * wc ** 3) / 6 + np.real(c(p[47], p[48]) * w ** 5 * wc))


def aberration_w(p, u, v, lam):
    """getW (:314-336): W = Re(A1 w*^2)/2 + C1 |w|^2 / 2 + chi."""
    w = lam * u + 1j * lam * v
    return np.real((p[26] + 1j * p[27]) * np.conj(w) ** 2) / 2 + p[1] * np.abs(w) ** 2 / 2 + chi(p, u, v, lam)


def gradient(p, u, v, la):
    """(g_u, g_v) exactly as the kernels write them (:338-519), including their A5 terms."""
    A1x, A1y, A2x, A2y, B2x, B2y = p[26], p[27], p[28], p[29], p[30], p[31]
    A3x, A3y, S3x, S3y, C3 = p[32], p[33], p[34], p[35], p[8]
    A4x, A4y, C1 = p[37], p[38], p[1]
    D4x, D4y, B4x, B4y, A5x, A5y = p[39], p[40], p[41], p[42], p[43], p[44]
    S5x, S5y, C5, D5x, D5y = p[45], p[46], p[21], p[47], p[48]
    r2 = u * u + v * v
    gu = (la * (A1x * u + A1y * v)
          + A2x * la ** 2 / 3 * (3 * u ** 2 - 3 * v ** 2) + A2y * la ** 2 / 3 * 6 * u * v
          + B2x * (3 * u ** 2 + v ** 2) * la ** 2 - B2y * la ** 2 * 2 * u * v
          + 1.0 / 4 * la ** 3 * (A3x * (4 * u ** 3 - 12 * u * v ** 2) + A3y * (12 * u ** 2 * v - 4 * v ** 3))
          + la ** 3 * (2 * u * (S3x * (u ** 2 - v ** 2) - S3y * 2 * u * v) + r2 * (2 * S3x * u - 2 * S3y * v))
          + C3 * la ** 3 * r2 * u
          + 1.0 / 5 * la ** 4 * (A4x * (5 * u ** 4 + 5 * v ** 4 - 30 * u ** 2 * v ** 2)
                                 + A4y * (20 * u ** 3 * v - 20 * u * v ** 3)))
    gu = (gu
          + la ** 4 * ((2 * u) * (D4x * (u ** 3 - 3 * u * v ** 2) - D4y * (3 * u ** 2 * v - v ** 3))
                       + r2 * (D4x * (3 * u ** 2 - 3 * v ** 2) - D4y * 6 * u * v))
          + la ** 4 * (2 * r2 * 2 * u * (B4x * u - B4y * v) + r2 ** 2 * B4x)
          + 1.0 / 6 * la ** 5 * (A5x * (4 * u ** 3 - 14 * v) * (u ** 2 - v ** 2)
                                 + A5x * (u ** 4 + v ** 4 - 14 * u * v) * 2 * u
                                 - A5y * 2 * v * (3 * u ** 4 + 3 * v ** 4 - 10 * u ** 2 * v ** 2)
                                 - A5y * 2 * v * u * (12 * u ** 3 - 20 * u * v ** 2))
          + la ** 5 * (2 * r2 * 2 * u * (S5x * (u ** 2 - v ** 2) - 2 * u * v * S5y) + r2 ** 2 * (2 * S5x * u - 2 * v * S5y))
          + 1.0 / 6 * la ** 5 * C5 * 3 * r2 ** 2 * 2 * u
          + la ** 5 * (2 * u * (D5x * (u ** 4 + v ** 4 - 6 * u ** 2 * v ** 2) - D5y * 4 * u * v * (u ** 2 - v ** 2))
                       + r2 * (D5x * (4 * u ** 3 - 12 * u * v ** 2) - D5y * (12 * u ** 2 * v - 4 * v ** 3)))
          + C1 * la * u)
    gv = (la * (A1y * u - A1x * v)
          - A3x * la ** 2 / 3 * 6 * u * v
          + A3y * la ** 2 / 3 * (-3 * v ** 2 + 3 * u ** 2)
          + B2x * la ** 2 * 2 * u * v
          - B2y * la ** 2 * (u ** 2 + 3 * v ** 2)
          + 1.0 / 4 * la ** 3 * (A3x * (4 * v ** 3 - 12 * u ** 2 * v) + A3y * (4 * u ** 3 - 12 * u * v ** 2))
          + la ** 3 * (2 * v * (S3x * (u ** 2 - v ** 2) - S3y * 2 * u * v) + r2 * (-2 * S3x * v - 2 * S3y * u))
          + C3 * la ** 3 * r2 * v
          + 1.0 / 5 * la ** 4 * (A4x * (20 * u * v ** 3 - 20 * u ** 3 * v) + A4y * (5 * v ** 4 + 5 * u ** 4 - 30 * u ** 2 * v ** 2)))
    gv = (gv
          + la ** 4 * ((2 * v) * (D4x * (u ** 3 - 3 * u * v ** 2) - D4y * (3 * u ** 2 * v - v ** 3))
                       + r2 * (-D4x * 6 * u * v - D4y * (3 * u ** 2 - 3 * v ** 2)))
          + la ** 4 * (2 * r2 * 2 * v * (B4x * u - B4y * v) - r2 ** 2 * B4y)
          + 1.0 / 6 * la ** 5 * (A5x * (4 * v ** 3 - 14 * u) * (u ** 2 - v ** 2)
                                 + A5x * (u ** 4 + v ** 4 - 14 * u * v) * (-2 * v)
                                 - A5y * 2 * u * (3 * u ** 4 + 3 * v ** 4 - 10 * u ** 2 * v ** 2)
                                 - A5y * 2 * v * u * (12 * v ** 3 - 20 * u ** 2 * v))
          + la ** 5 * (2 * r2 * 2 * v * (S5x * (u ** 2 - v ** 2) - 2 * u * v * S5y) + r2 ** 2 * (-2 * S5x * v - 2 * u * S5y))
          + 1.0 / 6 * la ** 5 * C5 * 3 * r2 ** 2 * 2 * v
          + la ** 5 * (2 * v * (D5x * (u ** 4 + v ** 4 - 6 * u ** 2 * v ** 2) - D5y * 4 * u * v * (u ** 2 - v ** 2))
                       + r2 * (D5x * (4 * v ** 3 - 12 * u ** 2 * v) - D5y * (4 * u ** 3 - 12 * u * v ** 2)))
          + C1 * la * v)
    return gu, gv


def envelope(p, u, v, lam, alafa):
    """Spatial-coherence envelope in single precision, transposed (getESCoh :522-533, :142)."""
    gu, gv = gradient(p, u, v, lam)
    arg = (-1 * alafa * alafa * np.pi * np.pi / lam / lam * (gu ** 2 + gv ** 2)).astype(np.float32)
    return np.exp(arg).astype(np.float64).T


def transfer(p, u, v, lam, alafa, mtotal=1, delta_yita=0.0, m=0):
    """tl_m for one parameter row: exp(-i pi (m+1-(M+1)/2) dEta |k|^2 lambda) exp(-2 pi i W / lambda) E_s."""
    t = np.exp(-1j * 2 * np.pi * aberration_w(p, u, v, lam) / lam)
    t1 = -1 * np.pi * (m + 1 - (mtotal + 1) * 1.0 / 2) * delta_yita * (u ** 2 + v ** 2) * lam
    return np.exp(1j * t1) * t * envelope(p, u, v, lam, alafa)


def images(mywave, paras, u, v, lam, alafa, mtotal=1, delta_yita=0.0, gfsf=1.0):
    """I_b = gfsf |ifft2(ifftshift(mywave * tl_0))|^2 for every parameter row (:681-700)."""
    out = np.zeros((len(paras),) + mywave.shape)
    for b, p in enumerate(paras):
        x = np.fft.ifft2(np.fft.ifftshift(mywave * transfer(p, u, v, lam, alafa, mtotal, delta_yita, 0)))
        out[b] = gfsf * np.abs(x) ** 2
    return out
