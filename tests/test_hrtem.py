"""Batch HRTEM imaging (the reference's Python tool, cuda_stem64.py).

CPU: pins of the FP64 oracle (oracle/hrtem.py) -- no aberrations give the plain image; its chi is
the MATLAB residual aberration function (optics.residual_chi, restated from another reference
file); its gradient is (1/lambda) grad W, by central differences, for every term the reference
writes as a derivative, and its two deviations are pinned the same way: the A5 terms are the
gradient of a stated polynomial P5 (not of W's A5 term), and the v-gradient's A2 slot is the A2
derivative evaluated with A3's coefficients; the envelope is transposed; the parameter draw; and
the golden images on the reference's pydata.mat inputs.
GPU: the engine (totem_hrtem_images) against the golden images (<= 1e-5 relative L2 per crop,
1e-5 on image sums), against the oracle with every aberration drawn, and the tool's
producer/consumer call shape.
"""
import os
import queue
import types

import numpy as np
import pytest

from oracle import hrtem as oh
from tests.golden import make_golden as mg
from totem_amd import optics

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LAM = 0.00417552


def _inputs():
    return dict(np.load(os.path.join(HERE, "pydata_inputs.npz")))


def _golden():
    return dict(np.load(os.path.join(HERE, "pydata_golden.npz")))


def _grid(n=64, sampling=0.01):
    f = (np.arange(n) - n // 2) / (n * sampling)
    U, V = np.meshgrid(f, f)
    return U, V


def test_no_aberrations_give_the_plain_image():
    U, V = _grid()
    rng = np.random.default_rng(0)
    w = rng.standard_normal(U.shape) + 1j * rng.standard_normal(U.shape)
    p = np.zeros(49)
    p[0] = LAM
    img = oh.images(w, [p], U, V, LAM, alafa=3e-4, gfsf=0.7)[0]
    np.testing.assert_allclose(img, 0.7 * np.abs(np.fft.ifft2(np.fft.ifftshift(w))) ** 2, rtol=1e-12, atol=0)


def test_chi_is_the_matlab_residual_aberration_function():
    U, V = _grid(48)
    rng = np.random.default_rng(1)
    p = np.zeros(49)
    idx = [8, 21] + list(range(28, 36)) + list(range(37, 49))
    p[idx] = rng.standard_normal(len(idx)) * 1e3
    ab = types.SimpleNamespace(A2x=p[28], A2y=p[29], B2x=p[30], B2y=p[31], A3x=p[32], A3y=p[33], S3x=p[34],
                               S3y=p[35], C3=p[8], A4x=p[37], A4y=p[38], D4x=p[39], D4y=p[40], B4x=p[41],
                               B4y=p[42], A5x=p[43], A5y=p[44], S5x=p[45], S5y=p[46], C5=p[21], D5x=p[47],
                               D5y=p[48])
    ref = optics.residual_chi(ab, U, V, LAM)
    np.testing.assert_allclose(oh.chi(p, U, V, LAM), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


EXACT = {1: 10, 8: 1e4, 21: 1e7, 26: 10, 27: 10, 28: 50, 29: 50, 30: 50, 31: 50, 32: 700, 33: 700, 34: 600,
         35: 600, 37: 5e4, 38: 5e4, 39: 5e4, 40: 5e4, 41: 5e4, 42: 5e4, 45: 1e6, 46: 1e6, 47: 1e6, 48: 1e6}
V_AS_WRITTEN = (28, 29, 32, 33)   # v-gradient: A3 sits in the A2 slot (cuda_stem64.py:453-455)


@pytest.mark.parametrize("i", sorted(EXACT))
def test_gradient_is_grad_w_over_lambda(i):
    u = np.array([[3.1, -7.3], [12.2, 5.5]])
    v = np.array([[-4.4, 9.1], [2.2, -11.7]])
    p = np.zeros(49)
    p[0], p[i] = LAM, EXACT[i]
    gu, gv = oh.gradient(p, u, v, LAM)
    h = 1e-5
    nu = (oh.aberration_w(p, u + h, v, LAM) - oh.aberration_w(p, u - h, v, LAM)) / (2 * h) / LAM
    nv = (oh.aberration_w(p, u, v + h, LAM) - oh.aberration_w(p, u, v - h, LAM)) / (2 * h) / LAM
    np.testing.assert_allclose(gu, nu, rtol=1e-8, atol=1e-8 * np.abs(nu).max())
    if i not in V_AS_WRITTEN:
        np.testing.assert_allclose(gv, nv, rtol=1e-8, atol=1e-8 * np.abs(nv).max())


def _p5(p, u, v, la):
    """la^5/6 [A5x (u^4 + v^4 - 14 u v)(u^2 - v^2) - A5y 2 u v (3 u^4 + 3 v^4 - 10 u^2 v^2)]: the
    polynomial whose gradient the reference writes for A5 (cuda_stem64.py:411, :504-505) in
    place of that of Re(A5/6 conj(w)^6)."""
    A5x, A5y = p[43], p[44]
    return la ** 5 / 6 * (A5x * (u ** 4 + v ** 4 - 14 * u * v) * (u ** 2 - v ** 2)
                          - A5y * 2 * u * v * (3 * u ** 4 + 3 * v ** 4 - 10 * u ** 2 * v ** 2))


@pytest.mark.parametrize("i", [43, 44])
def test_a5_gradient_terms_are_the_gradient_of_p5(i):
    u = np.array([[3.1, -7.3], [12.2, 5.5]])
    v = np.array([[-4.4, 9.1], [2.2, -11.7]])
    p = np.zeros(49)
    p[0], p[i] = LAM, 3e6
    gu, gv = oh.gradient(p, u, v, LAM)
    h = 1e-5
    nu = (_p5(p, u + h, v, LAM) - _p5(p, u - h, v, LAM)) / (2 * h)
    nv = (_p5(p, u, v + h, LAM) - _p5(p, u, v - h, LAM)) / (2 * h)
    np.testing.assert_allclose(gu, nu, rtol=1e-7, atol=1e-8 * np.abs(nu).max())
    np.testing.assert_allclose(gv, nv, rtol=1e-7, atol=1e-8 * np.abs(nv).max())
    # and it is not the gradient of the A5 term of W
    wu = (oh.aberration_w(p, u + h, v, LAM) - oh.aberration_w(p, u - h, v, LAM)) / (2 * h) / LAM
    assert np.max(np.abs(gu - wu)) > 0.1 * np.abs(wu).max()


def test_v_gradient_reads_a3_in_the_a2_slot():
    """cuda_stem64.py:461-462: the v-gradient's A2 term is d/dv of Re(A2/3 conj(w)^3) / lambda
    evaluated with A3's coefficients, so A2 adds nothing to g_v and A3 adds that term too."""
    u = np.array([[3.1, -7.3], [12.2, 5.5]])
    v = np.array([[-4.4, 9.1], [2.2, -11.7]])
    h = 1e-5
    for i in (28, 29):
        p = np.zeros(49)
        p[0], p[i] = LAM, 50.0
        _, gv = oh.gradient(p, u, v, LAM)
        assert np.all(gv == 0.0)
    p = np.zeros(49)
    p[0], p[32], p[33] = LAM, 700.0, -400.0
    q = np.zeros(49)
    q[0], q[28], q[29] = LAM, p[32], p[33]          # W's A2 term with A3's coefficients
    w = lambda vv: oh.aberration_w(p, u, vv, LAM) + oh.aberration_w(q, u, vv, LAM)
    nv = (w(v + h) - w(v - h)) / (2 * h) / LAM
    _, gv = oh.gradient(p, u, v, LAM)
    np.testing.assert_allclose(gv, nv, rtol=1e-8, atol=1e-8 * np.abs(nv).max())


def test_envelope_is_transposed():
    U, V = _grid(32)
    p = np.zeros(49)
    p[0], p[26], p[27], p[8] = LAM, 40.0, -25.0, 2e4     # an envelope without u <-> v symmetry
    gu, gv = oh.gradient(p, U, V, LAM)
    e = np.exp((-(3e-4) ** 2 * np.pi ** 2 / LAM ** 2 * (gu ** 2 + gv ** 2)).astype(np.float32)).astype(np.float64)
    assert not np.allclose(e, e.T)
    np.testing.assert_array_equal(oh.envelope(p, U, V, LAM, 3e-4), e.T)


def test_parameter_draw():
    inp = _inputs()
    pp2, mrr = inp["pp2"], inp["mrr"]
    mid = oh.new_pp2(pp2, mrr, np.full((1, 21), 0.5))[0]
    np.testing.assert_array_equal(mid[:26], pp2[:26])
    top = oh.new_pp2(pp2, mrr, np.ones((1, 21)))[0]
    np.testing.assert_allclose(top[1:22], pp2[1:22] + mrr, rtol=1e-15)
    np.testing.assert_allclose(top[26] + 1j * top[27], top[2] * np.exp(1j * top[3] / 180 * np.pi), rtol=1e-14)
    assert top[36] == top[8]


@pytest.mark.parametrize("name", sorted(mg.HRTEM_CASES))
def test_oracle_reproduces_golden_images(name):
    inp, gold = _inputs(), _golden()
    pp1, ai = mg.hrtem_case(inp, name)
    paras = oh.new_pp2(inp["pp2"], inp["mrr"], ai)
    img = oh.images(inp["mywave"], paras, inp["U"], inp["V"], pp1["lambda"], pp1["alafa"], int(pp1["mtotal"]),
                    pp1["delta_yita"], pp1["gfsf"])
    np.testing.assert_allclose(paras, gold[name + "_paras"], rtol=1e-14)
    np.testing.assert_allclose(mg.crop(inp, img), gold[name + "_crop"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(img.sum(axis=(1, 2)), gold[name + "_sum"], rtol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(mg.HRTEM_CASES))
def test_engine_matches_golden_images(engine, name):
    from totem_amd import hrtem
    print("case hrtem %s" % name, flush=True)
    inp, gold = _inputs(), _golden()
    pp1, ai = mg.hrtem_case(inp, name)
    img, paras = hrtem.images(pp1, inp["pp2"], inp["mrr"], inp["U"], inp["V"], inp["mywave"], ai)
    np.testing.assert_allclose(paras, gold[name + "_paras"], rtol=1e-12, atol=1e-12)
    got, ref = mg.crop(inp, img), gold[name + "_crop"]
    for b in range(len(ai)):
        err = np.linalg.norm(got[b] - ref[b]) / np.linalg.norm(ref[b])
        print("  image %d: crop rel L2 %.3e (FP64-oracle error)" % (b, err))
        assert err <= 1e-5
    np.testing.assert_allclose(img.sum(axis=(1, 2)), gold[name + "_sum"], rtol=1e-5)


@pytest.mark.gpu
def test_engine_matches_oracle_with_every_aberration_drawn(engine):
    """All 21 draw ranges non-zero (A5, D4, B4, S5, C5, D5 included, which the golden inputs do
    not draw) on a 128^2 grid with a 3-member focal series."""
    from totem_amd import hrtem
    print("case hrtem every aberration drawn", flush=True)
    n, lam, sampling = 128, 0.00417552, 0.02
    U, V = _grid(n, sampling)
    rng = np.random.default_rng(11)
    w = np.exp(1j * rng.uniform(-0.4, 0.4, (n, n)))
    mywave = np.fft.fftshift(np.fft.fft2(w))
    pp2 = np.zeros(49)
    pp2[0], pp2[1], pp2[2], pp2[8] = lam, 10.0, 10.0, 1.0e3
    mrr = np.abs(rng.normal(0.0, 5.0, 21)) + 1.0
    mrr[18] = 2e5                                    # A5 magnitude (param 19)
    pp1 = {"lambda": lam, "alafa": 3e-4, "delta_yita": 2.0, "gfsf": 0.7, "mtotal": 3}
    ai = rng.random((4, 21))
    img, paras = hrtem.images(pp1, pp2, mrr, U, V, mywave, ai)
    assert np.all(np.abs(paras[:, 43:45]).max(axis=1) > 1e3)
    ref = oh.images(mywave, paras, U, V, lam, pp1["alafa"], 3, pp1["delta_yita"], pp1["gfsf"])
    for b in range(len(ai)):
        err = np.linalg.norm(img[b] - ref[b]) / np.linalg.norm(ref[b])
        print("  image %d: rel L2 %.3e (FP64-oracle error)" % (b, err))
        assert err <= 1e-5


@pytest.mark.gpu
def test_producer_consumer_call_shape(engine, tmp_path):
    from totem_amd import hrtem
    inp = _inputs()
    pp1 = {k: float(inp["pp1_" + k]) for k in ("sampling", "mtotal", "alafa", "lambda", "gmax", "gfsf",
                                                "delta_yita")}
    s = hrtem.stem(pp1, inp["pp2"], inp["mrr"], inp["U"], inp["V"], inp["mywave"], 5, seed=3)
    q = queue.Queue()
    s.producer(q, 2)
    items = []
    while True:
        it = q.get()
        items.append(it)
        if it[2] == 0:
            break
    assert [it[2] for it in items] == [2, 2, 2, 0]
    assert items[0][0].shape == (2, 320, 320) and items[0][1].shape == (2, 49)
    for it in items:
        q.put(it)
    ov = inp["othervar"]
    position = [int(ov[4]), int(ov[4] + ov[5]), int(ov[3]), int(ov[3] + ov[6])]
    n = hrtem.consumer(q, inp["commpara"], 0, str(tmp_path), position)
    assert n == 6
    im = np.load(os.path.join(str(tmp_path), "0.npy"))
    assert im.shape == (int(ov[5]), int(ov[6])) and im.min() == 0 and abs(im.max() - 1) < 1e-6
