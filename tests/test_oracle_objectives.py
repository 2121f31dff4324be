"""Pins the oracle's FP64 objectives (CEC2014.cc:481-530, 590-612, 668-683,
907-953, 1373-1393, 1481-1519) against the reference's own CEC2014.cc (golden vectors
produced by tests/golden/make_cec_ref_vectors.py from oracle/_ref), closed
forms, and the properties a transposed operand or a wrong index would break.
"""
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "cec2014_ref_vectors.npz"


def _cases():
    d = np.load(GOLDEN)
    # the compositions ("comp_F24", "comp_F25") are checked end to end on the GPU
    return [str(c) for c in d["cases"] if not str(c).startswith("comp")]


@pytest.mark.parametrize("case", _cases())
def test_against_reference_cec2014(oracle, case):
    d = np.load(GOLDEN)
    fid = int(case.split("_F")[1].split("_")[0])
    obj = {1: oracle.OBJ_ELLIPTIC, 2: oracle.OBJ_BENT_CIGAR, 3: oracle.OBJ_DISCUS,
           5: oracle.OBJ_ACKLEY, 9: oracle.OBJ_RASTRIGIN, 11: oracle.OBJ_SCHWEFEL,
           13: oracle.OBJ_HAPPYCAT, 14: oracle.OBJ_HGBAT}[fid]
    x = d[case + "_x"]
    if case.startswith("haar"):
        o, M = oracle.gen_problem(int(d[case + "_seed"]), x.shape[1])
    else:
        o, M = d[case + "_o"], d[case + "_M"]
    f = oracle.evaluate(obj, x, o, M)
    # the reference returns f + 100 id - 100 id (CEC2014.cc:144): ~1e-13 absolute
    np.testing.assert_allclose(f, d[case + "_f"], rtol=1e-12, atol=1e-11)


def test_live_reference_if_available(oracle, tmp_path):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    o, M = oracle.gen_problem(555, 20)
    x = np.random.default_rng(0).uniform(-100, 100, (5, 20))
    for fid, obj in ((1, oracle.OBJ_ELLIPTIC), (2, oracle.OBJ_BENT_CIGAR), (3, oracle.OBJ_DISCUS),
                     (5, oracle.OBJ_ACKLEY)):
        oracle.write_cec_files(tmp_path, fid, o, M)
        fr = oracle.ref_eval(tmp_path, fid, x)
        np.testing.assert_allclose(oracle.evaluate(obj, x, o, M), fr, rtol=1e-12)


@pytest.mark.parametrize("obj", [1, 2, 3, 4, 5, 6, 7])
def test_zero_at_shift(oracle, obj):
    o, M = oracle.gen_problem(3, 30)
    f = oracle.evaluate(obj, o[None, :], o, M)[0]
    assert abs(f) < 1e-14


def test_elliptic_unit_vectors(oracle):
    # z = e_k gives 10^(6k/(D-1)): pins the weight index and the orientation
    # of M (z = M (x - o), rows of M, CEC2014.cc:1390).
    D = 7
    rng = np.random.default_rng(1)
    M = np.linalg.qr(rng.normal(size=(D, D)))[0]
    o = np.zeros(D)
    for k in range(D):
        x = M[k]                      # M x = e_k for orthogonal M
        f = oracle.evaluate(oracle.OBJ_ELLIPTIC, x[None, :], o, M)[0]
        assert f == pytest.approx(10.0 ** (6.0 * k / (D - 1)), rel=1e-12)


@pytest.mark.parametrize("obj,w0,wk", [(4, 1.0, 1.0e6), (5, 1.0e6, 1.0)],
                         ids=["bent_cigar", "discus"])
def test_weighted_sphere_unit_vectors(oracle, obj, w0, wk):
    # z = t e_0 gives w0 t^2 and z = t e_k (k >= 1) gives wk t^2: bent cigar
    # weights the first coordinate 1 and the rest 1e6, discus the reverse
    # (CEC2014.cc:507-515, 521-529), which a swapped weight or an off-by-one
    # first index fails.
    D = 6
    rng = np.random.default_rng(2)
    M = np.linalg.qr(rng.normal(size=(D, D)))[0]
    o = rng.uniform(-5, 5, D)
    for k in range(D):
        t = 0.75 + k
        x = o + t * M[k]                       # M (x - o) = t e_k
        f = oracle.evaluate(obj, x[None, :], o, M)[0]
        assert f == pytest.approx((w0 if k == 0 else wk) * t * t, rel=1e-12)


@pytest.mark.parametrize("t", [0.0, 0.5, -3.0, 40.0])
def test_happycat_hgbat_closed_forms(oracle, t):
    # z = 1 + t e_0: u = t e_0, r2 = t^2, s = t, so HappyCat is
    # |t^2 - D|^(1/4) + (t^2 / 2 + t) / D + 1/2 and HGBat |t^4 - t^2|^(1/2) +
    # the same tail (CEC2014.cc:919-928, 943-952): pins the shift by one, the
    # two exponents and which sum D divides.
    D = 5
    z = np.ones(D)
    z[0] += t
    tail = (0.5 * t * t + t) / D + 0.5
    f6 = oracle.evaluate(oracle.OBJ_HAPPYCAT, z[None, :], np.zeros(D), np.eye(D))[0]
    f7 = oracle.evaluate(oracle.OBJ_HGBAT, z[None, :], np.zeros(D), np.eye(D))[0]
    assert f6 == pytest.approx(abs(t * t - D) ** 0.25 + tail, rel=1e-13, abs=1e-13)
    assert f7 == pytest.approx(abs(t ** 4 - t * t) ** 0.5 + tail, rel=1e-13, abs=1e-13)


def test_schwefel_branches(oracle):
    # v = z + 420.9687462275036 at 300 (inside), 600 (folded down to
    # m = 400, penalty 1) and -700 (folded to m = 300 with the sign flipped,
    # penalty 4), D = 3 (CEC2014.cc:717-737).
    c = 4.209687462275036e+002
    z = np.array([300.0, 600.0, -700.0]) - c
    f = oracle.evaluate(oracle.OBJ_SCHWEFEL, z[None, :], np.zeros(3), np.eye(3))[0]
    want = (-300.0 * math.sin(math.sqrt(300.0)) - 400.0 * math.sin(20.0) + 1.0 / 3
            + 300.0 * math.sin(math.sqrt(300.0)) + 4.0 / 3 + 4.189828872724338e+002 * 3)
    assert f == pytest.approx(want, rel=1e-13)


def test_rastrigin_integer_lattice(oracle):
    D = 9
    z = np.array([0, 1, -2, 3, 0, -1, 4, 2, -3], np.float64)
    f = oracle.evaluate(oracle.OBJ_RASTRIGIN, z[None, :], np.zeros(D), np.eye(D))[0]
    assert f == pytest.approx(float((z * z).sum()), abs=1e-11)
    # half-integers: cos(2 pi z) = -1 -> each term is z^2 + 20
    h = z + 0.5
    f = oracle.evaluate(oracle.OBJ_RASTRIGIN, h[None, :], np.zeros(D), np.eye(D))[0]
    assert f == pytest.approx(float((h * h).sum() + 20 * D), abs=1e-10)


def test_ackley_closed_forms(oracle):
    D = 12
    f0 = oracle.basic(oracle.OBJ_ACKLEY, np.zeros(D))
    assert abs(f0) < 1e-14
    f1 = oracle.basic(oracle.OBJ_ACKLEY, np.ones(D))   # r = 1, mean cos = 1
    assert f1 == pytest.approx(20.0 * (1.0 - math.exp(-0.2)), rel=1e-13)
    h = np.full(D, 0.5)                                 # r = 0.5, mean cos = -1
    fh = oracle.basic(oracle.OBJ_ACKLEY, h)
    assert fh == pytest.approx(math.e - 20 * math.exp(-0.1) + 20 - math.exp(-1), rel=1e-13)


def test_rotation_preserves_norm(oracle):
    o, M = oracle.gen_problem(11, 50)
    x = np.random.default_rng(2).uniform(-100, 100, 50)
    # sum z^2 via Ackley's radius: r = sqrt(sum z^2 / D); compare with |x - o|
    z = M @ (x - o)
    assert np.linalg.norm(z) == pytest.approx(np.linalg.norm(x - o), rel=1e-13)
    f_ell_I = oracle.evaluate(oracle.OBJ_ELLIPTIC, z[None, :], np.zeros(50), np.eye(50))[0]
    f_ell = oracle.evaluate(oracle.OBJ_ELLIPTIC, x[None, :], o, M)[0]
    assert f_ell == pytest.approx(f_ell_I, rel=1e-12)


def test_composition_special_cases(oracle):
    D = 4
    objs = [oracle.OBJ_ELLIPTIC, oracle.OBJ_RASTRIGIN, oracle.OBJ_ACKLEY]
    os_ = np.stack([oracle.gen_shift(s, D) for s in (1, 2, 3)])
    Ms = np.stack([oracle.gen_rotation(s, D) for s in (1, 2, 3)])
    sigma, bias = np.array([10.0, 20.0, 30.0]), np.array([0.0, 100.0, 200.0])
    # at o_2 its weight is INF (CEC2014.cc:1498): f = fit_2 = g_2(0) + 100 = 100
    f = oracle.composition(os_[1][None, :].astype(np.float32), objs, os_, Ms, sigma, bias)[0]
    assert f == pytest.approx(100.0, abs=1e-9)
    # far from every o_i all weights underflow -> uniform (CEC2014.cc:1507-1512)
    x = np.full((1, D), 3.0e4, np.float32)
    f = oracle.composition(x, objs, os_, Ms, sigma, bias)[0]
    fits = [oracle.evaluate(objs[c], x.astype(np.float64), os_[c], Ms[c])[0] + bias[c]
            for c in range(3)]
    assert f == pytest.approx(np.mean(fits), rel=1e-12)
    # in between: w_i = d^-1/2 exp(-d / (2 D sigma^2)), normalised
    x = (os_[0] + 3.0).astype(np.float32)[None, :]
    d = ((x.astype(np.float64) - os_) ** 2).sum(1)
    w = d ** -0.5 * np.exp(-d / (2 * D * sigma ** 2))
    fits = np.array([oracle.evaluate(objs[c], x.astype(np.float64), os_[c], Ms[c])[0] + bias[c]
                     for c in range(3)])
    f = oracle.composition(x, objs, os_, Ms, sigma, bias)[0]
    assert f == pytest.approx(float((w / w.sum() * fits).sum()), rel=1e-12)


@pytest.mark.parametrize("fid", [24, 25])
def test_composition_against_reference_cf02_cf03(oracle, fid):
    # the oracle's composition with per-component rates and lambdas against
    # the reference's cf02 / cf03 (CEC2014.cc:1235-1267) on its own F24 / F25
    # data files: F24 = unrotated Schwefel (rate 10), Rastrigin (5.12/100),
    # HGBat (5/100), sigma 20; F25 = Schwefel, Rastrigin, elliptic with
    # lambda 1000/4e3, 1000/1e3, 1000/1e10, sigma 10, 30, 50.
    d = np.load(GOLDEN)
    for D in (10, 30):
        case = f"comp_F{fid}_D{D}"
        x, o, M = d[case + "_x"], d[case + "_o"], d[case + "_M"]
        if fid == 24:
            objs, rate, lam, sig = [8, 2, 7], [10.0, 0.0512, 0.05], [1.0, 1.0, 1.0], [20.0] * 3
            M = M.copy()
            M[0] = np.eye(D)
        else:
            objs, rate, lam, sig = [8, 2, 1], [10.0, 0.0512, 1.0], [0.25, 1.0, 1e-7], [10.0, 30.0, 50.0]
        f = oracle.composition(x, objs, o, M, sig, [0.0, 100.0, 200.0], lam, rate)
        np.testing.assert_allclose(f, d[case + "_f"], rtol=1e-12)
