"""The oracle's CEC2014 F1-F30 against the reference's own CEC2014.cc.

Golden values come from the reference's objective code compiled from its
source (oracle/_ref), written by tests/golden/make_cec_all_vectors.py:
the reference's data files at D = 10 and 30 (rounded to FP32), and seeded
Haar problems with seeded shuffles at D = 50 and 100.
"""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "cec2014_all_vectors.npz"


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def case_data(O, g, key):
    """x, o, M, shuffle of one golden case."""
    fid, D = int(key.split("_")[1][1:]), int(key.split("_")[2][1:])
    x = g[key + "_x"].astype(np.float64)
    if key.startswith("ref_"):
        o = g[key + "_o"].astype(np.float64)
        M = g[key + "_M"].astype(np.float64)
        S = g[key + "_S"] if (key + "_S") in g.files else None
    else:
        o, M, S = O.gen_cec_problem(int(g[key + "_seed"]), fid, D)
    return fid, D, x, o, M, S


def test_every_function_and_dimension_is_covered(gold):
    cases = set(gold["cases"].tolist())
    for fid in range(1, 31):
        for src, D in (("ref", 10), ("ref", 30), ("haar", 50), ("haar", 100)):
            assert f"{src}_F{fid}_D{D}" in cases


def test_oracle_matches_reference_cec2014(oracle, gold):
    worst = 0.0
    for key in gold["cases"].tolist():
        fid, D, x, o, M, S = case_data(oracle, gold, key)
        f = oracle.cec(fid, x, o, M, S)
        ref = gold[key + "_f"]
        err = np.max(np.abs(f - ref) / np.maximum(np.abs(ref), 1e-300))
        assert err < 1e-12, (key, err)
        worst = max(worst, err)
    assert worst < 1e-12


@pytest.mark.parametrize("fid", [4, 6, 7, 12, 15, 16])
def test_new_basic_functions_vanish_at_the_shift(oracle, fid):
    # Rosenbrock and Griewank-Rosenbrock reach 0 at z = 0 through their fixed
    # shift of 1; Weierstrass, Griewank, Katsuura, Schaffer F6 at z = 0.
    o, M, _ = oracle.gen_cec_problem(7, fid, 20)
    assert oracle.cec(fid, o, o, M)[0] == pytest.approx(0.0, abs=1e-12)


def test_hybrid_segments_follow_fp64_ceil(oracle):
    # ceil(p D) in FP64 (CEC2014.cc:961-967): 0.3 * 50 = 15 and 0.3 * 100 = 30
    # exactly, so the last segment takes the rest.
    assert oracle.hybrid_segments(1, 50) == [15, 15, 20]
    assert oracle.hybrid_segments(1, 100) == [30, 30, 40]
    assert oracle.hybrid_segments(5, 10) == [1, 2, 2, 2, 3]
    assert oracle.hybrid_segments(6, 30) == [3, 6, 6, 6, 9]


def test_hybrid_is_sum_of_segment_objectives(oracle):
    # F17 by hand: z = M (x - o), y = z[S - 1], then Schwefel(10 y[:9]) +
    # Rastrigin(0.0512 y[9:18]) + elliptic(y[18:]) at D = 30.
    D = 30
    o, M, S = oracle.gen_cec_problem(11, 17, D)
    x = np.random.default_rng(3).uniform(-100, 100, D)
    z = M[0] @ (x - o[0])
    y = z[S[0] - 1]
    want = (oracle.basic(oracle.OBJ_SCHWEFEL, 10.0 * y[:9])
            + oracle.basic(oracle.OBJ_RASTRIGIN, 0.0512 * y[9:18])
            + oracle.basic(oracle.OBJ_ELLIPTIC, y[18:]))
    assert oracle.cec(17, x[None, :], o, M, S)[0] == pytest.approx(want, rel=1e-13)


def test_katsuura_and_weierstrass_closed_forms(oracle):
    # Katsuura at one coordinate z = 1/2: each term |2^j / 2 - floor(2^j / 2 + 1/2)|
    # is 0 for j >= 1, so the product is 1 and f = 0; at z = 1/4 only j = 1
    # contributes 1/4, giving 10/D^2 ((1 + 1/4)^(10/D^1.2) - 1) for D = 1.
    assert oracle.basic(oracle.OBJ_KATSUURA, np.array([0.5])) == pytest.approx(0.0, abs=1e-15)
    assert oracle.basic(oracle.OBJ_KATSUURA, np.array([0.25])) == pytest.approx(
        10.0 * (1.25 ** 10.0 - 1.0), rel=1e-12)
    # Weierstrass at integer z: cos(2 pi 3^k (n + 1/2)) = cos(pi 3^k) for every
    # k, so each coordinate cancels its share of the constant term.
    assert oracle.basic(oracle.OBJ_WEIERSTRASS, np.array([1.0, -2.0, 3.0])) == pytest.approx(
        0.0, abs=1e-9)
