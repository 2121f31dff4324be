"""The seeded problem generator: o ~ U[-80, 80]^D rounded to FP32, M Haar
orthogonal (QR of a Philox/Box-Muller N(0,1) matrix, diag(R) > 0).  The
library's C++ generator (Householder QR) is checked against the oracle's
numpy one (LAPACK QR) on the same Gaussian matrix."""
import numpy as np
import pytest


@pytest.mark.parametrize("D", [2, 10, 100])
def test_oracle_generator_properties(oracle, D):
    o, M = oracle.gen_problem(42, D)
    assert ((o >= -80) & (o <= 80)).all()
    assert np.array_equal(o, o.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(M @ M.T, np.eye(D), atol=1e-12)
    G = oracle.gen_gauss(42, D)
    R = M.T @ G                         # G = M R with R upper triangular, diag > 0
    assert np.allclose(np.tril(R, -1), 0, atol=1e-10)
    assert (np.diag(R) > 0).all()


def test_gauss_moments(oracle):
    G = oracle.gen_gauss(7, 300)
    assert abs(G.mean()) < 0.01 and abs(G.std() - 1) < 0.01


def test_haar_entry_variance(oracle):
    M = oracle.gen_rotation(9, 200)
    assert abs(M.var() - 1.0 / 200) < 0.1 / 200


@pytest.mark.parametrize("D", [2, 3, 10, 64, 100])
def test_library_generator_matches_oracle(oracle, speo, D):
    seed = speo.problem_seed(1234, 5)
    assert seed == oracle.problem_seed(1234, 5)
    o, M = speo.generate_problem_data(seed, D)
    o2, M2 = oracle.gen_problem(seed, D)
    np.testing.assert_array_equal(o, o2)
    np.testing.assert_allclose(M, M2, atol=1e-12)
