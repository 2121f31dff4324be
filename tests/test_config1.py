"""Config 1 of the benchmark set: one island, NP=100, D=10, DE/rand/1/bin
with F=0.5, CR=0.9, bounds [-100, 100], shifted-rotated Rastrigin
(o ~ U[-80, 80]^D, Haar M), 1000 generations -- the reference's single-island
CPU path (DE_CUDA::Reproduce_CPU / SelectSurvival_CPU, CUDA_EA_library/
DE_CUDA.cc:121-198, with CEC2014::EvaluateFitness, CEC2014.cc:146-279).

CPU: the oracle runs the 1000 generations itself (FP32 DE operators, FP64
objective, the reference's CPU arithmetic) and must converge.
GPU: the engine runs the same 1000 generations and every generation is
checked in lockstep against the oracle started from the engine's own state:
trials and selection bit-exact, trial fitness within 1e-5 of FP64; the epoch
kernel then reproduces the whole run bit for bit."""
import numpy as np
import pytest

SEED = 1001
NP, D, GENS = 100, 10, 1000


def _problem(oracle):
    return oracle.gen_problem(oracle.problem_seed(SEED, 0), D)


def test_config1_oracle_run_converges(oracle):
    o, M = _problem(oracle)
    ld = 12
    X = oracle.init_population(NP, D, ld, -100.0, 100.0, SEED)
    f = oracle.evaluate(2, X[:, :D], o, M).astype(np.float32)
    best = [float(f.min())]
    for g in range(GENS):
        U, *_ = oracle.de_generate(X, ld, NP, D, seed=SEED, gen=g, U=np.zeros_like(X))
        fU = oracle.evaluate(2, U[:, :D], o, M).astype(np.float32)
        X, f, _ = oracle.de_select(X, f, U, fU, ld, NP, D)
        best.append(float(f.min()))
    best = np.array(best)
    assert np.all(np.diff(best) <= 0)            # elitist: best never worsens
    assert best[-1] < 0.05 * best[0]
    assert np.all(np.abs(X[:, :D]) <= 100.0)     # reflection keeps the bounds


@pytest.mark.gpu
def test_config1_gpu_lockstep(cuda, speo, oracle):
    o, M = _problem(oracle)
    prob = speo.Problem([2], o[None, :], M[None, :, :])
    e = speo.Engine([prob], [NP], [0], F=0.5, CR=0.9, seed=SEED, interval=0, migration_size=0,
                    mode=speo.MODE_PLAIN, rows_per_block=32)
    X, f = e.read()
    ld = X.shape[1]
    X0 = oracle.init_population(NP, D, ld, -100.0, 100.0, SEED)
    np.testing.assert_array_equal(X[:, :D], X0[:, :D])
    worst_rel = 0.0
    for g in range(GENS):
        e.run(1)
        U, fU = e.read_trials()
        X1, f1 = e.read()
        Uref, *_ = oracle.de_generate(X, ld, NP, D, seed=SEED, gen=g, U=np.zeros_like(X))
        np.testing.assert_array_equal(U[:, :D], Uref[:, :D], err_msg=f"generation {g}")
        ref = oracle.evaluate(2, Uref[:, :D], o, M)
        worst_rel = max(worst_rel, float(np.max(np.abs(fU - ref) / np.abs(ref))))
        Xs, fs, _ = oracle.de_select(X, f, U, fU, ld, NP, D)
        np.testing.assert_array_equal(X1, Xs, err_msg=f"generation {g}")
        np.testing.assert_array_equal(f1, fs, err_msg=f"generation {g}")
        X, f = X1, f1
    assert worst_rel < 1e-5
    # the epoch kernel (one launch for all 1000 generations) reproduces the run
    ep = speo.Engine([prob], [NP], [0], F=0.5, CR=0.9, seed=SEED, interval=0, migration_size=0,
                     mode=speo.MODE_EPOCH)
    ep.run(GENS)
    Xe, fe = ep.read()
    np.testing.assert_array_equal(Xe, X)
    np.testing.assert_array_equal(fe, f)
    _, _, f64 = e.best()
    assert f64[0] < 0.05 * float(oracle.evaluate(2, X0[:, :D], o, M).min())
