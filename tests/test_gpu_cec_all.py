"""GPU parity of CEC2014 F1-F30 through the drop-in API_evaluateFitness, the
problem API (speo_problem_create_cec / _ex, speo_evaluate, speo_evaluate_f64)
and the engine's plain mode, against the oracle (itself pinned to the
reference's CEC2014.cc by tests/test_oracle_cec_all.py) and the golden values.

The drop-in takes the reference's device layouts (CEC2014_CUDA.cc:68-289):
rows [n][next_pow2(D)], shifts [ncomp][next_pow2(D)], B = M^T row-padded
[ncomp][next_pow2(D)][D], 1-based shuffles [ncomp][next_pow2(D)], sigma
[ncomp].  FP32 outputs are held to 1e-5 relative of FP64; the fused FP32
path of F1-F3, F5, F8-F11, F13, F14 keeps its per-row bounds (tolerance.py).
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from tolerance import assert_rows, row_rtol, schwefel_rows

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "cec2014_all_vectors.npz"
FUSED = {1: (1, 1.0, True), 2: (4, 1.0, True), 3: (5, 1.0, True), 5: (3, 1.0, True),
         8: (2, 0.0512, False), 9: (2, 0.0512, True), 10: (8, 10.0, False),
         11: (8, 10.0, True), 13: (6, 0.05, True), 14: (7, 0.05, True)}
SIGMA = {23: [10, 20, 30, 40, 50], 24: [20, 20, 20], 25: [10, 30, 50], 26: [10] * 5,
         27: [10, 10, 10, 20, 20], 28: [10, 20, 30, 40, 50], 29: [10, 30, 50], 30: [10, 30, 50]}


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def dropin_eval(speo, dev, fid, X, o, M, S, bias=0.0, sigma=None):
    """API_evaluateFitness on the reference's layouts; X [n][D] float32."""
    import torch
    L = speo.lib()
    n, D = X.shape
    nc = o.shape[0]
    Dp = L.SPEO_NextPow2Dim(D)
    Xp = np.zeros((n, Dp), np.float32)
    Xp[:, :D] = X
    sh = np.zeros((nc, Dp), np.float32)
    sh[:, :D] = o
    B = np.zeros((nc, Dp, D), np.float32)
    for c in range(nc):
        B[c, :D, :] = M[c].T
    Sp = np.zeros((nc, Dp), np.int32)
    if S is not None:
        Sp[:, :D] = S
    sg = np.asarray(sigma if sigma is not None else SIGMA.get(fid, [10.0] * nc), np.float32)
    f = torch.zeros(n, dtype=torch.float32, device=dev)
    scratch = torch.zeros(n * Dp, dtype=torch.float32, device=dev)  # d_rotated_elements
    Xd, shd, Bd, Sd, sgd = (_t(a, dev) for a in (Xp, sh, B, Sp, sg))
    P = ctypes.c_void_p
    L.API_evaluateFitness(P(f.data_ptr()), None, P(scratch.data_ptr()), P(Xd.data_ptr()),
                          P(shd.data_ptr()), P(Bd.data_ptr()), P(Sd.data_ptr()), None, None,
                          P(sgd.data_ptr()), n, D, float(bias), fid)
    assert L.API_CheckCUDAError() == 0, speo.last_error()
    return f.cpu().numpy()


def gold_cases(D):
    g = np.load(GOLD)
    return g, [c for c in g["cases"].tolist() if c.endswith(f"_D{D}")]


def case_problem(oracle, g, key):
    fid = int(key.split("_")[1][1:])
    D = int(key.split("_")[2][1:])
    if key.startswith("ref_"):
        o = g[key + "_o"].astype(np.float64)
        M = g[key + "_M"].astype(np.float64)
        S = g[key + "_S"] if (key + "_S") in g.files else None
    else:
        o, M, S = oracle.gen_cec_problem(int(g[key + "_seed"]), fid, D)
    return fid, D, o, M, S


def check_fused(fid, f, X, o, M):
    obj, rate, rot = FUSED[fid]
    Mu = M[0] if rot else np.eye(o.shape[1])
    Xs = o[0][None, :] + rate * (X.astype(np.float64) - o[0][None, :])
    return obj, Mu, Xs


@pytest.mark.parametrize("D", [10, 30, 50, 100])
def test_dropin_every_function(cuda, speo, oracle, D):
    """All 30 IDs through API_evaluateFitness at the golden cases' points plus
    29 random rows (37 rows: a partial 16-row tile)."""
    g, cases = gold_cases(D)
    assert len(cases) == 30
    rng = np.random.default_rng(D)
    for key in cases:
        fid, D_, o, M, S = case_problem(oracle, g, key)
        o32 = o.astype(np.float32).astype(np.float64)
        M32 = M.astype(np.float32).astype(np.float64)
        X = np.concatenate([g[key + "_x"],
                            rng.uniform(-100, 100, (29, D)).astype(np.float32)]).astype(np.float32)
        print(f"dropin F{fid} D={D}", flush=True)
        f = dropin_eval(speo, cuda, fid, X, o32, M32, S)
        if fid in FUSED:
            obj, Mu, Xs = check_fused(fid, f, X, o32, M32)
            ref = oracle.evaluate(obj, Xs, o32[0], Mu)
            if obj == 8:
                keep = ~schwefel_rows(X, o32[0], Mu, FUSED[fid][1])[1]
                assert keep.sum() >= X.shape[0] - 2
                np.testing.assert_allclose(f[keep], ref[keep], rtol=1e-5, err_msg=key)
            else:
                assert_rows(f, ref, row_rtol(obj, Xs, o32[0], Mu))
        else:
            ref = oracle.cec(fid, X.astype(np.float64), o32, M32, S)
            np.testing.assert_allclose(f, ref, rtol=1e-5, atol=0, err_msg=key)
        if key.startswith("ref_") and fid not in FUSED:
            # the golden values were computed on these very FP32 data
            np.testing.assert_allclose(f[:g[key + "_f"].size], g[key + "_f"], rtol=1e-5,
                                       err_msg=key)


def test_dropin_bias_only_for_hybrids_and_compositions(cuda, speo, oracle):
    # F1-F16 ignore the call's bias; the hybrid and composition chains add it
    # in CombineFitness (kernel_CEC2014_CUDA.cu:205).
    D = 10
    X = np.random.default_rng(1).uniform(-100, 100, (20, D)).astype(np.float32)
    for fid in (4, 7, 17, 23):
        o, M, S = oracle.gen_cec_problem(40 + fid, fid, D)
        o, M = o.astype(np.float32).astype(np.float64), M.astype(np.float32).astype(np.float64)
        f0 = dropin_eval(speo, cuda, fid, X, o, M, S, bias=0.0)
        f1 = dropin_eval(speo, cuda, fid, X, o, M, S, bias=100.0)
        if fid >= 17:
            np.testing.assert_allclose(f1, (f0.astype(np.float64) + 100.0).astype(np.float32),
                                       rtol=1e-6)
        else:
            np.testing.assert_array_equal(f1, f0)


def test_dropin_sigma_comes_from_the_callers_array(cuda, speo, oracle):
    # CalW reads d_sigma (kernel_CEC2014_CUDA.cu:109): other widths, other values
    D, fid = 10, 26
    o, M, S = oracle.gen_cec_problem(77, fid, D)
    o, M = o.astype(np.float32).astype(np.float64), M.astype(np.float32).astype(np.float64)
    X = (o[1] + np.random.default_rng(2).normal(0, 3, (16, D))).astype(np.float32)
    f_ref_sigma = dropin_eval(speo, cuda, fid, X, o, M, S)
    f_wide = dropin_eval(speo, cuda, fid, X, o, M, S, sigma=[1000.0] * 5)
    assert np.max(np.abs(f_wide - f_ref_sigma) / np.abs(f_ref_sigma)) > 1e-3


@pytest.mark.parametrize("D", [10, 50, 100])
def test_problem_from_cec_files_every_function(cuda, speo, oracle, tmp_path, D):
    """speo_problem_create_cec reads shifts, matrices and shuffles written in
    the reference's text format; FP32 evaluation within 1e-5 of the oracle and
    FP64 re-verification within 1e-11."""
    import torch
    rng = np.random.default_rng(100 + D)
    X = rng.uniform(-100, 100, (40, D)).astype(np.float32)
    Xd = _t(X, cuda)
    for fid in range(1, 31):
        o, M, S = oracle.gen_cec_problem(300 + fid, fid, D)
        d = tmp_path / f"F{fid}"
        oracle.write_cec_files(d, fid, o, M, S)
        print(f"problem F{fid} D={D}", flush=True)
        p = speo.Problem.from_cec(d / "input_data", fid, D)
        assert p.ncomp == oracle.cec_ncomp(fid)
        f32 = torch.zeros(40, dtype=torch.float32, device=cuda)
        f64 = torch.zeros(40, dtype=torch.float64, device=cuda)
        speo.evaluate(p, Xd, f32)
        speo.evaluate_f64(p, Xd, f64)
        torch.cuda.synchronize()
        ref = oracle.cec(fid, X.astype(np.float64), o, M, S)
        np.testing.assert_allclose(f64.cpu().numpy(), ref, rtol=1e-11, err_msg=f"F{fid}")
        if fid not in FUSED and fid not in (24, 25):
            np.testing.assert_allclose(f32.cpu().numpy(), ref, rtol=1e-5, err_msg=f"F{fid}")
        p.close()


@pytest.mark.parametrize("D", [200, 1000])
def test_generic_kernel_large_dimension(cuda, speo, oracle, D):
    """The generic kernel's 16-row Z tile holds up to 1024 columns in LDS."""
    import torch
    rng = np.random.default_rng(D)
    X = rng.uniform(-100, 100, (20, D)).astype(np.float32)
    for obj in (speo.OBJ_ROSENBROCK, speo.OBJ_WEIERSTRASS, speo.OBJ_KATSUURA,
                speo.OBJ_HYBRID_6):
        fid = {speo.OBJ_ROSENBROCK: 4, speo.OBJ_WEIERSTRASS: 6, speo.OBJ_KATSUURA: 12,
               speo.OBJ_HYBRID_6: 22}[obj]
        o, M, S = oracle.gen_cec_problem(500 + fid, fid, D)
        rate = {4: 0.02048, 6: 0.005, 12: 0.05, 22: 1.0}[fid]
        p = speo.Problem([obj], o, M, rate=[rate], shuffle=S)
        f64 = torch.zeros(20, dtype=torch.float64, device=cuda)
        speo.evaluate_f64(p, _t(X, cuda), f64)
        torch.cuda.synchronize()
        ref = oracle.cec(fid, X.astype(np.float64), o, M, S)
        np.testing.assert_allclose(f64.cpu().numpy(), ref, rtol=1e-10, err_msg=f"F{fid} D={D}")
        p.close()


def test_engine_plain_mode_with_hybrids_and_compositions(cuda, speo, oracle):
    """Islands on F17 (hybrid), F23 (composition) and F30 (composition of
    hybrids) next to an F9 island: trials from the oracle's DE step bit for
    bit, their fitness within 1e-5 and the best rows' FP64 values within
    1e-11 of the oracle."""
    D, NP, seed = 30, 48, 777
    fids = [17, 23, 30, 9]
    data = [oracle.gen_cec_problem(900 + f, f, D) for f in fids]
    probs = []
    for f, (o, M, S) in zip(fids, data):
        nc = oracle.cec_ncomp(f)
        import tempfile
        d = Path(tempfile.mkdtemp())
        oracle.write_cec_files(d, f, o, M, S)
        probs.append(speo.Problem.from_cec(d / "input_data", f, D))
    eng = speo.Engine(probs, [NP] * 4, [0, 1, 2, 3], seed=seed, mode=speo.MODE_PLAIN,
                      rows_per_block=16)
    X0, f0 = eng.read()
    for i, f in enumerate(fids):
        o, M, S = data[i]
        rows = slice(i * NP, (i + 1) * NP)
        ref0 = oracle.cec(f, X0[rows, :D].astype(np.float64), o, M, S)
        rt = 1e-5 if f != 9 else 2e-5
        np.testing.assert_allclose(f0[rows], ref0, rtol=rt, err_msg=f"init F{f}")
    eng.run(1)
    U, fU = eng.read_trials()
    for i, f in enumerate(fids):
        o, M, S = data[i]
        rows = slice(i * NP, (i + 1) * NP)
        Xi = np.ascontiguousarray(X0[rows])
        Uref, *_ = oracle.de_generate(Xi, D, NP, D, seed=seed, gen=0, row_base=i * NP,
                                      U=np.zeros_like(Xi))
        np.testing.assert_array_equal(U[rows], Uref)
        ref = oracle.cec(f, Uref.astype(np.float64), o, M, S)
        rt = 1e-5 if f != 9 else 2e-5
        np.testing.assert_allclose(fU[rows], ref, rtol=rt, err_msg=f"trials F{f}")
    eng.run(4)
    rows_b, f32b, f64b = eng.best()
    Xn, fn = eng.read()
    for i, f in enumerate(fids):
        o, M, S = data[i]
        ref = oracle.cec(f, Xn[rows_b[i]:rows_b[i] + 1, :D].astype(np.float64), o, M, S)[0]
        assert f64b[i] == pytest.approx(ref, rel=1e-11), f"best F{f}"
    eng.close()
    for p in probs:
        p.close()


def test_engine_rejects_fused_mode_for_generic_objectives(cuda, speo, oracle, tmp_path):
    o, M, S = oracle.gen_cec_problem(5, 4, 10)
    p = speo.Problem([speo.OBJ_ROSENBROCK], o, M, rate=[0.02048])
    with pytest.raises(speo.SpeoError, match="SPEO_MODE_PLAIN") as e:
        speo.Engine([p], [32], [0], seed=1, mode=speo.MODE_FUSED)
    assert e.value.code == -3
    p.close()
