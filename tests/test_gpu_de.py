"""GPU parity of the DE operators through the C ABI: bit-exact against the
oracle (kernel_DE_CUDA.cu:160-327, kernel_EA_CUDA.cu:91-118 restated) on the
same seeded Philox draws and on injected edge draws."""
import ctypes

import numpy as np
import pytest

from tolerance import schwefel_rows

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _de_info(speo, dev, cr, f, strategy=4):
    import torch
    raw = speo.de_info_bytes(cr, f, strategy)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)


def _sync():
    import torch
    torch.cuda.synchronize()


@pytest.mark.parametrize("NP,D,ld,row_base", [(100, 37, 40, 0), (7, 3, 4, 11), (256, 100, 100, 5),
                                              (64, 1000, 1000, 0)])
def test_init_population_bit_exact(cuda, speo, oracle, NP, D, ld, row_base):
    import torch
    X = torch.full((NP, ld), -7.0, dtype=torch.float32, device=cuda)
    print(f"init NP={NP} D={D}", flush=True)
    speo.init_population(X, D, -100.0, 100.0, seed=0xBEEF, row_base=row_base)
    _sync()
    ref = oracle.init_population(NP, D, ld, -100.0, 100.0, 0xBEEF, row_base)
    np.testing.assert_array_equal(X.cpu().numpy(), ref)


CASES = [
    # NP, D, ld, subislands, F, CR, strategy, permute
    (4, 10, 12, 1, 0.5, 0.9, 4, False),
    (100, 10, 12, 1, 0.5, 0.9, 4, False),      # NP not a multiple of 16
    (128, 100, 100, 1, 0.7, 0.1, 4, False),
    (512, 100, 128, 4, 0.5, 0.9, 4, True),
    (96, 1000, 1000, 2, 0.5, 0.9, 4, True),
] + [(64, 33, 36, 4, 0.6, 0.5, st, True) for st in range(9)]


@pytest.mark.parametrize("NP,D,ld,subs,F,CR,st,permute", CASES)
def test_de_generate_bit_exact(cuda, speo, oracle, NP, D, ld, subs, F, CR, st, permute):
    import torch
    rng = np.random.default_rng(NP * 1000 + D + st)
    X = np.zeros((NP, ld), np.float32)
    X[:, :D] = rng.uniform(-100, 100, (NP, D)).astype(np.float32)
    perm = rng.permutation(NP).astype(np.int32) if permute else None
    S = NP // subs
    cr = np.full(subs, CR, np.float32)
    f = np.full(subs, F, np.float32)
    best = np.array([perm[s * S] if permute else s * S for s in range(subs)], np.int32)
    U0 = rng.uniform(-1, 1, (NP, ld)).astype(np.float32)
    Xd, Ud = _t(X, cuda), _t(U0, cuda)
    r_out = torch.zeros((NP, 5), dtype=torch.int32, device=cuda)
    jr_out = torch.zeros(NP, dtype=torch.int32, device=cuda)
    print(f"de_generate NP={NP} D={D} st={st}", flush=True)
    speo.de_generate(Ud, Xd, D, _de_info(speo, cuda, cr, f, st), subisland_num=subs,
                     perm=None if perm is None else _t(perm, cuda), best=_t(best, cuda),
                     seed=77, gen=5, row_base=3, r_out=r_out, jrand_out=jr_out)
    _sync()
    Uref, r, jr, _ = oracle.de_generate(X, ld, NP, D, subisland_num=subs, perm=perm, cr=cr, f=f,
                                        strategy=np.full(subs, st, np.int32), best=best,
                                        seed=77, gen=5, row_base=3, U=U0)
    np.testing.assert_array_equal(r_out.cpu().numpy(), r)
    np.testing.assert_array_equal(jr_out.cpu().numpy(), jr)
    np.testing.assert_array_equal(Ud.cpu().numpy(), Uref)


def test_de_generate_injected_edge_draws(cuda, speo, oracle):
    # u = 1.0 draws (== S decrement, j_rand == D), forced collisions and
    # crossover draws exactly at CR.
    import torch
    NP, D, ld = 8, 5, 8
    rng = np.random.default_rng(5)
    X = np.zeros((NP, ld), np.float32)
    X[:, :D] = rng.uniform(-100, 100, (NP, D)).astype(np.float32)
    # row 0: r0 = 7 (1.0 -> S -> 7), r1: 1.0 collides -> 0.5 -> 4, r2: 0.5
    # collides -> 0.01 -> 0, r3: 0.01 collides -> 0.26 -> 2, r4: 0.9 -> 7
    # collides -> 1.0 -> 8 -> 7 (after the loop), j_rand: 1.0 -> D.
    inj_r = np.array([[1.0, 1.0, 0.5, 0.5, 0.01, 0.01, 0.26, 0.9, 1.0, 1.0]] * NP, np.float32)
    inj_r[3] = [0.3, 0.3, 0.3, 0.6, 0.6, 0.75, 0.99, 0.2, 0.4, 0.5]
    inj_cr = np.full((NP, D), np.float32(0.9), np.float32)   # u == CR is not < CR
    inj_cr[1] = [0.1, 1.0, 0.89999, 0.9, 2 ** -24]
    Ud = torch.zeros((NP, ld), dtype=torch.float32, device=cuda)
    r_out = torch.zeros((NP, 5), dtype=torch.int32, device=cuda)
    jr_out = torch.zeros(NP, dtype=torch.int32, device=cuda)
    speo.de_generate(Ud, _t(X, cuda), D, _de_info(speo, cuda, [0.9], [0.5]), seed=1, gen=0,
                     inj_r=_t(inj_r, cuda), inj_cr=_t(inj_cr, cuda), r_out=r_out,
                     jrand_out=jr_out)
    _sync()
    Uref, r, jr, _ = oracle.de_generate(X, ld, NP, D, seed=1, gen=0, inj_r=inj_r, inj_cr=inj_cr)
    np.testing.assert_array_equal(r_out.cpu().numpy(), r)
    np.testing.assert_array_equal(jr_out.cpu().numpy(), jr)
    np.testing.assert_array_equal(Ud.cpu().numpy(), Uref)
    assert (jr == D).any()          # the u = 1.0 j_rand case occurred


def test_de_select_bit_exact(cuda, speo, oracle):
    import torch
    NP, D, ld = 300, 50, 52
    rng = np.random.default_rng(9)
    X = rng.uniform(-1, 1, (NP, ld)).astype(np.float32)
    U = rng.uniform(-1, 1, (NP, ld)).astype(np.float32)
    fX = rng.uniform(0, 1, NP).astype(np.float32)
    fU = rng.uniform(0, 1, NP).astype(np.float32)
    fU[::7] = fX[::7]                       # ties keep the parent
    fU[3::11] = np.nan
    fX[5::13] = np.nan
    Xd, fXd = _t(X, cuda), _t(fX, cuda)
    acc = torch.zeros(NP, dtype=torch.uint8, device=cuda)
    speo.de_select(Xd, fXd, _t(U, cuda), _t(fU, cuda), D, accept_out=acc)
    _sync()
    X2, fX2, acc2 = oracle.de_select(X, fX, U, fU, ld, NP, D)
    np.testing.assert_array_equal(Xd.cpu().numpy(), X2)
    np.testing.assert_array_equal(fXd.cpu().numpy(), fX2)
    np.testing.assert_array_equal(acc.cpu().numpy(), acc2)


def test_segment_argmin(cuda, speo):
    import torch
    rng = np.random.default_rng(2)
    f = rng.uniform(0, 1, 1000).astype(np.float32)
    f[100:120] = 0.001                      # tie: the lowest row wins
    f[500:600] = np.nan
    f[600] = np.nan
    off = np.array([0, 50, 130, 500, 601, 1000], np.int32)
    best = torch.zeros(5, dtype=torch.int32, device=cuda)
    speo.segment_argmin(_t(f, cuda), _t(off, cuda), best)
    _sync()
    expect = []
    for s in range(5):
        seg = f[off[s]:off[s + 1]]
        if np.isnan(seg).all():
            expect.append(int(off[s]))
        else:
            expect.append(int(off[s] + np.nanargmin(seg)))
    assert best.cpu().numpy().tolist() == expect


# ---------------------------------------------------------------- drop-in

def _pinfo(speo, D, seed, fid=1):
    from speo_de._lib import ProblemInfo
    return ProblemInfo(D, fid, 1, 10000, seed, 0, 0, 100.0, -100.0)


def test_dropin_generation_cycle(cuda, speo, oracle):
    """API_Initilize -> API_evaluateFitness -> API_DE_GenerateNewPopulation ->
    API_evaluateFitness -> API_DE_SelectSurvival in the reference's layout."""
    import torch
    from speo_de._lib import GPU_IslandInfo, GPU_Population, SPEO_RandState
    L = speo.lib()
    NP, D, subs = 100, 30, 4
    Dp = L.SPEO_NextPow2Dim(D)
    assert Dp == 32
    pop = torch.zeros((NP, Dp), dtype=torch.float32, device=cuda)
    fpop = torch.zeros(NP, dtype=torch.float32, device=cuda)
    cand = torch.zeros((NP, Dp), dtype=torch.float32, device=cuda)
    fcand = torch.zeros(NP, dtype=torch.float32, device=cuda)
    perm = np.random.default_rng(3).permutation(NP).astype(np.int32)
    permd = _t(perm, cuda)
    P = GPU_Population(pop.data_ptr(), fpop.data_ptr())
    C = GPU_Population(cand.data_ptr(), fcand.data_ptr())
    info = GPU_IslandInfo(NP, subs, permd.data_ptr())
    pi = _pinfo(speo, D, 4242, fid=1)
    rs = SPEO_RandState()
    assert L.API_Initilize(P, ctypes.byref(rs), info, pi) == 0
    # reference layouts: shift [Dp], M^T row-padded [Dp][D] (CEC2014_CUDA.cc:255)
    o, M = oracle.gen_problem(99, D)
    shift = np.zeros(Dp, np.float32)
    shift[:D] = o
    B = np.zeros((Dp, D), np.float32)
    B[:D, :] = M.T
    sd, Bd = _t(shift, cuda), _t(B, cuda)

    def evaluate(Xd, fd):
        L.API_evaluateFitness(ctypes.c_void_p(fd.data_ptr()), None, None,
                              ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(sd.data_ptr()),
                              ctypes.c_void_p(Bd.data_ptr()), None, None, None, None, NP, D,
                              0.0, 1)
    evaluate(pop, fpop)
    assert L.API_CheckCUDAError() == 0
    X0 = pop.cpu().numpy()
    np.testing.assert_array_equal(X0, oracle.init_population(NP, D, Dp, -100.0, 100.0, 4242))
    f0 = fpop.cpu().numpy()
    np.testing.assert_allclose(f0, oracle.evaluate(oracle.OBJ_ELLIPTIC, X0[:, :D], o, M),
                               rtol=1e-5)
    de = _de_info(speo, cuda, [0.9] * subs, [0.5] * subs)
    best = torch.zeros(subs, dtype=torch.int32, device=cuda)
    assert L.API_DE_GenerateNewPopulation(C, P, ctypes.c_void_p(best.data_ptr()), info,
                                          ctypes.c_void_p(de.data_ptr()), ctypes.byref(rs),
                                          pi) == 0
    assert rs.generation == 1
    evaluate(cand, fcand)
    assert L.API_CheckCUDAError() == 0
    U = cand.cpu().numpy()
    Uref, *_ = oracle.de_generate(X0, Dp, NP, D, subisland_num=subs, perm=perm, seed=4242,
                                  gen=0, U=np.zeros_like(X0))
    np.testing.assert_array_equal(U, Uref)
    # each sub-island's best slot (the reference intends this; its kernel
    # scans group 0 for every block, kernel_DE_CUDA.cu:133)
    S = NP // subs
    exp_best = [int(perm[s * S + np.argmin(f0[perm[s * S:(s + 1) * S]])]) for s in range(subs)]
    assert best.cpu().numpy().tolist() == exp_best
    fU = fcand.cpu().numpy()
    np.testing.assert_allclose(fU, oracle.evaluate(oracle.OBJ_ELLIPTIC, U[:, :D], o, M),
                               rtol=1e-5)
    assert L.API_DE_SelectSurvival(P, C, info, pi) == 0
    assert L.API_CheckCUDAError() == 0
    X2, f2, _ = oracle.de_select(X0, f0, U, fU, Dp, NP, D)
    np.testing.assert_array_equal(pop.cpu().numpy(), X2)
    np.testing.assert_array_equal(fpop.cpu().numpy(), f2)


@pytest.mark.parametrize("fid,obj,rate,rotate", [(1, 1, 1.0, True), (2, 4, 1.0, True),
                                                 (3, 5, 1.0, True), (5, 3, 1.0, True),
                                                 (8, 2, 0.0512, False), (9, 2, 0.0512, True),
                                                 (10, 8, 10.0, False), (11, 8, 10.0, True),
                                                 (13, 6, 0.05, True), (14, 7, 0.05, True)])
def test_dropin_evaluate_functions(cuda, speo, oracle, fid, obj, rate, rotate):
    import torch
    L = speo.lib()
    NP, D = 77, 100
    Dp = L.SPEO_NextPow2Dim(D)
    o, M = oracle.gen_problem(fid, D)
    Mu = M if rotate else np.eye(D)
    rng = np.random.default_rng(fid)
    X = np.zeros((NP, Dp), np.float32)
    X[:, :D] = rng.uniform(-100, 100, (NP, D)).astype(np.float32)
    shift = np.zeros(Dp, np.float32)
    shift[:D] = o
    B = np.zeros((Dp, D), np.float32)
    B[:D, :] = Mu.T
    f = torch.zeros(NP, dtype=torch.float32, device=cuda)
    Xd, sd, Bd = _t(X, cuda), _t(shift, cuda), _t(B, cuda)
    L.API_evaluateFitness(ctypes.c_void_p(f.data_ptr()), None, None, ctypes.c_void_p(Xd.data_ptr()),
                          ctypes.c_void_p(sd.data_ptr()), ctypes.c_void_p(Bd.data_ptr()), None,
                          None, None, None, NP, D, 0.0, fid)
    assert L.API_CheckCUDAError() == 0
    # z = rate * M (x - o): the oracle at x' = o + rate (x - o)
    xs = o[None, :] + rate * (X[:, :D].astype(np.float64) - o[None, :])
    ref = oracle.evaluate(obj, xs, o, Mu)
    keep = np.ones(NP, bool)
    if obj == 8:   # rows at a jump of Schwefel have no well-defined FP32 value
        keep = ~schwefel_rows(X[:, :D], o, Mu, rate)[1]
        assert keep.sum() >= NP - 2
    np.testing.assert_allclose(f.cpu().numpy()[keep], ref[keep], rtol=1e-5)


def test_dropin_unsupported_function_reports(cuda, speo):
    import torch
    L = speo.lib()
    f = torch.zeros(16, dtype=torch.float32, device=cuda)
    L.API_evaluateFitness(ctypes.c_void_p(f.data_ptr()), None, None,
                          ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                          ctypes.c_void_p(f.data_ptr()), None, None, None, None, 16, 10, 0.0, 31)
    assert L.API_CheckCUDAError() == -3
    assert "31" in speo.last_error()
    assert L.API_CheckCUDAError() == 0       # reported once
