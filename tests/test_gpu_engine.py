"""GPU parity of the island engine (IslandEA::Execute / DE_CUDA::Run semantics,
src/island_EA.cc:78-113, CUDA_EA_library/DE_CUDA.cc:200-216) against the
oracle: initial population, trial vectors and selection bit-exact, fitness
within 1e-5 of FP64 (the per-row bound of tolerance.py for ill-conditioned
weighted-sphere rows), ring migration bit-exact, the fused one-launch generation
equal to the plain three-launch one, and graph replays equal to eager runs."""
import numpy as np
import pytest

from tolerance import assert_rows, row_rtol, schwefel_rows

pytestmark = pytest.mark.gpu

SEED = 0x5EED


def _problems(speo, oracle, dims_objs):
    probs = []
    for k, (D, obj) in enumerate(dims_objs):
        o, M = oracle.gen_problem(1000 + k, D)
        probs.append(speo.Problem([obj], o[None, :], M[None, :, :]))
    return probs


def _engine(speo, probs, sizes, pidx, **kw):
    args = dict(F=0.5, CR=0.9, seed=SEED, rows_per_block=32)
    args.update(kw)
    return speo.Engine(probs, sizes, pidx, **args)


LAYOUT = dict(sizes=[16, 24, 40, 20], pidx=[0, 0, 1, 0])
DIMS = [(10, 2), (30, 3)]


def test_plain_generation_matches_oracle(cuda, speo, oracle):
    probs = _problems(speo, oracle, DIMS)
    e = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], mode=speo.MODE_PLAIN)
    X0, f0 = e.read()
    e.run(1)
    U, fU = e.read_trials()
    X1, f1 = e.read()
    off = e.row_off
    for i, (np_i, p) in enumerate(zip(LAYOUT["sizes"], LAYOUT["pidx"])):
        D = DIMS[p][0]
        obj = DIMS[p][1]
        rows = slice(off[i], off[i + 1])
        ld = e.max_dim
        Xi = np.ascontiguousarray(X0[rows])
        ref0 = oracle.init_population(np_i, D, ld, -100.0, 100.0, SEED, int(off[i]))
        np.testing.assert_array_equal(Xi[:, :D], ref0[:, :D])
        o = probs[p].shifts[0]
        M = probs[p].rotations[0]
        np.testing.assert_allclose(f0[rows], oracle.evaluate(obj, Xi[:, :D], o, M), rtol=1e-5)
        Uref, *_ = oracle.de_generate(Xi, ld, np_i, D, seed=SEED, gen=0, row_base=int(off[i]),
                                      U=np.zeros_like(Xi))
        np.testing.assert_array_equal(U[rows][:, :D], Uref[:, :D])
        np.testing.assert_allclose(fU[rows], oracle.evaluate(obj, Uref[:, :D], o, M), rtol=1e-5)
        X2, f2, _ = oracle.de_select(Xi, f0[rows], np.ascontiguousarray(U[rows]), fU[rows], ld,
                                     np_i, D)
        np.testing.assert_array_equal(X1[rows], X2)
        np.testing.assert_array_equal(f1[rows], f2)
    st = e.stats()
    assert st["generations"] == 1 and st["evaluations"] == 2 * sum(LAYOUT["sizes"])


@pytest.mark.parametrize("rt", [16, 32, 64])
def test_fused_equals_plain(cuda, speo, oracle, rt):
    probs = _problems(speo, oracle, DIMS)
    a = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], mode=speo.MODE_PLAIN,
                rows_per_block=rt)
    b = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], mode=speo.MODE_FUSED,
                rows_per_block=rt)
    for chunk in (1, 2, 5):
        a.run(chunk)
        b.run(chunk)
    Ua, fUa = a.read_trials()
    Ub, fUb = b.read_trials()
    np.testing.assert_array_equal(Ua, Ub)
    np.testing.assert_array_equal(fUa, fUb)
    Xa, fa = a.read()
    Xb, fb = b.read()
    np.testing.assert_array_equal(Xa, Xb)
    np.testing.assert_array_equal(fa, fb)


def test_weighted_sphere_objectives_all_modes(cuda, speo, oracle):
    # bent cigar (4) and discus (5) run through the kernels' weighted-sphere
    # path with their own column weights: the first generation's trials
    # against the FP64 oracle (no migration, which in fused mode clears the
    # trial fitness), then plain, fused, epoch (whole epochs and split
    # islands) bit-identical with migration; D = 200 exercises the large-D
    # path (plain and fused only).
    dims_objs = [(30, 4), (100, 5), (100, 4)]
    probs = _problems(speo, oracle, dims_objs)
    sizes, pidx = [32, 48, 64, 40], [0, 1, 2, 1]
    first = _engine(speo, probs, sizes, pidx, mode=speo.MODE_PLAIN)
    first.run(1)
    U, fU = first.read_trials()
    off = first.row_off
    for i, p in enumerate(pidx):
        D, obj = dims_objs[p]
        rows = slice(off[i], off[i + 1])
        o, M = probs[p].shifts[0], probs[p].rotations[0]
        ref = oracle.evaluate(obj, U[rows][:, :D], o, M)
        assert_rows(fU[rows], ref, row_rtol(obj, U[rows][:, :D], o, M))
    kw = dict(interval=4, migration_size=2)
    plain = _engine(speo, probs, sizes, pidx, mode=speo.MODE_PLAIN, **kw)
    others = [_engine(speo, probs, sizes, pidx, mode=speo.MODE_FUSED, **kw),
              _engine(speo, probs, sizes, pidx, mode=speo.MODE_EPOCH, epoch_parts=1, **kw),
              _engine(speo, probs, sizes, pidx, mode=speo.MODE_EPOCH, epoch_parts=2, **kw)]
    for chunk in (1, 3, 8):
        print(f"weighted sphere chunk {chunk}", flush=True)
        plain.run(chunk)
        for e in others:
            e.run(chunk)
    Xa, fa = plain.read()
    for e in others:
        Xb, fb = e.read()
        np.testing.assert_array_equal(fa, fb)
        np.testing.assert_array_equal(Xa, Xb)
    big = _problems(speo, oracle, [(200, 5), (200, 4)])
    a = _engine(speo, big, [64, 48], [0, 1], mode=speo.MODE_PLAIN)
    b = _engine(speo, big, [64, 48], [0, 1], mode=speo.MODE_FUSED)
    a.run(1)
    b.run(1)
    U, fU = a.read_trials()
    Ub, fUb = b.read_trials()
    np.testing.assert_array_equal(fU, fUb)
    np.testing.assert_array_equal(U, Ub)
    for i, p in enumerate([0, 1]):
        rows = slice(a.row_off[i], a.row_off[i + 1])
        o, M = big[p].shifts[0], big[p].rotations[0]
        ref = oracle.evaluate([5, 4][p], U[rows][:, :200], o, M)
        assert_rows(fU[rows], ref, row_rtol([5, 4][p], U[rows][:, :200], o, M))
    a.run(6)
    b.run(6)
    np.testing.assert_array_equal(a.read()[0], b.read()[0])


def test_happycat_hgbat_plain_and_fused(cuda, speo, oracle):
    # HappyCat (6) and HGBat (7): first-generation trials against the FP64
    # oracle, plain and fused bit-identical over migrating runs (D = 40 and
    # the large-D path at D = 160); epoch mode declines them.
    dims_objs = [(40, 6), (40, 7), (160, 6), (160, 7)]
    probs = _problems(speo, oracle, dims_objs)
    sizes, pidx = [32, 40, 48, 24], [0, 1, 2, 3]
    first = _engine(speo, probs, sizes, pidx, mode=speo.MODE_PLAIN)
    first.run(1)
    U, fU = first.read_trials()
    for i, p in enumerate(pidx):
        D, obj = dims_objs[p]
        rows = slice(first.row_off[i], first.row_off[i + 1])
        ref = oracle.evaluate(obj, U[rows][:, :D], probs[p].shifts[0], probs[p].rotations[0])
        np.testing.assert_allclose(fU[rows], ref, rtol=1e-5)
    kw = dict(interval=3, migration_size=2)
    a = _engine(speo, probs, sizes, pidx, mode=speo.MODE_PLAIN, **kw)
    b = _engine(speo, probs, sizes, pidx, mode=speo.MODE_FUSED, **kw)
    for chunk in (1, 4, 9):
        a.run(chunk)
        b.run(chunk)
    Xa, fa = a.read()
    Xb, fb = b.read()
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(Xa, Xb)
    with pytest.raises(speo.SpeoError, match="objectives 1-5"):
        _engine(speo, probs[:1], [32], [0], mode=speo.MODE_EPOCH)


def test_schwefel_plain_and_fused(cuda, speo, oracle):
    # Schwefel (8) at rate 10, D = 50 and the large-D path at D = 200:
    # first-generation trials against the FP64 oracle (rows at a jump of the
    # function left out), plain and fused bit-identical over migrating runs.
    probs, dims = [], [50, 200]
    for k, D in enumerate(dims):
        o, M = oracle.gen_problem(1900 + k, D)
        probs.append(speo.Problem([8], o[None, :], M[None, :, :], rate=[10.0]))
    sizes, pidx = [40, 64, 48], [0, 1, 0]
    first = _engine(speo, probs, sizes, pidx, mode=speo.MODE_PLAIN)
    first.run(1)
    U, fU = first.read_trials()
    for i, p in enumerate(pidx):
        D = dims[p]
        rows = slice(first.row_off[i], first.row_off[i + 1])
        o, M = probs[p].shifts[0], probs[p].rotations[0]
        Ui = U[rows][:, :D].astype(np.float64)
        ref = oracle.evaluate(8, o[None, :] + 10.0 * (Ui - o[None, :]), o, M)
        keep = ~schwefel_rows(Ui, o, M, 10.0)[1]
        np.testing.assert_allclose(fU[rows][keep], ref[keep], rtol=1e-5)
    kw = dict(interval=3, migration_size=2)
    a = _engine(speo, probs, sizes, pidx, mode=speo.MODE_PLAIN, **kw)
    b = _engine(speo, probs, sizes, pidx, mode=speo.MODE_FUSED, **kw)
    for chunk in (2, 7):
        a.run(chunk)
        b.run(chunk)
    Xa, fa = a.read()
    Xb, fb = b.read()
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(Xa, Xb)


@pytest.mark.parametrize("mode", [0, 1])
def test_migration_matches_oracle(cuda, speo, oracle, mode):
    probs = _problems(speo, oracle, DIMS)
    k = 2
    base = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], mode=mode)
    mig = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], mode=mode, interval=3,
                  migration_size=k)
    base.run(1)      # generation 0, then migration (0 % 3 == 0) in `mig`
    mig.run(1)
    X, f = base.read()
    Xm, fm = mig.read()
    off = base.row_off.astype(np.int32)
    dims = np.array([DIMS[p][0] for p in LAYOUT["pidx"]], np.int32)
    # ring per problem: 0 -> 1 -> 3 -> 0 (problem 0), island 2 alone
    dst = np.array([1, 3, -1, 0], np.int32)
    Xo, fo = oracle.migrate_ring(X, f, X.shape[1], off, dims, dst, k)
    np.testing.assert_array_equal(Xm, Xo)
    np.testing.assert_array_equal(fm, fo)
    assert mig.stats()["migrations"] == 1
    # generations 1, 2 do not migrate; 3 does
    base.run(3)
    mig.run(2)
    assert mig.stats()["migrations"] == 1
    mig.run(1)
    assert mig.stats()["migrations"] == 2


@pytest.mark.parametrize("mode", [0, 1])
def test_graph_replay_equals_eager(cuda, speo, oracle, mode):
    probs = _problems(speo, oracle, DIMS)
    kw = dict(mode=mode, interval=7, migration_size=2)
    eager = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], use_graph=False, **kw)
    graph = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], use_graph=True, **kw)
    eager.run(40)
    for chunk in (1, 13, 7, 19):          # several replays of the same graphs
        graph.run(chunk)
        _ = graph.stats()
    Xe, fe = eager.read()
    Xg, fg = graph.read()
    np.testing.assert_array_equal(Xe, Xg)
    np.testing.assert_array_equal(fe, fg)
    st = graph.stats()
    assert st["graph_launches"] >= 6 and st["generations"] == 40
    assert st["migrations"] == eager.stats()["migrations"] == 6   # after 0, 7, ..., 35


@pytest.mark.parametrize("mode", ["fused", "epoch"])
def test_prepare_then_run_equals_eager(cuda, speo, oracle, mode):
    # speo_engine_prepare captures the graphs of the next run without running
    # anything: the generation count and the populations are unchanged, and
    # the run that follows gives the eager result bit for bit.
    m = {"fused": speo.MODE_FUSED, "epoch": speo.MODE_EPOCH}[mode]
    probs = _problems(speo, oracle, DIMS)
    kw = dict(mode=m, interval=7, migration_size=2)
    eager = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], use_graph=False, **kw)
    graph = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], use_graph=True, **kw)
    eager.run(5)
    graph.run(5)
    X0, f0 = graph.read()
    graph.prepare(20)
    st = graph.stats()
    assert st["generations"] == 5
    X1, f1 = graph.read()
    np.testing.assert_array_equal(X0, X1)
    np.testing.assert_array_equal(f0, f1)
    eager.run(20)
    graph.run(20)
    Xe, fe = eager.read()
    Xg, fg = graph.read()
    np.testing.assert_array_equal(Xe, Xg)
    np.testing.assert_array_equal(fe, fg)
    assert graph.stats()["generations"] == 25


def test_best_reverified_in_fp64(cuda, speo, oracle):
    probs = _problems(speo, oracle, DIMS)
    e = _engine(speo, probs, LAYOUT["sizes"], LAYOUT["pidx"], mode=speo.MODE_FUSED)
    e.run(25)
    rows, f32, f64 = e.best()
    X, f = e.read()
    for i, p in enumerate(LAYOUT["pidx"]):
        seg = f[e.row_off[i]:e.row_off[i + 1]]
        assert rows[i] == e.row_off[i] + int(np.argmin(seg))
        assert f32[i] == seg.min()
        D, obj = DIMS[p]
        ref = oracle.evaluate(obj, X[rows[i], :D][None, :], probs[p].shifts[0],
                              probs[p].rotations[0])[0]
        assert f64[i] == pytest.approx(ref, rel=1e-12)
        assert f32[i] == pytest.approx(ref, rel=1e-5)


def test_uneven_islands_and_precisions(cuda, speo, oracle):
    dims_objs = [(10, 1), (30, 2), (50, 3), (100, 2)]
    probs = _problems(speo, oracle, dims_objs)
    rng = np.random.default_rng(4)
    sizes = rng.integers(50, 200, 9).tolist()
    pidx = [i % 4 for i in range(9)]
    for prec in (speo.PREC_FP32, speo.PREC_BF16X3, speo.PREC_BF16X6):
        e = _engine(speo, probs, sizes, pidx, precision=prec, interval=5, migration_size=2,
                    use_graph=True, rows_per_block=0)
        _, f0, _ = e.best()
        e.run(30)
        rows, f32, f64 = e.best()
        # the FP32 best never increases; its FP64 value may move by the FP32
        # evaluation error of the split precisions
        assert (f64 <= f0 * (1 + 1e-3) + 1e-9).all()
        assert np.isfinite(f64).all()
        st = e.stats()
        assert st["evaluations"] == sum(sizes) * 31
        assert st["rows_per_block"] in (16, 32, 64)


def test_convergence_on_small_problem(cuda, speo, oracle):
    probs = _problems(speo, oracle, [(10, 1)])
    e = _engine(speo, probs, [64, 64], [0, 0], interval=20, migration_size=2, use_graph=True)
    _, _, f_start = e.best()
    e.run(600)
    _, _, f_end = e.best()
    assert f_end.min() < 1e-4 * f_start.min()


@pytest.mark.parametrize("case", ["no_islands", "small_island", "problem_index", "bounds", "mode",
                                  "precision", "rows_per_block", "migration_size", "epoch_parts",
                                  "surrogate_dims", "surrogate_mode", "run_negative"])
def test_engine_rejects_bad_configs(cuda, speo, oracle, case):
    # every rejected configuration returns SPEO_ERR_ARG with a message naming
    # the problem, before any device work
    probs = _problems(speo, oracle, DIMS)
    sizes, pidx, kw = [16, 24], [0, 1], {}
    match = "speo"
    if case == "no_islands":
        sizes, pidx, match = [], [], "islands and problems are required"
    elif case == "small_island":
        sizes, match = [7, 24], "island size must be >= 8"
    elif case == "problem_index":
        pidx, match = [0, 2], "problem index in range"
    elif case == "bounds":
        kw, match = dict(lo=5.0, hi=5.0), "bad migration or bounds"
    elif case == "mode":
        kw, match = dict(mode=7), "unknown mode"
    elif case == "precision":
        kw, match = dict(precision=9), "unknown precision"
    elif case == "rows_per_block":
        kw, match = dict(rows_per_block=48), "rows_per_block must be"
    elif case == "migration_size":
        kw, match = dict(migration_size=9, interval=5), ">= 2 \\* migration_size"
    elif case == "epoch_parts":
        kw, match = dict(mode=speo.MODE_EPOCH, epoch_parts=3), "epoch_parts must be"
    elif case == "surrogate_dims":
        kw, match = dict(surrogate_candidates=4, mode=speo.MODE_PLAIN), "share one dimension"
    elif case == "surrogate_mode":
        kw, match = dict(surrogate_candidates=4, mode=speo.MODE_FUSED), "surrogate runs in SPEO_MODE_PLAIN"
    if case == "run_negative":
        e = _engine(speo, probs, sizes, pidx)
        with pytest.raises(speo.SpeoError, match="bad arguments"):
            e.run(-1)
        e.run(0)
        return
    with pytest.raises(speo.SpeoError, match=match):
        _engine(speo, probs, sizes, pidx, **kw)


def test_run_fes_budget_loop(cuda, speo, oracle):
    # IslandEA::Execute (src/island_EA.cc:86-104): generations while the
    # evaluation count (initial population included) is below the budget,
    # migration at generation % interval == 0; then the record row
    probs = _problems(speo, oracle, DIMS)
    sizes, pidx = [16, 24], [0, 1]
    e = _engine(speo, probs, sizes, pidx, interval=5, migration_size=2, use_graph=True)
    g = e.run_fes(1000)
    st = e.stats()
    assert g == 24 == st["generations"]           # 40 rows: 40 + 24 * 40 = 1000
    assert st["evaluations"] == 1000 and st["migrations"] == 5   # generations 0, 5, ..., 20
    assert e.run_fes(1000) == 0 and e.stats()["generations"] == 24
    assert e.run_fes(1001) == 1 and e.stats()["evaluations"] == 1040
    _, _, f64 = e.best()
    row = speo.result_record(function_ID=9, run_ID=1, dim=10, best_fitness=float(f64.min()),
                             time_period=0.0, communication_percentage=0.0, total_FEs=1001,
                             node_num=1, GPU_num=1, island_num=2, total_pop_size=40,
                             subisland_size=-1, interval=5, migration_size=2, migration_rate=0.5,
                             buffer_capacity=32, migration_topology="ring", buffer_manage="none",
                             regroup_option="static_and_ordered", configure_EA="constant",
                             EA_parameters=speo.de_parameters(0.9, 0.5, 4))
    cols = row.split(",")
    assert len(cols) == 21 and float(cols[3]) == pytest.approx(float(f64.min()), rel=1e-5)
