"""GPU parity of the epoch kernel (one workgroup per island running every
generation between two migrations, the loop of src/island_EA.cc:84-99 around
CUDA_EA_library/DE_CUDA.cc:200-216) against the plain three-launch engine,
whose generate/evaluate/select steps are checked against the oracle in
test_gpu_engine.py: populations, fitness, the last trials and their fitness
must be bit-identical, for islands whose rows stay in LDS and for islands
too large for that (rows in HBM, trials staged per pass)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 0xE90C


def _problems(speo, oracle, dims_objs, base=2000):
    probs = []
    for k, (D, obj) in enumerate(dims_objs):
        o, M = oracle.gen_problem(base + k, D)
        probs.append(speo.Problem([obj], o[None, :], M[None, :, :]))
    return probs


def _pair(speo, probs, sizes, pidx, **kw):
    args = dict(F=0.5, CR=0.9, seed=SEED, rows_per_block=32, interval=7, migration_size=2)
    args.update(kw)
    plain = speo.Engine(probs, sizes, pidx, mode=speo.MODE_PLAIN, **args)
    epoch = speo.Engine(probs, sizes, pidx, mode=speo.MODE_EPOCH, **args)
    return plain, epoch


def _assert_same(a, b):
    Xa, fa = a.read()
    Xb, fb = b.read()
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(Xa, Xb)
    Ua, fUa = a.read_trials()
    Ub, fUb = b.read_trials()
    np.testing.assert_array_equal(fUa, fUb)
    np.testing.assert_array_equal(Ua, Ub)
    sa, sb = a.stats(), b.stats()
    for k in ("generations", "evaluations", "migrations"):
        assert sa[k] == sb[k], k


def test_epoch_equals_plain_small(cuda, speo, oracle):
    probs = _problems(speo, oracle, [(10, 2), (30, 3), (3, 1)])
    sizes, pidx = [16, 24, 40, 20, 12], [0, 0, 1, 0, 2]
    a, b = _pair(speo, probs, sizes, pidx)
    for chunk in (1, 2, 5, 13):
        print(f"chunk {chunk}", flush=True)
        a.run(chunk)
        b.run(chunk)
        _assert_same(a, b)


@pytest.mark.parametrize("obj", [1, 2, 3, 4, 5])
def test_epoch_equals_plain_d100(cuda, speo, oracle, obj):
    # NP=128 at D=100 keeps both slots in LDS; NP=200 and 333 do not (HBM slots,
    # trials staged in passes of rows); NP=17 leaves a ragged last row tile.
    probs = _problems(speo, oracle, [(100, obj)], base=2100 + obj)
    sizes, pidx = [128, 200, 17, 333], [0, 0, 0, 0]
    a, b = _pair(speo, probs, sizes, pidx, interval=5)
    for chunk in (3, 9):
        print(f"objective {obj} chunk {chunk}", flush=True)
        a.run(chunk)
        b.run(chunk)
    _assert_same(a, b)


def test_epoch_uneven_batch(cuda, speo, oracle):
    dims = [(10, 1), (30, 2), (50, 3), (100, 1), (64, 2), (128, 3)]
    probs = _problems(speo, oracle, dims, base=2200)
    rng = np.random.default_rng(5)
    n = 24
    pidx = [i % len(dims) for i in range(n)]
    sizes = [int(v) for v in rng.integers(50, 513, n)]
    a, b = _pair(speo, probs, sizes, pidx, interval=4, use_graph=True)
    a.run(6)
    b.run(6)
    _assert_same(a, b)


@pytest.mark.parametrize("prec", [1, 2, 3, 4, 5])
def test_epoch_precisions_equal_plain(cuda, speo, oracle, prec):
    probs = _problems(speo, oracle, [(100, 2), (30, 1)], base=2300)
    sizes, pidx = [64, 48, 160, 40], [0, 1, 0, 1]
    a, b = _pair(speo, probs, sizes, pidx, precision=prec)
    a.run(8)
    b.run(8)
    _assert_same(a, b)


def test_epoch_graph_replay_equals_eager(cuda, speo, oracle):
    probs = _problems(speo, oracle, [(30, 3), (100, 2)], base=2400)
    sizes, pidx = [40, 72, 40, 72], [0, 1, 0, 1]
    for parts in (1, 2):   # whole epochs per launch; split islands, a launch per generation
        kw = dict(F=0.5, CR=0.9, seed=SEED, rows_per_block=32, interval=6, migration_size=2,
                  mode=speo.MODE_EPOCH, epoch_parts=parts)
        eager = speo.Engine(probs, sizes, pidx, use_graph=False, **kw)
        graph = speo.Engine(probs, sizes, pidx, use_graph=True, **kw)
        eager.run(37)
        for chunk in (1, 11, 6, 19):
            graph.run(chunk)
            _ = graph.stats()
        _assert_same(eager, graph)
        assert graph.stats()["graph_launches"] >= 6


def test_epoch_profile_reports_epoch_time(cuda, speo, oracle):
    probs = _problems(speo, oracle, [(30, 2)], base=2500)
    e = speo.Engine(probs, [64, 64], [0, 0], mode=speo.MODE_EPOCH, interval=10, migration_size=2,
                    seed=SEED)
    kms = e.profile(25)
    assert kms["epoch"] > 0 and kms["donors"] > 0 and kms["fused"] == 0
    assert e.stats()["generations"] == 25 and e.stats()["migrations"] == 3


def test_epoch_rejects_unsupported(cuda, speo, oracle):
    o, M = oracle.gen_problem(2600, 200)
    big = speo.Problem([2], o[None, :], M[None, :, :])
    with pytest.raises(speo.SpeoError, match="dim <= 128"):
        speo.Engine([big], [32, 32], [0, 0], mode=speo.MODE_EPOCH)


@pytest.mark.parametrize("parts", [2, 4, 8])
def test_epoch_split_islands_equal_plain(cuda, speo, oracle, parts):
    # islands split over `parts` workgroups, one generation per launch, each
    # workgroup loading the whole island and writing its share of the rows;
    # NP=136 leaves a short share (and, at 8 parts, empty ones)
    probs = _problems(speo, oracle, [(100, 2), (30, 3), (64, 1)], base=2700)
    sizes, pidx = [128, 136, 160, 144, 128], [0, 1, 2, 0, 1]
    kw = dict(F=0.5, CR=0.9, seed=SEED, rows_per_block=32, interval=6, migration_size=2)
    plain = speo.Engine(probs, sizes, pidx, mode=speo.MODE_PLAIN, **kw)
    split = speo.Engine(probs, sizes, pidx, mode=speo.MODE_EPOCH, epoch_parts=parts, **kw)
    assert split.stats()["epoch_parts"] == parts
    for chunk in (1, 7, 13):
        print(f"parts {parts} chunk {chunk}", flush=True)
        plain.run(chunk)
        split.run(chunk)
    _assert_same(plain, split)


@pytest.mark.parametrize("prec", [0, 3])
def test_epoch_split_edge_dims_equal_plain(cuda, speo, oracle, prec):
    # split islands at the edges of the row layouts: D = 128 (every K block and
    # column tile in use, LDS stride 132), D = 3 (one float4 group per row),
    # D = 96 (stride padded to 100), D = 10 (ragged last group)
    probs = _problems(speo, oracle, [(128, 1), (3, 2), (96, 3), (10, 2)], base=2900)
    sizes, pidx = [96, 64, 80, 72, 64, 96], [0, 1, 2, 3, 0, 2]
    kw = dict(F=0.5, CR=0.9, seed=SEED, rows_per_block=32, interval=5, migration_size=2, precision=prec)
    plain = speo.Engine(probs, sizes, pidx, mode=speo.MODE_PLAIN, **kw)
    split = speo.Engine(probs, sizes, pidx, mode=speo.MODE_EPOCH, epoch_parts=2, **kw)
    assert split.stats()["epoch_parts"] == 2
    for chunk in (1, 6, 4):
        print(f"precision {prec} chunk {chunk}", flush=True)
        plain.run(chunk)
        split.run(chunk)
    _assert_same(plain, split)


def test_epoch_automatic_split(cuda, speo, oracle, tmp_path, monkeypatch):
    # epoch_parts = 0: the tuner times every feasible split on throwaway
    # engines, keeps the fastest and persists it; a second engine of the same
    # shape reads the choice back
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SPEO_TUNE_CACHE", str(cache))
    probs = _problems(speo, oracle, [(100, 1)], base=2800)
    e = speo.Engine(probs, [128] * 8, [0] * 8, mode=speo.MODE_EPOCH, interval=10, migration_size=2,
                    seed=SEED)
    chosen = e.stats()["epoch_parts"]
    assert chosen in (1, 2, 4, 8)
    lines = cache.read_text().split()
    assert lines[0].startswith("epoch_parts:m2:p0:n8:100/128*8") and int(lines[1]) == chosen
    cache.write_text(lines[0] + " 2\n")          # a stored choice is used as it is
    e2 = speo.Engine(probs, [128] * 8, [0] * 8, mode=speo.MODE_EPOCH, interval=10,
                     migration_size=2, seed=SEED)
    assert e2.stats()["epoch_parts"] == 2
    e2.close()
    one = speo.Engine(probs, [128] * 8, [0] * 8, mode=speo.MODE_EPOCH, interval=10,
                      migration_size=2, seed=SEED, epoch_parts=1)
    e.run(25)
    one.run(25)
    _assert_same(one, e)
