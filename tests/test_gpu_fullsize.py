"""GPU parity at the benchmark workloads' full sizes.  The plain three-launch
engine (CUDA_EA_library/DE_CUDA.cc:200-216 per generation, checked against the
oracle step by step in test_gpu_engine.py) is the reference here: the fast
paths must reproduce it bit for bit on the real shapes, where the launch plans
differ from the small cases (islands split over 4 and 8 workgroups, 256 uneven
islands on the HBM plan and in fused tiles).  Size-independent properties are
checked on the way: every row inside the bounds, each island's best fitness
never worse than at the previous check (greedy selection,
kernel_DE_CUDA.cu:305-327, and migration replaces only the worst rows), and the
evaluation count of island_EA.cc:81/105."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 0xF011


def _round_robin(speo, oracle, D, objs, islands, base):
    probs = []
    for k, obj in enumerate(objs):
        o, M = oracle.gen_problem(base + k, D)
        probs.append(speo.Problem([obj], o[None, :], M[None, :, :]))
    return probs, [i % len(objs) for i in range(islands)]


def _assert_same(a, b):
    # the trials first: reading the population applies the fused engine's
    # pending selection, which clears the trial fitness
    Ua, fUa = a.read_trials()
    Ub, fUb = b.read_trials()
    np.testing.assert_array_equal(fUa, fUb)
    np.testing.assert_array_equal(Ua, Ub)
    Xa, fa = a.read()
    Xb, fb = b.read()
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(Xa, Xb)


def _island_best(e):
    _, f = e.read()
    return np.array([f[e.row_off[i]:e.row_off[i + 1]].min() for i in range(e.num_islands)])


def _run_pair(speo, plain, fast, chunks, lo=-100.0, hi=100.0):
    prev = _island_best(plain)
    for k, c in enumerate(chunks):
        print(f"chunk {c}", flush=True)
        plain.run(c)
        fast.run(c)
        if k + 1 < len(chunks):
            best = _island_best(plain)
            assert (best <= prev).all()
            prev = best
    _assert_same(plain, fast)
    assert (_island_best(plain) <= prev).all()
    X, _ = fast.read()
    assert np.isfinite(X).all() and (X >= lo).all() and (X <= hi).all()


def test_headline_workload_epoch_equals_plain(cuda, speo, oracle):
    # 64 islands x NP=128, D=100, elliptic / Rastrigin / Ackley round-robin,
    # ring best-2 every 50 generations; epoch mode splits every island over 4
    # workgroups on the 256 compute units
    probs, pidx = _round_robin(speo, oracle, 100, [1, 2, 3], 64, 4100)
    kw = dict(F=0.5, CR=0.9, seed=SEED, interval=50, migration_size=2, use_graph=True)
    plain = speo.Engine(probs, [128] * 64, pidx, mode=speo.MODE_PLAIN, **kw)
    fast = speo.Engine(probs, [128] * 64, pidx, mode=speo.MODE_EPOCH, epoch_parts=4, **kw)
    assert fast.stats()["epoch_parts"] == 4
    _run_pair(speo, plain, fast, [1, 40, 60, 19])
    st = fast.stats()
    assert st["evaluations"] == 64 * 128 * 121 and st["migrations"] == 3


@pytest.mark.parametrize("prec", [0, 3])
def test_time_to_error_workload_epoch_equals_plain(cuda, speo, oracle, prec):
    # 32 islands x NP=256, D=100, Rastrigin / elliptic: islands split over 8
    # workgroups, FP32 and BF16x3
    probs, pidx = _round_robin(speo, oracle, 100, [2, 1], 32, 4200)
    kw = dict(F=0.5, CR=0.9, seed=SEED, interval=50, migration_size=2, use_graph=True,
              precision=prec)
    plain = speo.Engine(probs, [256] * 32, pidx, mode=speo.MODE_PLAIN, **kw)
    fast = speo.Engine(probs, [256] * 32, pidx, mode=speo.MODE_EPOCH, epoch_parts=8, **kw)
    assert fast.stats()["epoch_parts"] == 8
    _run_pair(speo, plain, fast, [1, 49, 30])


@pytest.mark.parametrize("mode", ["fused", "epoch"])
def test_uneven_batch_equals_plain(cuda, speo, oracle, mode):
    # 256 islands, D in {10, 30, 50, 100}, NP uniform in [50, 512], migration
    # every 25: the fused kernel's autotuned row tiles and the epoch kernel's
    # LDS and HBM plans against the plain engine
    rng = np.random.default_rng(4300)
    dims = rng.choice([10, 30, 50, 100], 256)
    sizes = rng.integers(50, 513, 256).tolist()
    keys = sorted({(int(d), i % 3) for i, d in enumerate(dims)})
    probs, index = [], {}
    for k, (d, o_i) in enumerate(keys):
        o, M = oracle.gen_problem(4400 + k, d)
        index[(d, o_i)] = len(probs)
        probs.append(speo.Problem([o_i + 1], o[None, :], M[None, :, :]))
    pidx = [index[(int(d), i % 3)] for i, d in enumerate(dims)]
    kw = dict(F=0.5, CR=0.9, seed=SEED, interval=25, migration_size=2, use_graph=True)
    plain = speo.Engine(probs, sizes, pidx, mode=speo.MODE_PLAIN, rows_per_block=32, **kw)
    fast = speo.Engine(probs, sizes, pidx, rows_per_block=0,
                       mode=speo.MODE_FUSED if mode == "fused" else speo.MODE_EPOCH, **kw)
    # (45 generations: the run ends between two migrations, so the last
    # trials are still there to compare)
    _run_pair(speo, plain, fast, [1, 24, 20])
