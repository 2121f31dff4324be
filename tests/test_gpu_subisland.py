"""GPU parity of the engine's sub-islands, DE strategies 0-8 and the
reference's regroup / configure_EA options (plain mode): every generation's
trial vectors bit for bit against the oracle's DE step (kernel_DE_CUDA.cu:
172-289 restated) under the partition and DEInfo the oracle's regroup
restatement (migrate.cc:147-186, DE_CUDA.cc:89-119) predicts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(speo, oracle, D, objs):
    out = []
    for k, obj in enumerate(objs):
        o, M = oracle.gen_problem(600 + k, D)
        out.append(speo.Problem([obj], o[None, :], M[None, :, :]))
    return out


def _best_slots(f, perm, subs):
    S = perm.size // subs
    best = []
    for s in range(subs):
        m = perm[s * S:(s + 1) * S]
        best.append(int(min(m, key=lambda r: (bool(np.isnan(f[r])), f[r], r))))
    return np.array(best, np.int32)


def _check_generation(oracle, eng, X0, f0, U, seed, gen, D):
    for i in range(eng.num_islands):
        r0, r1 = int(eng.row_off[i]), int(eng.row_off[i + 1])
        subs, perm, info = eng.current[i]
        cr = np.array([t[0] for t in info], np.float32)
        f = np.array([t[1] for t in info], np.float32)
        st = np.array([t[2] for t in info], np.int32)
        best = _best_slots(f0[r0:r1], perm, subs)
        Xi = np.ascontiguousarray(X0[r0:r1])
        Uref, *_ = oracle.de_generate(Xi, D, r1 - r0, D, subisland_num=subs, perm=perm, cr=cr,
                                      f=f, strategy=st, best=best, seed=seed, gen=gen,
                                      row_base=r0, U=np.zeros_like(Xi))
        np.testing.assert_array_equal(U[r0:r1], Uref, err_msg=f"island {i} gen {gen}")


@pytest.mark.parametrize("strategy", range(9))
def test_every_strategy_with_sub_islands(cuda, speo, oracle, strategy):
    D, NP, seed = 10, 32, 4000 + strategy
    probs = _problems(speo, oracle, D, [speo.OBJ_RASTRIGIN])
    eng = speo.Engine(probs, [NP, NP], [0, 0], seed=seed, mode=speo.MODE_PLAIN,
                      rows_per_block=16, subisland_size=8, strategy=strategy)
    for g in range(3):
        eng.current = [eng.subislands(i) for i in range(2)]
        assert eng.current[0][0] == 4
        assert all(t[2] == strategy for t in eng.current[0][2])
        X0, f0 = eng.read()
        eng.run(1)
        U, fU = eng.read_trials()
        print(f"strategy {strategy} gen {g}", flush=True)
        _check_generation(oracle, eng, X0, f0, U, seed, g, D)
    eng.close()
    probs[0].close()


@pytest.mark.parametrize("option", [speo_opt for speo_opt in range(4)])
def test_regroups_and_dynamic_configuration_in_lockstep(cuda, speo, oracle, option):
    D, NP, seed, interval = 10, 64, 77 + option, 3
    probs = _problems(speo, oracle, D, [speo.OBJ_ELLIPTIC, speo.OBJ_ACKLEY])
    eng = speo.Engine(probs, [NP, NP, NP], [0, 1, 0], seed=seed, mode=speo.MODE_PLAIN,
                      rows_per_block=16, interval=interval, migration_size=2, subisland_size=16,
                      configure_EA=speo.CONFIGURE_DYNAMIC, regroup_option=option)
    base = (0.9, 0.5, 4)
    # the initial configuration: event 0 of every island
    expect = []
    for i in range(3):
        s, p, inf = oracle.island_regroup(seed, i, 0, option, 1, NP, NP // 16, list(range(NP)),
                                          base, None)
        expect.append((s, list(p), inf))
    events = 0
    for g in range(10):
        eng.current = [eng.subislands(i) for i in range(3)]
        for i in range(3):
            s, p, inf = eng.current[i]
            assert s == expect[i][0] and list(p) == expect[i][1], (option, g, i)
            assert np.allclose([t[:3] for t in inf], [t[:3] for t in expect[i][2]])
        X0, f0 = eng.read()
        eng.run(1)
        U, fU = eng.read_trials()
        print(f"regroup option {option} gen {g}", flush=True)
        _check_generation(oracle, eng, X0, f0, U, seed, g, D)
        if g % interval == 0:
            events += 1
            expect = [oracle.island_regroup(seed, i, events, option, 1, NP, expect[i][0],
                                            expect[i][1], base, expect[i][2]) for i in range(3)]
            expect = [(s, list(p), inf) for s, p, inf in expect]
    st = eng.stats()
    assert st["generations"] == 10 and st["migrations"] == 4
    eng.close()
    for p in probs:
        p.close()


def test_sub_island_options_need_plain_mode(cuda, speo, oracle):
    probs = _problems(speo, oracle, 10, [speo.OBJ_RASTRIGIN])
    with pytest.raises(speo.SpeoError, match="SPEO_MODE_PLAIN") as e:
        speo.Engine(probs, [32], [0], mode=speo.MODE_FUSED, subisland_size=8)
    assert e.value.code == -3
    with pytest.raises(speo.SpeoError, match="divide"):
        speo.Engine(probs, [36], [0], mode=speo.MODE_PLAIN, subisland_size=8)
    with pytest.raises(speo.SpeoError, match="4 \\* 2\\^k"):
        speo.Engine(probs, [48], [0], mode=speo.MODE_PLAIN,
                    regroup_option=speo.REGROUP_DYNAMIC_RANDOM)
    probs[0].close()
