"""GPU tests of the island driver: the function x run sweep with the result
file (src/main.cc:124-150, src/island_EA.cc:13-171) and the measured budget
loop (speo_engine_run_fes_timed)."""
import csv

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["plain", "fused", "epoch"])
def test_timed_budget_loop_equals_untimed(cuda, speo, oracle, mode):
    # migrations taken out of the chunks and run between synchronisations
    # give the same populations as the untimed budget loop
    m = {"plain": speo.MODE_PLAIN, "fused": speo.MODE_FUSED, "epoch": speo.MODE_EPOCH}[mode]
    D = 20
    o, M = oracle.gen_problem(3, D)
    prob = speo.Problem([speo.OBJ_RASTRIGIN], o[None, :], M[None, :, :])
    kw = dict(seed=5, mode=m, interval=5, migration_size=2, use_graph=(mode != "plain"),
              rows_per_block=16)
    a = speo.Engine([prob], [32] * 4, [0] * 4, **kw)
    b = speo.Engine([prob], [32] * 4, [0] * 4, **kw)
    ga = a.run_fes(32 * 4 * 24)
    gb, secs, comm = b.run_fes_timed(32 * 4 * 24)
    assert ga == gb == 23
    assert secs > 0.0 and 0.0 <= comm <= secs
    Xa, fa = a.read()
    Xb, fb = b.read()
    np.testing.assert_array_equal(Xa, Xb)
    np.testing.assert_array_equal(fa, fb)
    assert a.stats()["migrations"] == b.stats()["migrations"] == 5   # after 0, 5, 10, 15, 20
    a.close()
    b.close()
    prob.close()


def test_driver_sweep_and_record_file(cuda, speo, oracle, tmp_path):
    D, fids, runs = 10, [1, 17, 23], [1, 2]
    data = {}
    for fid in fids:
        o, M, S = oracle.gen_cec_problem(700 + fid, fid, D)
        oracle.write_cec_files(tmp_path, fid, o, M, S)
        data[fid] = (o, M, S)
    rec = tmp_path / "Results.csv"
    rows = speo.island_driver(tmp_path / "input_data", fids, runs, D, island_num=4,
                              island_size=32, max_base_FEs=100, record_file=rec, node_ID=1,
                              interval=3, migration_size=2)
    lines = list(csv.reader(open(rec)))
    assert ",".join(lines[0]) == speo.result_header()
    assert len(lines) == 1 + len(fids) * len(runs)
    for k, (fid, run) in enumerate([(f, r) for f in fids for r in runs]):
        row = rows[k]
        assert (row["function_ID"], row["run_ID"], row["dim"]) == (fid, run, D)
        assert row["total_FEs"] == 100 * D and row["total_pop_size"] == 128
        assert row["subisland_size"] == 32 and row["island_num"] == 4
        assert row["EA_parameters"] == "CR/F=0.9/0.5_rand/1/bin"
        assert row["time_period"] > 0 and 0 <= row["communication_percentage"] <= 1
        assert lines[1 + k][0] == str(fid) and lines[1 + k][1] == str(run)
        # the same run by hand: seed (node + f) (run + f), budget max_base_FEs * dim
        p = speo.Problem.from_cec(tmp_path / "input_data", fid, D)
        e = speo.Engine([p], [32] * 4, [0] * 4, seed=speo.run_seed(1, fid, run), interval=3,
                        migration_size=2, mode=speo.MODE_PLAIN, rows_per_block=0)
        e.run_fes(100 * D)
        _, _, f64 = e.best()
        assert row["best_fitness"] == pytest.approx(float(f64.min()), rel=1e-12)
        o, M, S = data[fid]
        X, f = e.read()
        i = int(np.argmin(f))
        assert float(f64.min()) == pytest.approx(oracle.cec(fid, X[i:i + 1, :D], o, M, S)[0],
                                                 rel=1e-10)
        e.close()
        p.close()
    # appending: a second sweep adds rows under the same header
    speo.island_driver(tmp_path / "input_data", [1], [3], D, island_num=4, island_size=32,
                       max_base_FEs=100, record_file=rec)
    lines2 = list(csv.reader(open(rec)))
    assert len(lines2) == len(lines) + 1 and lines2[0] == lines[0]


def test_driver_dynamic_regroup_records_minus_one(cuda, speo, oracle, tmp_path):
    D = 10
    o, M, S = oracle.gen_cec_problem(800, 9, D)
    oracle.write_cec_files(tmp_path, 9, o, M, S)
    rows = speo.island_driver(tmp_path / "input_data", [9], [1], D, island_num=2, island_size=32,
                              max_base_FEs=50, interval=2, migration_size=1, subisland_size=8,
                              configure_EA=speo.CONFIGURE_DYNAMIC,
                              regroup_option=speo.REGROUP_DYNAMIC_RANDOM)
    assert rows[0]["subisland_size"] == -1
    assert rows[0]["regroup_option"] == "dynamic_and_random"
    assert rows[0]["configure_EA"] == "dynamic"
