"""The loader of the reference's problem data files (speo_load_cec_data;
CEC2014.cc:353-420, CEC2014_CUDA.cc:195-277) on the CPU: the first dim values
of each shift line (the reference's lines hold 100), ncomp * dim * dim matrix
values, one line per composition component, missing and short files; and,
when the reference tree is present, its own files against the values the
committed golden vectors hold for them."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "cec2014_ref_vectors.npz"
REF_DATA = Path("/root/reference/SPEO-GPU-single-control/bin/input_data")


def _write(dirpath, fid, shifts, Ms, width=100, seed=0):
    """Shift lines padded to `width` values and matrices one row per line,
    in the reference's text format."""
    rng = np.random.default_rng(seed)
    d = Path(dirpath)
    with open(d / f"shift_data_{fid}.txt", "w") as f:
        for o in shifts:
            row = np.concatenate([o, rng.uniform(-80, 80, width - o.size)])
            f.write(" ".join(f"{v: .16e}" for v in row) + "\n")
    D = shifts.shape[1]
    with open(d / f"M_{fid}_D{D}.txt", "w") as f:
        for M in Ms:
            for r in M:
                f.write(" ".join(f"{v: .16e}" for v in r) + "\n")


def test_single_function_files(speo, tmp_path):
    rng = np.random.default_rng(1)
    for D in (2, 10, 30, 100):
        o = rng.uniform(-80, 80, (1, D))
        M = rng.normal(size=(1, D, D))
        _write(tmp_path, 7, o, M)
        o2, M2 = speo.load_cec_data(tmp_path, 7, D)
        np.testing.assert_array_equal(o2, o)
        np.testing.assert_array_equal(M2, M)


def test_composition_files(speo, tmp_path):
    # component c's shift is the first D values of line c; the matrix file
    # holds the components' matrices one after another
    rng = np.random.default_rng(2)
    D, nc = 10, 4
    o = rng.uniform(-80, 80, (nc, D))
    M = rng.normal(size=(nc, D, D))
    _write(tmp_path, 24, o, M)
    o2, M2 = speo.load_cec_data(tmp_path, 24, D, nc)
    np.testing.assert_array_equal(o2, o)
    np.testing.assert_array_equal(M2, M)


def test_missing_and_short_files(speo, tmp_path):
    with pytest.raises(speo.SpeoError, match="shift_data_5.txt|M_5_D10.txt"):
        speo.load_cec_data(tmp_path, 5, 10)
    rng = np.random.default_rng(3)
    _write(tmp_path, 6, rng.uniform(-80, 80, (1, 10)), rng.normal(size=(1, 10, 10)))
    with pytest.raises(speo.SpeoError, match="too few values in .*M_6_D10.txt"):
        speo.load_cec_data(tmp_path, 6, 10, 2)
    (tmp_path / "M_9_D4.txt").write_text(" ".join(["0.5"] * 16) + "\n")
    (tmp_path / "shift_data_9.txt").write_text("1 2 3\n")
    with pytest.raises(speo.SpeoError, match="too few values in .*shift_data_9.txt"):
        speo.load_cec_data(tmp_path, 9, 4)


def test_unknown_function_id(speo, tmp_path):
    with pytest.raises(speo.SpeoError, match="function 31 is not in 1..30") as e:
        speo.Problem.from_cec(tmp_path, 31, 10)
    assert e.value.code == -1


def test_missing_shuffle_file_is_reported(speo, oracle, tmp_path):
    # F17 needs shuffle_data_17_D10.txt besides its shift and matrix files
    o, M, S = oracle.gen_cec_problem(5, 17, 10)
    oracle.write_cec_files(tmp_path, 17, o, M)
    with pytest.raises(speo.SpeoError, match="shuffle_data_17_D10.txt") as e:
        speo.Problem.from_cec(tmp_path / "input_data", 17, 10)
    assert e.value.code == -1


def test_hybrid_shuffle_must_be_a_permutation(speo, oracle):
    o, M, S = oracle.gen_cec_problem(5, 17, 10)
    S = S.copy()
    S[0, 0] = S[0, 1]
    with pytest.raises(speo.SpeoError, match="permutation") as e:
        speo.Problem([speo.OBJ_HYBRID_1], o, M, shuffle=S)
    assert e.value.code == -1
    with pytest.raises(speo.SpeoError, match="needs a shuffle"):
        speo.Problem([speo.OBJ_HYBRID_1], o, M)


def test_reference_files_if_present(speo):
    if not REF_DATA.is_dir():
        pytest.skip("reference tree absent")
    g = np.load(GOLDEN)
    cases = [str(c) for c in g["cases"] if str(c).startswith("refdata")]
    assert cases
    for case in cases:
        fid = int(case.split("_F")[1].split("_")[0])
        D = int(case.split("_D")[1])
        o, M = speo.load_cec_data(REF_DATA, fid, D)
        np.testing.assert_array_equal(o[0], g[case + "_o"])
        np.testing.assert_array_equal(M[0], g[case + "_M"])
