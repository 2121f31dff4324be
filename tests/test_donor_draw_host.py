"""The library's donor draw and reflective bound (speo-gpu-single-control_amd/
csrc/de.h, common.h), built for the host, against the oracle
(oracle_draw_donors follows kernel_DE_CUDA.cu:187-222, oracle_check_bound
kernel_DE_CUDA.cu:160-170): the same draws, donors, j_rand and draw count for
random rows, for tiny sub-islands where redraws are frequent (down to S = 1,
which runs into the 4096-draw cap), and for injected sequences; and the same
bits of the bound for in-range, reflected, far-out, infinite and NaN values.
Also the integer form of the crossover test (exhaustive over all 2^24 draw
values for a set of CR) and the accuracy of the objective's sin^2(pi z)
polynomial and of the Schwefel terms' bounded sine against FP64."""
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import REPO

PKG = REPO / "speo-gpu-single-control_amd"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not shutil.which(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("donor") / "donor_draw_check"
    subprocess.run([hipcc, "-O2", "--offload-arch=gfx950", "-x", "hip", "-std=c++17",
                    f"-I{REPO / 'include'}", f"-I{PKG / 'csrc'}",
                    str(REPO / "tests" / "cpp" / "donor_draw_check.cpp"), "-o", str(out)],
                   check=True, capture_output=True)
    return out


def _run(exe, lines):
    res = subprocess.run([str(exe)], input=("\n".join(lines) + "\n").encode(),
                         capture_output=True, check=True)
    return res.stdout.decode().splitlines()


def _bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_donor_draws_match_oracle(exe, oracle):
    rng = np.random.default_rng(11)
    cases = []
    for S in (1, 2, 3, 4, 5, 6, 8, 16, 50, 128, 512):
        for _ in range(40 if S > 1 else 3):
            cases.append((int(rng.integers(0, 2**64, dtype=np.uint64)), int(rng.integers(0, 2**31)),
                          int(rng.integers(0, 5000)), S, int(rng.integers(1, 200)), None))
    for _ in range(60):
        n = int(rng.integers(1, 12))
        inj = rng.choice([0.05, 0.25, 0.3, 0.5, 0.6, 0.99, 1.0, 1.2, -0.1], n).astype(np.float32)
        cases.append((int(rng.integers(0, 2**32)), int(rng.integers(0, 1000)), 3, int(rng.integers(2, 9)),
                      10, inj))
    lines = []
    for seed, row, gen, S, D, inj in cases:
        tail = "0" if inj is None else f"{inj.size} " + " ".join(str(_bits(float(v))) for v in inj)
        lines.append(f"D {seed} {row} {gen} {S} {D} {tail}")
    got = _run(exe, lines)
    assert len(got) == len(cases)
    capped = 0
    for (seed, row, gen, S, D, inj), out in zip(cases, got):
        r, jr, used = oracle.draw_donors(seed, row, gen, S, D, inj)
        vals = [int(v) for v in out.split()]
        assert vals[:5] == [int(v) for v in r], (seed, row, gen, S, D)
        assert vals[5] == int(jr) and vals[6] == int(used)
        capped += int(used) >= 4096
    assert capped >= 1   # S = 1 exercised the draw cap


def test_check_bound_matches_oracle(exe, oracle):
    rng = np.random.default_rng(3)
    vals = list(rng.uniform(-100, 100, 200)) + list(rng.uniform(-300, 300, 200)) + \
        list(rng.uniform(-1e6, 1e6, 50)) + [np.inf, -np.inf, np.nan, 100.0, -100.0, 100.00001,
                                              -100.00001, 3e38, -3e38, 0.0, -0.0]
    lines = [f"B {_bits(float(v))} -100 100" for v in vals] + \
            [f"B {_bits(float(v))} -5 17.5" for v in vals[:100]]
    got = [int(x) for x in _run(exe, lines)]
    ref = [_bits(oracle.check_bound(float(v), -100.0, 100.0)) for v in vals] + \
          [_bits(oracle.check_bound(float(v), -5.0, 17.5)) for v in vals[:100]]
    assert got == ref


def _f32(bits):
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def test_crossover_threshold_exhaustive(exe):
    crs = [0.9, 0.5, 0.1, 1.0, 0.0, -0.5, 1e-9, 2.0 ** -24, 3 * 2.0 ** -25, 0.999999,
           1.0 + 2.0 ** -23, 1.5, float("inf"), float("nan"), 0.3333333]
    got = [int(x) for x in _run(exe, [f"T {_bits(c)}" for c in crs])]
    n = np.arange(1 << 24, dtype=np.uint32)
    u = (n + np.uint32(1)).astype(np.float32) * np.float32(2.0 ** -24)   # u01 of x = n << 8
    for cr, thr in zip(crs, got):
        want = u < np.float32(cr)
        assert np.array_equal(n < thr, want), cr


def test_sin2pi_accuracy(exe):
    rng = np.random.default_rng(9)
    z = np.concatenate([rng.uniform(-3, 3, 20000), rng.uniform(-200, 200, 20000),
                        rng.normal(0, 1e-3, 2000), np.arange(-8, 8.5, 0.5),
                        [1e-30, -1e-20, 2.0 ** 23, 2.0 ** 23 + 2, 1e7 + 0.25]]).astype(np.float32)
    got = np.array([_f32(int(x)) for x in _run(exe, [f"S {_bits(float(v))}" for v in z])],
                   np.float64)
    zd = z.astype(np.float64)
    ref = np.sin(np.pi * (zd - np.rint(zd))) ** 2
    tiny = ref < 1e-30
    assert np.all(got[tiny] <= 1e-30)
    rel = np.abs(got[~tiny] - ref[~tiny]) / ref[~tiny]
    assert rel.max() < 3e-7, rel.max()


def test_sin_small_accuracy(exe):
    # sin x on [0, 23] (the Schwefel terms' sin(sqrt v), v <= 500): absolute
    # error against FP64 within a few FP32 ulps of 1, including the
    # reduction points k pi +- pi/2 and the ends of the range.
    rng = np.random.default_rng(10)
    k = np.arange(8)
    x = np.concatenate([rng.uniform(0, 23, 40000), np.sqrt(rng.uniform(0, 500, 20000)),
                        k * np.pi, k * np.pi + np.pi / 2, k * np.pi - np.pi / 2,
                        [0.0, 1e-30, 1e-7, 22.36068, 23.0]]).astype(np.float32)
    x = x[(x >= 0) & (x <= 23)]
    got = np.array([_f32(int(v)) for v in _run(exe, [f"N {_bits(float(t))}" for t in x])],
                   np.float64)
    err = np.abs(got - np.sin(x.astype(np.float64)))
    assert err.max() < 4e-7, err.max()
