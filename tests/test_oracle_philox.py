"""Pins the oracle's Philox4x32-10 and its uniform conversion.

Checks: the Random123 known-answer triplets (Salmon et al., SC'11, kat_vectors),
the open-source rocRAND engine (rocrand_philox4x32_10.h compiled for the host,
test-only) on random counters and keys, and exactness of the u32 -> (0,1]
float conversion at its ends.
"""
import shutil
import subprocess

import numpy as np
import pytest

from conftest import REPO

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_random123_kat(oracle, ctr, key, expect):
    assert tuple(int(v) for v in oracle.philox(ctr, key)) == expect
    seed = key[0] | (key[1] << 32)
    w = oracle.philox_np(seed, *ctr)
    assert tuple(int(v) for v in w) == expect


def test_rocrand_cross_check(oracle, tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not shutil.which(hipcc):
        pytest.skip("hipcc not available to build the rocRAND cross-check")
    exe = tmp_path / "philox_kat"
    subprocess.run([hipcc, "-O1", "--offload-arch=gfx950", "-x", "hip",
                    str(REPO / "tests" / "cpp" / "philox_rocrand_kat.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    rng = np.random.default_rng(7)
    rows = rng.integers(0, 2**32, size=(200, 6), dtype=np.uint64)
    text = "\n".join(" ".join(str(int(v)) for v in r) for r in rows) + "\n"
    res = subprocess.run([str(exe)], input=text.encode(), capture_output=True, check=True)
    got = np.array([[int(v) for v in line.split()] for line in res.stdout.decode().splitlines()],
                   dtype=np.uint64)
    seeds = [int(r[4]) | (int(r[5]) << 32) for r in rows]
    for i, r in enumerate(rows):
        mine = oracle.philox([int(v) for v in r[:4]], [int(r[4]), int(r[5])])
        assert [int(v) for v in mine] == [int(v) for v in got[i]]
    w = np.stack([np.stack(oracle.philox_np(seeds[i], *[int(v) for v in rows[i, :4]]))
                  for i in range(len(rows))])
    assert (w.astype(np.uint64) == got).all()


def test_u01_range_and_exactness(oracle):
    assert oracle.u01(0) == 2.0 ** -24
    assert oracle.u01(0xFFFFFFFF) == 1.0
    assert oracle.u01(0xFF) == 2.0 ** -24          # low 8 bits are dropped
    assert oracle.u01(0x100) == 2.0 * 2.0 ** -24
    x = np.random.default_rng(1).integers(0, 2**32, 1000, dtype=np.uint64)
    u = np.array([oracle.u01(int(v)) for v in x])
    assert ((u > 0) & (u <= 1)).all()
    assert np.array_equal(u, ((x >> 8) + 1).astype(np.float64) * 2.0 ** -24)


def test_draw_addressing(oracle):
    # draw idx of (purpose, a, b) = word idx%4 of block {idx/4, a, b, purpose}
    seed = 0x0123456789ABCDEF
    for idx in (0, 1, 2, 3, 4, 7, 100):
        w = oracle.philox([idx // 4, 11, 22, 1], [seed & 0xFFFFFFFF, seed >> 32])
        assert oracle.draw(seed, 1, 11, 22, idx) == oracle.u01(int(w[idx % 4]))
