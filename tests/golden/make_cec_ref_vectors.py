"""Writes tests/golden/cec2014_ref_vectors.npz from the reference's own objective
code (CUDA_EA_library/CEC2014.cc, built from its source into oracle/_ref by
oracle/Makefile).  Run in the build container, where the reference tree is
mounted; the fixture is committed and the tests only read it.

Cases (func_id -> objective; CEC2014.cc:154-171, 186):
  F1 rotated elliptic, rate 1;  F5 rotated Ackley, rate 1;
  F2 rotated bent cigar, rate 1;  F3 rotated discus, rate 1;
  F13 rotated HappyCat and F14 rotated HGBat, rate 5/100 -- evaluated at a
  rate-adjusted input like F9;
  F11 rotated Schwefel, rate 10 -- at the rate-adjusted input x' = o + (x - o)/10,
  with x in [-1000, 1000] so that |z| reaches the folded range beyond 500;
  F9 rotated Rastrigin, rate 5.12/100 -- evaluated at x' = o + (x - o)/0.0512
  so that the reference computes Rastrigin(M (x - o)) at rate 1.
Two data sources:
  "refdata": o and M from the reference's bin/input_data files (D = 10, 30);
  "haar":    this project's seeded problems (o, M regenerated from the seed by
             the oracle's generator; only the seed is stored), written in the
             reference's text format (D = 10, 100).
Compositions "comp": F24 (cf02) and F25 (cf03) on the reference's data files
(D = 10, 30), at the points themselves.
Points x are float32-representable, uniform in [-100, 100].
"""
from __future__ import annotations

import sys
import tempfile
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO / "oracle"))
import oracle as O  # noqa: E402

REF_BIN = Path("/root/reference/SPEO-GPU-single-control/bin")
# Functions added later come last so that the random points of the earlier
# cases keep their values.
FUNCS = {1: O.OBJ_ELLIPTIC, 5: O.OBJ_ACKLEY, 9: O.OBJ_RASTRIGIN, 2: O.OBJ_BENT_CIGAR,
         3: O.OBJ_DISCUS, 13: O.OBJ_HAPPYCAT, 14: O.OBJ_HGBAT, 11: O.OBJ_SCHWEFEL}
# point range per function (default [-100, 100])
XRANGE = {11: 1000.0}
NPTS = 12


def ref_inputs(fid, x, o):
    rate = O.REF_RATE[fid]
    if rate == 1.0:
        return x
    return o[None, :] + (x - o[None, :]) / rate


def main():
    O.build()
    rng = np.random.default_rng(20260512)
    out = {}
    cases = []
    for fid in FUNCS:
        for D in (10, 30):
            o = np.loadtxt(REF_BIN / "input_data" / f"shift_data_{fid}.txt").reshape(-1)[:D]
            M = np.loadtxt(REF_BIN / "input_data" / f"M_{fid}_D{D}.txt").reshape(D, D)
            r = XRANGE.get(fid, 100.0)
            x = rng.uniform(-r, r, (NPTS, D)).astype(np.float32).astype(np.float64)
            f = O.ref_eval(REF_BIN, fid, ref_inputs(fid, x, o))
            key = f"refdata_F{fid}_D{D}"
            out[key + "_x"], out[key + "_o"], out[key + "_M"], out[key + "_f"] = x, o, M, f
            cases.append(key)
        for D, seed in ((10, 101), (100, 102)):
            o, M = O.gen_problem(seed, D)
            r = XRANGE.get(fid, 100.0)
            x = rng.uniform(-r, r, (NPTS, D)).astype(np.float32).astype(np.float64)
            with tempfile.TemporaryDirectory() as td:
                O.write_cec_files(td, fid, o, M)
                f = O.ref_eval(td, fid, ref_inputs(fid, x, o))
            key = f"haar_F{fid}_D{D}"
            out[key + "_x"], out[key + "_f"] = x, f
            out[key + "_seed"] = np.array(seed)
            cases.append(key)
    # Compositions F24 (cf02) and F25 (cf03) on the reference's own data files:
    # three shift lines and three matrices; points taken as they are (each
    # component applies its own rate).
    for fid in (24, 25):
        for D in (10, 30):
            o = np.loadtxt(REF_BIN / "input_data" / f"shift_data_{fid}.txt")[:3, :D]
            M = np.loadtxt(REF_BIN / "input_data" / f"M_{fid}_D{D}.txt").reshape(-1)[:3 * D * D]
            M = M.reshape(3, D, D)
            x = rng.uniform(-100, 100, (NPTS, D)).astype(np.float32).astype(np.float64)
            f = O.ref_eval(REF_BIN, fid, x)
            key = f"comp_F{fid}_D{D}"
            out[key + "_x"], out[key + "_o"], out[key + "_M"], out[key + "_f"] = x, o, M, f
            cases.append(key)
    out["cases"] = np.array(cases)
    dst = Path(__file__).resolve().parent / "cec2014_ref_vectors.npz"
    np.savez_compressed(dst, **out)
    print(f"wrote {dst} ({dst.stat().st_size} bytes, {len(cases)} cases)")


if __name__ == "__main__":
    main()
