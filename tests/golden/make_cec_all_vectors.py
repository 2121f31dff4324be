"""Writes tests/golden/cec2014_all_vectors.npz: CEC2014 F1-F30 values from the
reference's own objective code (CUDA_EA_library/CEC2014.cc, built from its
source into oracle/_ref by oracle/Makefile), for every function the reference
defines (CEC2014.cc:14-145).  Run in the build container, where the reference
tree is mounted; the fixture is committed and the tests only read it.

Cases:
  "ref":  the reference's bin/input_data files (shifts, matrices, shuffles) at
          D = 10 and 30, with shifts and matrices rounded to FP32 (the values
          the reference's GPU_SINGLE_PRECISION build reads, CEC2014_CUDA.cc:
          195-277); the rounded data are written back in the reference's text
          format and evaluated there.  Stored: x, o, M, shuffle, f.
  "haar": this project's seeded problems (oracle.gen_cec_problem: shifts in
          [-80, 80], Haar rotations, seeded shuffles) at D = 50 and 100; only
          the seed is stored with x and f.
Points x are FP32 values, uniform in [-100, 100].
"""
from __future__ import annotations

import sys
import tempfile
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO / "oracle"))
import oracle as O  # noqa: E402

REF_DATA = Path("/root/reference/SPEO-GPU-single-control/bin/input_data")
NPTS = 8
OUT = Path(__file__).resolve().parent / "cec2014_all_vectors.npz"


def load_ref(fid, D):
    """The reference's files as CEC2014::Initilize reads them (CEC2014.cc:353-460)."""
    nc = O.cec_ncomp(fid)
    lines = (REF_DATA / f"shift_data_{fid}.txt").read_text().split("\n")
    if nc == 1:
        o = np.array(" ".join(lines).split(), float)[:D][None, :]
    else:
        o = np.stack([np.array(lines[c].split(), float)[:D] for c in range(nc)])
    M = np.array((REF_DATA / f"M_{fid}_D{D}.txt").read_text().split(), float)[:nc * D * D]
    SS = None
    if O.is_hybrid(fid):
        SS = np.array((REF_DATA / f"shuffle_data_{fid}_D{D}.txt").read_text().split(),
                      np.int32)[:nc * D].reshape(nc, D)
    return o, M.reshape(nc, D, D), SS


def ref_values(fid, x, o, M, SS):
    with tempfile.TemporaryDirectory() as td:
        O.write_cec_files(td, fid, o, M, SS)
        return O.ref_eval(td, fid, x)


def main():
    O.build()
    rng = np.random.default_rng(20261203)
    out, cases = {}, []
    for D in (10, 30):
        for fid in range(1, 31):
            o, M, SS = load_ref(fid, D)
            o = o.astype(np.float32).astype(np.float64)
            M = M.astype(np.float32).astype(np.float64)
            x = rng.uniform(-100, 100, (NPTS, D)).astype(np.float32).astype(np.float64)
            key = f"ref_F{fid}_D{D}"
            out[key + "_x"] = x.astype(np.float32)
            out[key + "_o"] = o.astype(np.float32)
            out[key + "_M"] = M.astype(np.float32)
            if SS is not None:
                out[key + "_S"] = SS
            out[key + "_f"] = ref_values(fid, x, o, M, SS)
            cases.append(key)
    for D in (50, 100):
        for fid in range(1, 31):
            seed = 1000 + fid
            o, M, SS = O.gen_cec_problem(seed, fid, D)
            x = rng.uniform(-100, 100, (NPTS, D)).astype(np.float32).astype(np.float64)
            key = f"haar_F{fid}_D{D}"
            out[key + "_x"] = x.astype(np.float32)
            out[key + "_seed"] = np.array(seed)
            out[key + "_f"] = ref_values(fid, x, o, M, SS)
            cases.append(key)
    out["cases"] = np.array(cases)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(cases)} cases)")


if __name__ == "__main__":
    main()
