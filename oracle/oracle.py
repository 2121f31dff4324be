"""CPU oracle for the island-model DE/rand/1/bin hot path -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module, and only as the checker (or as the timed CPU baseline).  The
product library never calls it.

It wraps ``libspeo_oracle.so`` (plain C, see speo_oracle.c for the reference
file:line each function follows) and adds, in numpy:

* the seeded problem generator: shift o ~ U[-80, 80]^D rounded to FP32, and a
  Haar-random orthogonal M = Q of the QR of a seeded N(0,1) matrix, sign-fixed
  so that diag(R) > 0.  The reference loads o and M from text files instead
  (CUDA_EA_library/CEC2014_CUDA.cc:160-289, CEC2014.cc:353-420);
* writers of those text files, so the reference's own CEC2014.cc (built into
  oracle/_ref by oracle/Makefile) can evaluate the same problems.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libspeo_oracle.so"
REF_DIR = HERE / "_ref"

OBJ_ELLIPTIC, OBJ_RASTRIGIN, OBJ_ACKLEY, OBJ_BENT_CIGAR, OBJ_DISCUS = 1, 2, 3, 4, 5
OBJ_HAPPYCAT, OBJ_HGBAT, OBJ_SCHWEFEL = 6, 7, 8
OBJ_ROSENBROCK, OBJ_WEIERSTRASS, OBJ_GRIEWANK, OBJ_KATSUURA = 9, 10, 11, 12
OBJ_GRIE_ROSEN, OBJ_ESCAFFER6 = 13, 14
PURPOSE_DONOR, PURPOSE_CROSS, PURPOSE_INIT = 0, 1, 2
PURPOSE_SHIFT, PURPOSE_GAUSS, PURPOSE_PROBLEM = 16, 17, 18

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile the C oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if Path("/root/reference/SPEO-GPU-single-control/CUDA_EA_library/CEC2014.cc").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.oracle_u01.argtypes = [ctypes.c_uint32]
        L.oracle_u01.restype = ctypes.c_float
        L.oracle_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_draw.restype = ctypes.c_float
        L.oracle_check_bound.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.oracle_check_bound.restype = ctypes.c_float
        L.oracle_draw_donors.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_int, _i32p, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
        L.oracle_de_generate.argtypes = [
            _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_void_p, _f32p, _f32p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_float, ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p]
        L.oracle_de_generate.restype = ctypes.c_int
        L.oracle_de_select.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p]
        L.oracle_de_select.restype = ctypes.c_int
        L.oracle_init_population.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                             ctypes.c_uint32]
        L.oracle_basic.argtypes = [ctypes.c_int, _f64p, ctypes.c_int]
        L.oracle_basic.restype = ctypes.c_double
        L.oracle_eval.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, _f64p, _f64p,
                                  ctypes.c_int, _f64p]
        L.oracle_eval_f32.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, _f64p,
                                      _f64p, ctypes.c_int, _f64p]
        L.oracle_composition.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                         _f64p, _f64p, _f64p, _f64p, _f64p, ctypes.c_int, _f64p]
        L.oracle_composition_rated.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                               ctypes.c_int, _f64p]
        L.oracle_cec.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 _f64p, _f64p, ctypes.c_void_p, _f64p]
        L.oracle_cec.restype = ctypes.c_int
        L.oracle_cec_ncomp.argtypes = [ctypes.c_int]
        L.oracle_cec_ncomp.restype = ctypes.c_int
        L.oracle_hybrid_segments.argtypes = [ctypes.c_int, ctypes.c_int, _i32p]
        L.oracle_hybrid_segments.restype = ctypes.c_int
        L.oracle_migrate_ring.argtypes = [_f32p, _f32p, ctypes.c_int, _i32p, _i32p, _i32p,
                                          ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ Philox

def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def u01(x: int) -> float:
    return lib().oracle_u01(x)


def draw(seed, purpose, a, b, idx) -> float:
    return lib().oracle_draw(seed, purpose, a, b, idx)


def philox_np(seed: int, c0, c1, c2, c3):
    """Vectorised Philox4x32-10 over counter arrays (same algorithm, numpy)."""
    c = [np.asarray(v, np.uint64) & 0xFFFFFFFF for v in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c0, c1, c2, c3 = [v.astype(np.uint64) for v in c]
    k0 = np.uint64(seed & 0xFFFFFFFF)
    k1 = np.uint64((seed >> 32) & 0xFFFFFFFF)
    m32 = np.uint64(0xFFFFFFFF)
    for rnd in range(10):
        p0 = np.uint64(0xD2511F53) * c0
        p1 = np.uint64(0xCD9E8D57) * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ k0
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ k1
        c0, c1, c2, c3 = n0 & m32, p1 & m32, n2 & m32, p0 & m32
        if rnd < 9:
            k0 = (k0 + np.uint64(0x9E3779B9)) & m32
            k1 = (k1 + np.uint64(0xBB67AE85)) & m32
    return [v.astype(np.uint32) for v in (c0, c1, c2, c3)]


def u53(wa, wb):
    """Two u32 words -> float64 in (0,1]: (a>>5)*2^26 + (b>>6) + 1, times 2^-53."""
    wa = np.asarray(wa, np.uint64)
    wb = np.asarray(wb, np.uint64)
    k = (wa >> np.uint64(5)) * np.uint64(1 << 26) + (wb >> np.uint64(6)) + np.uint64(1)
    return k.astype(np.float64) * 2.0 ** -53


# --------------------------------------------------------- problem generator

def problem_seed(base_seed: int, index: int) -> int:
    w = philox_np(base_seed, index, 0, 0, PURPOSE_PROBLEM)
    return int(w[0]) | (int(w[1]) << 32)


def gen_shift(seed: int, D: int, lo: float = -80.0, hi: float = 80.0) -> np.ndarray:
    """o_j = fl32(lo + (hi-lo) * u53), double j%2 of Philox block j/2."""
    j = np.arange(D)
    w = philox_np(seed, j // 2, D, 0, PURPOSE_SHIFT)
    first = (j % 2) == 0
    wa = np.where(first, w[0], w[2])
    wb = np.where(first, w[1], w[3])
    o = lo + (hi - lo) * u53(wa, wb)
    return o.astype(np.float32).astype(np.float64)


def gen_gauss(seed: int, D: int) -> np.ndarray:
    """N(0,1) D x D matrix, Box-Muller on Philox block e/2 (e = i*D + j)."""
    e = np.arange(D * D)
    w = philox_np(seed, e // 2, D, 0, PURPOSE_GAUSS)
    u1 = u53(w[0], w[1])
    u2 = u53(w[2], w[3])
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.where(e % 2 == 0, r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2))
    return g.reshape(D, D)


def gen_rotation(seed: int, D: int) -> np.ndarray:
    """Haar orthogonal M = Q diag(sign(diag R)) from G = QR."""
    G = gen_gauss(seed, D)
    Q, R = np.linalg.qr(G)
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return np.ascontiguousarray(Q * s[None, :])


def gen_problem(seed: int, D: int):
    return gen_shift(seed, D), gen_rotation(seed, D)


# ------------------------------------------------------------- DE operators

def check_bound(v, lo, hi) -> float:
    return lib().oracle_check_bound(v, lo, hi)


def draw_donors(seed, row, gen, S, D, inj=None):
    r = np.zeros(5, np.int32)
    jr = ctypes.c_int()
    used = ctypes.c_int()
    inj_a = None if inj is None else np.ascontiguousarray(inj, np.float32)
    n_inj = 0 if inj_a is None else inj_a.size
    lib().oracle_draw_donors(seed, row, gen, S, D, _ptr(inj_a), n_inj, r, ctypes.byref(jr),
                             ctypes.byref(used))
    return r, jr.value, used.value


def de_generate(X, ld, NP, D, subisland_num=1, perm=None, cr=None, f=None, strategy=None,
                best=None, lo=-100.0, hi=100.0, seed=0, gen=0, row_base=0, inj_r=None,
                inj_cr=None, U=None):
    """FP32 DE generation (kernel_DE_CUDA.cu:172-289).  Returns (U, r, jrand, mask)."""
    X = np.ascontiguousarray(X, np.float32)
    U = np.zeros_like(X) if U is None else np.array(U, np.float32, copy=True)
    S_num = subisland_num
    cr = np.full(S_num, 0.9, np.float32) if cr is None else np.ascontiguousarray(cr, np.float32)
    f = np.full(S_num, 0.5, np.float32) if f is None else np.ascontiguousarray(f, np.float32)
    perm_a = None if perm is None else np.ascontiguousarray(perm, np.int32)
    st_a = None if strategy is None else np.ascontiguousarray(strategy, np.int32)
    best_a = None if best is None else np.ascontiguousarray(best, np.int32)
    inj_r_a = None if inj_r is None else np.ascontiguousarray(inj_r, np.float32)
    n_inj = 0 if inj_r_a is None else inj_r_a.shape[1]
    inj_cr_a = None if inj_cr is None else np.ascontiguousarray(inj_cr, np.float32)
    r = np.zeros((NP, 5), np.int32)
    jr = np.zeros(NP, np.int32)
    mask = np.zeros((NP, D), np.uint8)
    rc = lib().oracle_de_generate(U, X, ld, NP, D, S_num, _ptr(perm_a), cr, f, _ptr(st_a),
                                  _ptr(best_a), lo, hi, seed, gen, row_base, _ptr(inj_r_a), n_inj,
                                  _ptr(inj_cr_a), _ptr(r), _ptr(jr), _ptr(mask))
    if rc != 0:
        raise ValueError("oracle_de_generate rejected its arguments")
    return U, r, jr, mask


def de_select(X, fX, U, fU, ld, NP, D):
    X = np.array(X, np.float32, copy=True)
    fX = np.array(fX, np.float32, copy=True)
    acc = np.zeros(NP, np.uint8)
    lib().oracle_de_select(X, fX, np.ascontiguousarray(U, np.float32),
                           np.ascontiguousarray(fU, np.float32), ld, NP, D, _ptr(acc))
    return X, fX, acc


def init_population(NP, D, ld, lo, hi, seed, row_base=0):
    X = np.zeros((NP, ld), np.float32)
    lib().oracle_init_population(X, ld, NP, D, lo, hi, seed, row_base)
    return X


# --------------------------------------------------------------- objectives

def basic(objective, z):
    z = np.ascontiguousarray(z, np.float64)
    return lib().oracle_basic(objective, z, z.size)


def evaluate(objective, x, o, M):
    """FP64 shifted-rotated objective of the rows of x (float32 or float64)."""
    x = np.asarray(x)
    x2 = x.reshape(-1, x.shape[-1])
    D = o.size
    out = np.zeros(x2.shape[0], np.float64)
    o = np.ascontiguousarray(o, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    if x2.dtype == np.float32:
        lib().oracle_eval_f32(objective, np.ascontiguousarray(x2), x2.shape[1], x2.shape[0], o, M,
                              D, out)
    else:
        lib().oracle_eval(objective, np.ascontiguousarray(x2, np.float64), x2.shape[1],
                          x2.shape[0], o, M, D, out)
    return out


def composition(x, objectives, os_, Ms, sigma, bias, lam=None, rate=None):
    x2 = np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1, np.asarray(x).shape[-1]))
    nc = len(objectives)
    D = os_.shape[1]
    lam = np.ones(nc) if lam is None else np.asarray(lam, np.float64)
    rate = np.ones(nc) if rate is None else np.asarray(rate, np.float64)
    out = np.zeros(x2.shape[0], np.float64)
    lib().oracle_composition_rated(x2, x2.shape[1], x2.shape[0], nc,
                                   np.ascontiguousarray(objectives, np.int32),
                                   np.ascontiguousarray(os_, np.float64),
                                   np.ascontiguousarray(Ms, np.float64),
                                   np.ascontiguousarray(sigma, np.float64),
                                   np.ascontiguousarray(bias, np.float64),
                                   np.ascontiguousarray(lam, np.float64),
                                   np.ascontiguousarray(rate, np.float64), D, out)
    return out


def cec_ncomp(func_id: int) -> int:
    return lib().oracle_cec_ncomp(func_id)


def is_hybrid(func_id: int) -> bool:
    return 17 <= func_id <= 22 or func_id in (29, 30)


def hybrid_segments(h: int, D: int):
    out = np.zeros(5, np.int32)
    n = lib().oracle_hybrid_segments(h, D, out)
    return out[:n].tolist()


def cec(func_id, x, os_, Ms, SS=None):
    """CEC2014 F<func_id> (1..30) in FP64 at the rows of x, as the reference's
    CEC2014::EvaluateFitness minus its 100 * func_id.  os_ [ncomp][D],
    Ms [ncomp][D][D], SS [ncomp][D] 1-based shuffle (hybrids only)."""
    x2 = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, np.asarray(x).shape[-1]))
    D = x2.shape[1]
    nc = cec_ncomp(func_id)
    os_ = np.ascontiguousarray(np.asarray(os_, np.float64).reshape(nc, D))
    Ms = np.ascontiguousarray(np.asarray(Ms, np.float64).reshape(nc, D, D))
    ss = None
    if is_hybrid(func_id):
        ss = np.ascontiguousarray(np.asarray(SS, np.int32).reshape(nc, D))
    out = np.zeros(x2.shape[0], np.float64)
    rc = lib().oracle_cec(func_id, x2, D, x2.shape[0], D, os_, Ms, _ptr(ss), out)
    assert rc == 0
    return out


def gen_shuffle(seed: int, D: int, ncomp: int = 1) -> np.ndarray:
    """Seeded 1-based permutations of 1..D, one per component (the format of
    the reference's shuffle_data files)."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.permutation(D) + 1 for _ in range(ncomp)]).astype(np.int32)


def gen_cec_problem(seed: int, func_id: int, D: int):
    """Seeded data of F<func_id> at dimension D: shifts, Haar rotations and
    (hybrids) shuffles, one per component."""
    nc = cec_ncomp(func_id)
    os_ = np.zeros((nc, D))
    Ms = np.zeros((nc, D, D))
    for c in range(nc):
        os_[c], Ms[c] = gen_problem(seed * 16 + c, D)
    SS = gen_shuffle(seed, D, nc) if is_hybrid(func_id) else None
    return os_, Ms, SS


# ------------------------------------------------- sub-island regroups

PURPOSE_REGROUP = 19
REGROUP_STATIC_ORDERED, REGROUP_STATIC_RANDOM, REGROUP_DYNAMIC_RANDOM, REGROUP_DYNAMIC_ORDERED = range(4)
CONFIGURE_CONSTANT, CONFIGURE_DYNAMIC = 0, 1


class HostDraws:
    """Host-side draws of one regroup event: word k of Philox with counter
    {k/4, island, event, PURPOSE_REGROUP}; rand_int(a, b) = a + w % (b - a + 1)
    with no draw when a == b, as Random::RandIntUnif (src/random.cc:14-19)
    with rand() replaced by the Philox word.  `words` (optional) injects the
    words instead."""

    def __init__(self, seed, island, event, words=None):
        self.seed, self.island, self.event, self.k = seed, island, event, 0
        self.words = words

    def rand_int(self, a, b):
        if a == b:
            return a
        if self.words is not None:
            w = int(self.words[self.k])
        else:
            key = [self.seed & 0xFFFFFFFF, self.seed >> 32]
            w = int(philox([self.k >> 2, self.island, self.event, PURPOSE_REGROUP], key)[self.k & 3])
        self.k += 1
        return a + w % (b - a + 1)


def permutate(draws, n):
    """Random::Permutate(n, n), src/random.cc:29-53: pick a random remaining
    index, move the last remaining one into its place."""
    g = list(range(n))
    out = []
    for i in range(n, 0, -1):
        t = draws.rand_int(0, i - 1)
        out.append(g[t])
        g[t] = g[i - 1]
    return out


def island_regroup(seed, island, event, option, configure, np_, subs, perm, base, info,
                   words=None):
    """One regroup event (migrate.cc:147-186, DE_CUDA.cc:89-119): returns
    (subisland_num, perm, [(CR, F, strategy, pop_size)] per sub-island).
    Event 0 is DE_CUDA::Initilize's ConfigureEA.  base = (CR, F, strategy)."""
    d = HostDraws(seed, island, event, words)
    perm = list(perm)
    configure_now = event == 0
    if event > 0:
        if option in (REGROUP_DYNAMIC_RANDOM, REGROUP_DYNAMIC_ORDERED):
            max_index, p = 0, np_ // 4
            while p > 1:
                p //= 2
                max_index += 1
            size = 4 << d.rand_int(0, max_index)
            subs = np_ // size
            configure_now = True
        if option in (REGROUP_DYNAMIC_RANDOM, REGROUP_STATIC_RANDOM):
            perm = permutate(d, np_)
    S = np_ // subs
    out = []
    for s in range(subs):
        if configure_now and configure == CONFIGURE_DYNAMIC:
            cr = np.float32([0.1, 0.9][d.rand_int(0, 1)])
            st = [4, 6][d.rand_int(0, 1)]
            out.append((float(cr), float(np.float32(base[1])), st, S))
        elif configure_now:
            out.append((float(np.float32(base[0])), float(np.float32(base[1])), int(base[2]), S))
        else:
            cr, f, st, _ = info[s]
            out.append((cr, f, st, S))
    return subs, perm, out


def migrate_ring(X, fX, ld, off, dims, dst, k):
    X = np.array(X, np.float32, copy=True)
    fX = np.array(fX, np.float32, copy=True)
    lib().oracle_migrate_ring(X, fX, ld, np.ascontiguousarray(off, np.int32),
                              np.ascontiguousarray(dims, np.int32),
                              np.ascontiguousarray(dst, np.int32), len(dims), k)
    return X, fX


# ------------------------------------------------- the reference's CEC2014.cc

REF_FUNC = {OBJ_ELLIPTIC: 1, OBJ_BENT_CIGAR: 2, OBJ_DISCUS: 3, OBJ_ACKLEY: 5, OBJ_RASTRIGIN: 9,
            OBJ_HAPPYCAT: 13, OBJ_HGBAT: 14, OBJ_SCHWEFEL: 11}
# CEC2014.cc:154-208: F1, F2, F3 and F5 use rate 1; F9 scales x - o by
# 5.12/100, F11 by 10, F13 and F14 by 5/100.
REF_RATE = {1: 1.0, 2: 1.0, 3: 1.0, 5: 1.0, 9: 0.0512, 11: 10.0, 13: 0.05, 14: 0.05}


def write_cec_files(directory, func_id, o, M, SS=None):
    """Write o, M (and the shuffle SS) in the reference's text format
    (CEC2014.cc:353-460): one shift line per component, ncomp * D * D matrix
    values, ncomp * D 1-based shuffle indices."""
    d = Path(directory) / "input_data"
    d.mkdir(parents=True, exist_ok=True)
    o = np.asarray(o, np.float64)
    D = o.shape[-1]
    np.savetxt(d / f"shift_data_{func_id}.txt", o.reshape(-1, D), fmt="%.17e")
    np.savetxt(d / f"M_{func_id}_D{D}.txt", np.asarray(M, np.float64).reshape(-1, D),
               fmt="%.17e")
    if SS is not None:
        np.savetxt(d / f"shuffle_data_{func_id}_D{D}.txt",
                   np.asarray(SS, np.int64).reshape(1, -1), fmt="%d")


def ref_binary(opt="O2") -> Path:
    return REF_DIR / f"cec2014_ref_{opt}"


def ref_eval(workdir, func_id, x, opt="O2"):
    """Evaluate rows of x (float64) with the reference's CEC2014::EvaluateFitness."""
    x = np.ascontiguousarray(x, np.float64)
    n, D = x.shape
    res = subprocess.run([str(ref_binary(opt)), "eval", str(func_id), str(D), str(n)],
                         input=x.tobytes(), cwd=str(workdir), capture_output=True, check=True)
    return np.frombuffer(res.stdout, np.float64).copy()


def ref_time(workdir, func_id, x, seconds, opt="O2"):
    """Evals/s of the reference's CEC2014::EvaluateFitness on one point."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    res = subprocess.run([str(ref_binary(opt)), "time", str(func_id), str(x.size),
                          str(seconds)], input=x.tobytes(), cwd=str(workdir),
                         capture_output=True, check=True)
    evals, el, _ = res.stdout.decode().split()
    return int(evals) / float(el)


def ref_available() -> bool:
    return ref_binary("O2").exists() and os.access(ref_binary("O2"), os.X_OK)
