"""Python host mirror of the MI355X island-model DE engine.

Thin wrappers over the C ABI (include/speo_de.h, libspeo_de.so).  Device
buffers are torch tensors on ``cuda`` (PyTorch is plumbing here: memory and
streams); every computation runs in the library's HIP kernels.  The classes
follow the reference's host roles:

* ``Problem``  -- CEC2014_CUDA (CUDA_EA_library/CEC2014_CUDA.h:38-82): a
  shifted-rotated objective held on the device.
* ``Engine``   -- EA_CUDA / DE_CUDA + IslandEA (EA_CUDA.h:46-103,
  src/island_EA.cc:78-113): islands in HBM, generations, migration, best.
* module functions ``init_population``, ``de_generate``, ``de_select``,
  ``evaluate`` -- the kernel-level entry points (API_Initilize,
  API_DE_GenerateNewPopulation, API_DE_SelectSurvival, API_evaluateFitness).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int, c_void_p

import numpy as np

from ._lib import (DEInfo, DriverConfig, EngineConfig, RunRecord, EngineStats, GPU_IslandInfo, GPU_Population,
                   MODE_EPOCH, MODE_FUSED, MODE_PLAIN, OBJ_ACKLEY, OBJ_BENT_CIGAR, OBJ_DISCUS,
                   OBJ_ELLIPTIC, OBJ_ESCAFFER6, OBJ_GRIE_ROSEN, OBJ_GRIEWANK, OBJ_HAPPYCAT,
                   OBJ_HGBAT, OBJ_HYBRID_1, OBJ_HYBRID_2, OBJ_HYBRID_3, OBJ_HYBRID_4,
                   OBJ_HYBRID_5, OBJ_HYBRID_6, OBJ_KATSUURA, OBJ_RASTRIGIN, OBJ_ROSENBROCK,
                   OBJ_SCHWEFEL, OBJ_WEIERSTRASS, PREC_BF16, REGROUP_STATIC_ORDERED,
                   REGROUP_STATIC_RANDOM, REGROUP_DYNAMIC_RANDOM, REGROUP_DYNAMIC_ORDERED,
                   CONFIGURE_CONSTANT, CONFIGURE_DYNAMIC,
                   PREC_BF16X3, PREC_BF16X6, PREC_FP16, PREC_FP16X3, PREC_FP32, PREC_NAMES,
                   ProblemInfo, SPEO_RandState, SpeoError, check, last_error, lib)

__all__ = [
    "Problem", "Engine", "generate_problem_data", "problem_seed", "init_population",
    "de_generate", "de_select", "evaluate", "evaluate_f64", "evaluate_noalloc", "segment_argmin", "SpeoError",
    "OBJ_ELLIPTIC", "OBJ_RASTRIGIN", "OBJ_ACKLEY", "OBJ_BENT_CIGAR", "OBJ_DISCUS", "OBJ_HAPPYCAT",
    "OBJ_HGBAT", "OBJ_SCHWEFEL", "OBJ_ROSENBROCK", "OBJ_WEIERSTRASS", "OBJ_GRIEWANK",
    "OBJ_KATSUURA", "OBJ_GRIE_ROSEN", "OBJ_ESCAFFER6", "OBJ_HYBRID_1", "OBJ_HYBRID_2",
    "OBJ_HYBRID_3", "OBJ_HYBRID_4", "OBJ_HYBRID_5", "OBJ_HYBRID_6", "PREC_FP32", "PREC_BF16", "PREC_FP16",
    "PREC_BF16X3", "PREC_FP16X3", "PREC_BF16X6", "PREC_NAMES", "MODE_PLAIN", "MODE_FUSED",
    "MODE_EPOCH",    "DEInfo", "GPU_Population", "GPU_IslandInfo", "ProblemInfo", "SPEO_RandState", "lib",
    "REGROUP_STATIC_ORDERED", "REGROUP_STATIC_RANDOM", "REGROUP_DYNAMIC_RANDOM",
    "REGROUP_DYNAMIC_ORDERED", "CONFIGURE_CONSTANT", "CONFIGURE_DYNAMIC", "island_regroup",
    "island_driver", "generations_for_fes", "run_seed", "load_cec_data", "de_parameters", "result_header", "result_record",
]


def _p(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else c_void_p(t.data_ptr())


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _stream(stream):
    if stream is None:
        return None
    return c_void_p(stream.cuda_stream)


# ------------------------------------------------------------------ problems

def problem_seed(base_seed: int, index: int) -> int:
    return int(lib().speo_problem_seed(base_seed, index))


def generate_problem_data(seed: int, dim: int):
    """Seeded o ~ U[-80, 80]^D (FP32-exact) and Haar orthogonal M (FP64)."""
    o = np.zeros(dim, np.float64)
    M = np.zeros((dim, dim), np.float64)
    check(lib().speo_generate_problem_data(seed, dim, _np_ptr(o), _np_ptr(M)))
    return o, M


class Problem:
    """A shifted-rotated problem (single component or cf_cal composition)."""

    def __init__(self, objectives, shifts, rotations=None, sigma=None, bias=None, lam=None,
                 rate=None, shuffle=None):
        objectives = np.atleast_1d(np.asarray(objectives, np.int32))
        nc = objectives.size
        shifts = np.ascontiguousarray(np.asarray(shifts, np.float64).reshape(nc, -1))
        self.dim = shifts.shape[1]
        self.ncomp = nc
        rot = None
        if rotations is not None:
            rot = np.ascontiguousarray(np.asarray(rotations, np.float64).reshape(nc, self.dim,
                                                                                self.dim))
        arrs = [None if v is None else np.ascontiguousarray(np.asarray(v, np.float64).reshape(nc))
                for v in (sigma, bias, lam, rate)]
        shuf = None
        if shuffle is not None:
            shuf = np.ascontiguousarray(np.asarray(shuffle, np.int32).reshape(nc, self.dim))
        h = c_void_p()
        check(lib().speo_problem_create_ex(self.dim, nc, _np_ptr(objectives), _np_ptr(shifts),
                                           _np_ptr(rot), _np_ptr(shuf),
                                           *[_np_ptr(a) for a in arrs], byref(h)))
        self.handle = h
        self.shuffle = shuf
        self.objectives, self.shifts, self.rotations = objectives, shifts, rot
        self.sigma, self.bias, self.lam, self.rate = arrs

    @classmethod
    def from_cec(cls, directory, func_id: int, dim: int):
        """CEC2014 function `func_id` from the reference's data files in
        `directory` (speo_problem_create_cec)."""
        self = cls.__new__(cls)
        h = c_void_p()
        check(lib().speo_problem_create_cec(str(directory).encode(), func_id, dim, byref(h)))
        self.handle, self.dim = h, dim
        self.ncomp = int(lib().speo_problem_ncomp(h))
        self.objectives = self.shifts = self.rotations = self.shuffle = None
        self.sigma = self.bias = self.lam = self.rate = None
        return self

    @classmethod
    def seeded(cls, objective: int, dim: int, seed: int):
        o, M = generate_problem_data(seed, dim)
        return cls([objective], o[None, :], M[None, :, :])

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().speo_problem_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def island_regroup(seed, island, event, option, configure, np_, subs, perm, base, info=None):
    """speo_island_regroup: (subisland_num, perm, [(CR, F, strategy, pop_size)]).
    base = (CR, F, strategy); info = the current per-sub-island configuration."""
    perm_a = np.ascontiguousarray(perm, np.int32).copy()
    arr = (DEInfo * max(1, np_ // 4))()
    if info is not None:
        for s, (cr, f, st, ps) in enumerate(info):
            arr[s] = DEInfo(cr, f, st, ps)
    b = DEInfo(base[0], base[1], base[2], -1)
    out = c_int()
    check(lib().speo_island_regroup(seed, island, event, option, configure, np_, subs,
                                    _np_ptr(perm_a), byref(b), ctypes.cast(arr, c_void_p),
                                    byref(out)))
    return out.value, perm_a, [(arr[s].CR, arr[s].F, arr[s].strategy_ID, arr[s].pop_size)
                               for s in range(out.value)]


def island_driver(data_dir, functions, runs, dim, island_num, island_size, max_base_FEs,
                  record_file=None, node_ID=0, node_num=1, GPU_num=1, interval=10,
                  migration_size=2, migration_rate=0.5, buffer_capacity=32, subisland_size=0,
                  configure_EA=CONFIGURE_CONSTANT, regroup_option=REGROUP_STATIC_ORDERED,
                  mode=MODE_PLAIN, precision=PREC_FP32, use_graph=False, lo=-100.0, hi=100.0):
    """speo_island_driver: the function x run sweep; returns the records as dicts."""
    f = np.ascontiguousarray(functions, np.int32)
    r = np.ascontiguousarray(runs, np.int32)
    c = DriverConfig()
    c.data_dir = str(data_dir).encode()
    c.record_file = None if record_file is None else str(record_file).encode()
    c.dim, c.num_functions, c.num_runs = dim, f.size, r.size
    c.function_IDs, c.run_IDs = _np_ptr(f), _np_ptr(r)
    c.node_ID, c.node_num, c.GPU_num = node_ID, node_num, GPU_num
    c.island_num, c.island_size, c.max_base_FEs = island_num, island_size, max_base_FEs
    c.interval, c.migration_size = interval, migration_size
    c.migration_rate, c.buffer_capacity = migration_rate, buffer_capacity
    c.subisland_size, c.configure_EA, c.regroup_option = subisland_size, configure_EA, regroup_option
    c.mode, c.precision, c.use_graph = mode, precision, int(use_graph)
    c.min_bound, c.max_bound = lo, hi
    rows = (RunRecord * max(1, f.size * r.size))()
    check(lib().speo_island_driver(byref(c), ctypes.cast(rows, c_void_p)))
    out = []
    for k in range(f.size * r.size):
        d = {name: getattr(rows[k], name) for name, _ in RunRecord._fields_}
        out.append({k2: (v.decode() if isinstance(v, bytes) else v) for k2, v in d.items()})
    return out


def load_cec_data(directory, func_id: int, dim: int, ncomp: int = 1):
    """(shift [ncomp][dim], rotation [ncomp][dim][dim]) from the reference's
    data files (speo_load_cec_data)."""
    o = np.zeros((ncomp, dim), np.float64)
    M = np.zeros((ncomp, dim, dim), np.float64)
    check(lib().speo_load_cec_data(str(directory).encode(), func_id, dim, ncomp,
                                   o.ctypes.data_as(c_void_p), M.ctypes.data_as(c_void_p)))
    return o, M


def evaluate(problem: Problem, X, f_out, precision=PREC_FP32, ld=None, n=None, stream=None):
    """f_out[i] = f(X[i]) (FP32 values, rotation GEMM in `precision`)."""
    ld = X.shape[1] if ld is None else ld
    n = X.shape[0] if n is None else n
    check(lib().speo_evaluate(problem.handle, _p(X), ld, n, _p(f_out), precision,
                              _stream(stream)))


def wide_eval(enable=-1):
    """Selects the kernel of the 16-bit evaluations wider than 128 columns on
    at least 1024 rows (1: 256-column tiles, 0: 64-row tiles, -1: unchanged;
    the same bits either way).  Returns (previous setting, launches of the
    256-column kernel so far)."""
    n = ctypes.c_longlong(0)
    prev = lib().speo_wide_eval(enable, ctypes.byref(n))
    return prev, n.value


def evaluate_noalloc(problem: Problem, X, f_out, precision=PREC_FP32, ld=None, n=None,
                     stream=None):
    """evaluate() with no device allocation: the workspace is a fixed 2^20-float
    array per GPU, shared by every call on that GPU (overlapping calls race)."""
    ld = X.shape[1] if ld is None else ld
    n = X.shape[0] if n is None else n
    check(lib().speo_evaluate_noalloc(problem.handle, _p(X), ld, n, _p(f_out), precision,
                                      _stream(stream)))


def evaluate_f64(problem: Problem, X, f_out, ld=None, n=None, stream=None):
    ld = X.shape[1] if ld is None else ld
    n = X.shape[0] if n is None else n
    check(lib().speo_evaluate_f64(problem.handle, _p(X), ld, n, _p(f_out), _stream(stream)))


# ------------------------------------------------------------- DE operators

def init_population(X, dim, lo, hi, seed, row_base=0, n=None, stream=None):
    n = X.shape[0] if n is None else n
    check(lib().speo_init_population(_p(X), X.shape[1], n, dim, lo, hi, seed, row_base,
                                     _stream(stream)))


def de_generate(U, X, dim, de_info, subisland_num=1, perm=None, best=None, lo=-100.0, hi=100.0,
                seed=0, gen=0, row_base=0, inj_r=None, inj_cr=None, r_out=None, jrand_out=None,
                n=None, stream=None):
    """DE generation X -> U (device tensors; de_info a device DEInfo array)."""
    n = X.shape[0] if n is None else n
    rng = SPEO_RandState(seed, gen, row_base)
    n_inj = 0 if inj_r is None else inj_r.shape[1]
    check(lib().speo_de_generate(_p(U), _p(X), X.shape[1], n, dim, subisland_num, _p(perm),
                                 _p(de_info), _p(best), lo, hi, rng, _p(inj_r), n_inj,
                                 _p(inj_cr), _p(r_out), _p(jrand_out), _stream(stream)))


def de_select(X, fX, U, fU, dim, accept_out=None, n=None, stream=None):
    n = X.shape[0] if n is None else n
    check(lib().speo_de_select(_p(X), _p(fX), _p(U), _p(fU), X.shape[1], n, dim,
                               _p(accept_out), _stream(stream)))


def segment_argmin(f, off, best_out, stream=None):
    check(lib().speo_segment_argmin(_p(f), _p(off), off.numel() - 1, _p(best_out),
                                    _stream(stream)))


def generations_for_fes(max_fes: int, current: int, per_generation: int) -> int:
    return int(lib().speo_generations_for_fes(int(max_fes), int(current), int(per_generation)))


def run_seed(node_id: int, function_id: int, run_id: int) -> int:
    """The reference driver's per-run seed (src/main.cc:138)."""
    return int(lib().speo_run_seed(node_id, function_id, run_id))


def de_parameters(cr: float, f: float, strategy: int = 4) -> str:
    """DE_CUDA::GetParameters' string, e.g. "CR/F=0.9/0.5_rand/1/bin"."""
    buf = ctypes.create_string_buffer(128)
    n = lib().speo_de_parameters(float(cr), float(f), int(strategy), buf, len(buf))
    check(n if n < 0 else 0)
    return buf.value.decode()


def result_header() -> str:
    """The header line of the reference's SPEO_GPU.csv (island_EA.cc:23)."""
    return lib().speo_result_header().decode()


def result_record(**fields) -> str:
    """One result row (island_EA.cc:146-171) from the 21 named columns of
    speo_run_record; string columns are str."""
    from ._lib import RunRecord
    rec = RunRecord()
    for name, _ in RunRecord._fields_:
        if name not in fields:
            raise TypeError(f"result_record: missing column {name}")
        v = fields[name]
        setattr(rec, name, v.encode() if isinstance(v, str) else v)
    buf = ctypes.create_string_buffer(1024)
    n = lib().speo_result_record(byref(rec), buf, len(buf))
    check(n if n < 0 else 0)
    return buf.value.decode()


def de_info_bytes(cr, f, strategy=4, pop_size=-1) -> bytes:
    """Packed DEInfo array (EA_CUDA.h:33-39) for upload to the device."""
    cr = np.atleast_1d(np.asarray(cr, np.float32))
    f = np.broadcast_to(np.asarray(f, np.float32), cr.shape)
    st = np.broadcast_to(np.asarray(strategy, np.int32), cr.shape)
    arr = (DEInfo * cr.size)()
    for i in range(cr.size):
        arr[i] = DEInfo(float(cr[i]), float(f[i]), int(st[i]), pop_size)
    return bytes(arr)


# ------------------------------------------------------------------- engine

class Engine:
    """Islands of DE/rand/1/bin in one GPU's HBM (see speo_engine_* in the header)."""

    def __init__(self, problems, island_size, island_problem, F=0.5, CR=0.9, lo=-100.0,
                 hi=100.0, seed=1, interval=0, migration_size=0, precision=PREC_FP32,
                 mode=MODE_FUSED, use_graph=False, rows_per_block=0, island_F=None,
                 island_CR=None, surrogate_candidates=0, surrogate_train_interval=10,
                 surrogate_archive=65536, surrogate_steps=40, surrogate_batch=1024,
                 surrogate_lr=1e-3, surrogate_rho_min=0.5, epoch_parts=0, subisland_size=0,
                 strategy=4, configure_EA=CONFIGURE_CONSTANT,
                 regroup_option=REGROUP_STATIC_ORDERED):
        self.problems = list(problems)
        self.island_size = np.ascontiguousarray(island_size, np.int32)
        self.island_problem = np.ascontiguousarray(island_problem, np.int32)
        self._probs = (c_void_p * len(self.problems))(*[p.handle.value for p in self.problems])
        self._iF = None if island_F is None else np.ascontiguousarray(island_F, np.float32)
        self._iCR = None if island_CR is None else np.ascontiguousarray(island_CR, np.float32)
        cfg = EngineConfig()
        cfg.num_islands = self.island_size.size
        cfg.island_size = self.island_size.ctypes.data_as(POINTER(c_int))
        cfg.island_problem = self.island_problem.ctypes.data_as(POINTER(c_int))
        cfg.num_problems = len(self.problems)
        cfg.problems = ctypes.cast(self._probs, POINTER(c_void_p))
        cfg.island_F = None if self._iF is None else self._iF.ctypes.data_as(POINTER(c_float))
        cfg.island_CR = None if self._iCR is None else self._iCR.ctypes.data_as(POINTER(c_float))
        cfg.F, cfg.CR, cfg.min_bound, cfg.max_bound = F, CR, lo, hi
        cfg.seed = seed
        cfg.interval, cfg.migration_size = interval, migration_size
        cfg.precision, cfg.mode, cfg.use_graph = precision, mode, int(use_graph)
        cfg.rows_per_block = rows_per_block
        cfg.surrogate_candidates = surrogate_candidates
        cfg.surrogate_train_interval = surrogate_train_interval
        cfg.surrogate_archive = surrogate_archive
        cfg.surrogate_steps = surrogate_steps
        cfg.surrogate_batch = surrogate_batch
        cfg.surrogate_lr = surrogate_lr
        cfg.surrogate_rho_min = surrogate_rho_min
        cfg.epoch_parts = epoch_parts
        cfg.subisland_size = subisland_size
        cfg.de_strategy = 0 if strategy == 4 else strategy + 1
        cfg.configure_EA = configure_EA
        cfg.regroup_option = regroup_option
        self.cfg = cfg
        h = c_void_p()
        check(lib().speo_engine_create(byref(cfg), None, byref(h)))
        self.handle = h
        self.total_rows = lib().speo_engine_total_rows(h)
        self.max_dim = lib().speo_engine_max_dim(h)
        self.num_islands = int(self.island_size.size)
        self.row_off = np.concatenate([[0], np.cumsum(self.island_size)]).astype(np.int64)

    def run(self, generations: int):
        check(lib().speo_engine_run(self.handle, generations))

    def surrogate_dump(self, problem: int, candidates: int, archive: int):
        """(cands, cand [rows][cands][Kp], pred [rows][cands], archive X, archive y,
        head, appended) of one problem's surrogate."""
        rows = int(sum(self.island_size[self.island_problem == problem]))
        Kp = (self.max_dim + 31) // 32 * 32
        ce = c_int()
        cand = np.zeros((rows, candidates, Kp), np.float32)
        pred = np.zeros((rows, candidates), np.float32)
        AX = np.zeros((archive, Kp), np.float32)
        Ay = np.zeros(archive, np.float32)
        head = c_int()
        app = ctypes.c_int64()
        check(lib().speo_engine_surrogate_dump(self.handle, problem, byref(ce), _np_ptr(cand),
                                               _np_ptr(pred), _np_ptr(AX), _np_ptr(Ay),
                                               byref(head), byref(app)))
        n = ce.value
        flat_c = cand.reshape(-1)[:rows * n * Kp].reshape(rows, n, Kp)
        flat_p = pred.reshape(-1)[:rows * n].reshape(rows, n)
        return n, flat_c, flat_p, AX, Ay, head.value, app.value

    def surrogate_set_labels(self, problem: int, y):
        y = np.ascontiguousarray(y, np.float32)
        check(lib().speo_engine_surrogate_set_labels(self.handle, problem, _np_ptr(y)))

    def subislands(self, island: int):
        """(subisland_num, perm [island_size], [(CR, F, strategy, pop_size)])"""
        n = int(self.island_size[island])
        subs = c_int()
        perm = np.zeros(n, np.int32)
        info = (DEInfo * max(1, n // 4))()
        check(lib().speo_engine_subislands(self.handle, island, byref(subs), _np_ptr(perm),
                                           ctypes.cast(info, c_void_p)))
        return subs.value, perm, [(info[s].CR, info[s].F, info[s].strategy_ID, info[s].pop_size)
                                  for s in range(subs.value)]

    def prepare(self, generations: int):
        """Capture the graphs the next run(generations) launches (no work)."""
        check(lib().speo_engine_prepare(self.handle, generations))

    def run_fes_timed(self, max_fes: int):
        """(generations, seconds, communication seconds) of the budget loop."""
        g, t, c = c_int(0), c_double(0.0), c_double(0.0)
        check(lib().speo_engine_run_fes_timed(self.handle, int(max_fes), byref(g), byref(t),
                                              byref(c)))
        return g.value, t.value, c.value

    def run_fes(self, max_fes: int) -> int:
        """Generations until the evaluation count reaches max_fes (the budget
        loop of IslandEA::Execute, src/island_EA.cc:86-104); returns them."""
        g = c_int(0)
        check(lib().speo_engine_run_fes(self.handle, int(max_fes), byref(g)))
        return g.value

    def sync(self):
        check(lib().speo_engine_sync(self.handle))

    def best(self):
        I = self.num_islands
        rows = np.zeros(I, np.int32)
        f32 = np.zeros(I, np.float32)
        f64 = np.zeros(I, np.float64)
        check(lib().speo_engine_best(self.handle, _np_ptr(rows), _np_ptr(f32), _np_ptr(f64)))
        return rows, f32, f64

    def read(self):
        X = np.zeros((self.total_rows, self.max_dim), np.float32)
        f = np.zeros(self.total_rows, np.float32)
        check(lib().speo_engine_read(self.handle, _np_ptr(X), _np_ptr(f)))
        return X, f

    def read_trials(self):
        U = np.zeros((self.total_rows, self.max_dim), np.float32)
        f = np.zeros(self.total_rows, np.float32)
        check(lib().speo_engine_read_trials(self.handle, _np_ptr(U), _np_ptr(f)))
        return U, f

    def write(self, X=None, f=None):
        Xa = None if X is None else np.ascontiguousarray(X, np.float32)
        fa = None if f is None else np.ascontiguousarray(f, np.float32)
        check(lib().speo_engine_write(self.handle, _np_ptr(Xa), _np_ptr(fa)))

    def profile(self, generations: int):
        """Average kernel ms per generation and role: generate, evaluate, select
        (plain mode), fused (fused mode), epoch and donors (epoch mode: the epoch
        kernel and the donor draws ahead of it, per generation)."""
        out = np.zeros(6, np.float64)
        check(lib().speo_engine_profile(self.handle, generations, _np_ptr(out)))
        return dict(zip(("generate", "evaluate", "select", "fused", "epoch", "donors"), out.tolist()))

    def surrogate_stats(self):
        n = len(self.problems)
        rho = np.zeros(n, np.float32)
        act = np.zeros(n, np.int32)
        loss = np.zeros(n, np.float32)
        check(lib().speo_engine_surrogate_stats(self.handle, _np_ptr(rho), _np_ptr(act),
                                                _np_ptr(loss)))
        return rho, act, loss

    def stats(self):
        s = EngineStats()
        check(lib().speo_engine_get_stats(self.handle, byref(s)))
        return {k: getattr(s, k) for k, _ in EngineStats._fields_}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().speo_engine_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
