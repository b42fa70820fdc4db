"""ctypes face of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of
__graft_entry__.py and by the cpu_baseline leg of bench.py, as the checker.
The product (pagmo2_amd) never imports it.

Every function here forwards to oracle_capi.cpp, which restates the pagmo
2.19.1 reference path (see the citations in oracle.hpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PGH_ORACLE_LIB names another build of the same sources (the sanitized one
# of tests/test_sanitize_cpu.py)
LIB_PATH = os.environ.get("PGH_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

ROSENBROCK, RASTRIGIN, ACKLEY, SCHWEFEL, GRIEWANK, LENNARD_JONES, MO_SPHERE, ZDT, DTLZ = range(1, 10)

_lib = None


def build() -> None:
    """Compile oracle/liboracle.so with g++ (the recipe is oracle/Makefile)."""
    subprocess.check_call(["make", "-s", "-B", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64, u32, dbl, cint = C.c_uint64, C.c_uint32, C.c_double, C.c_int


def _declare(L):
    L.orc_last_error.restype = C.c_char_p
    sigs = {
        "orc_mt19937_u32": [u32, u64, u64, _u32p],
        "orc_uniform_real": [u32, dbl, dbl, u64, _dp],
        "orc_problem_info": [cint, u64, u64, C.POINTER(u64), C.POINTER(u64), C.c_void_p, C.c_void_p],
        "orc_fitness_one": [cint, u64, C.c_void_p, u64, _dp, u64, _dp],
        "orc_bfe": [cint, u64, C.c_void_p, u64, _dp, u64, u64, _dp, u32],
        "orc_bfe_timed": [cint, u64, C.c_void_p, u64, _dp, u64, u64, _dp, u32, u32, C.POINTER(dbl)],
        "orc_gradient": [cint, u64, _dp, u64, u64, _dp],
        "orc_mo_info": [cint, u32, u64, u64, u32, C.POINTER(u64), C.POINTER(u64), C.c_void_p, C.c_void_p],
        "orc_mo_fitness": [cint, u32, u64, u64, u32, _dp, u64, u64, _dp],
        "orc_local_descent": [cint, u64, _dp, u64, _dp, C.c_void_p, C.c_void_p, u32, dbl],
        "orc_estimate_gradient": [cint, u64, C.c_void_p, u64, _dp, u64, dbl, cint, _dp],
        "orc_estimate_gradient_kat": [_dp, dbl, cint, _dp],
        "orc_pareto_dominance": [_dp, _dp, u64, C.POINTER(cint)],
        "orc_fnds_timed": [_dp, u64, u64, u32, C.POINTER(dbl)],
        "orc_population_init": [cint, u64, C.c_void_p, u64, u64, u32, _dp, _dp],
        "orc_batch_random_dv": [cint, u64, u64, u32, _dp],
        "orc_batch_uniform": [_dp, _dp, u64, u64, u32, _dp],
        "orc_pso_gen_evolve": [cint, u64, u64, _dp, _dp, u32, dbl, dbl, dbl, dbl, u32, u32, u32, cint, u32,
                               cint, u32, u32, C.c_void_p, C.c_void_p],
        "orc_pso_gen_timed": [cint, u64, u64, _dp, _dp, u32, u32, u32, u32, C.POINTER(dbl)],
        "orc_crowding_distance": [_dp, u64, u64, cint, _dp],
        "orc_select_best_N_mo": [_dp, u64, u64, u64, cint, _u64p, C.POINTER(u64)],
        "orc_sort_population_mo": [_dp, u64, u64, cint, _u64p],
        "orc_sort_population_mo_timed": [_dp, u64, u64, u32, C.POINTER(dbl)],
        "orc_select_best": [u64, u64, _u64p, _dp, _dp, cint, u64, dbl, cint, C.POINTER(u64), _u64p, _dp, _dp],
        "orc_fair_replace": [u64, u64, _u64p, _dp, _dp, u64, _u64p, _dp, _dp, cint, u64, dbl, cint, C.POINTER(u64),
                             _u64p, _dp, _dp],
        "orc_topology": [cint, u64, dbl, _u64p, _u64p, _dp],
        "orc_archi_migrate": [u64, C.c_void_p, _u64p, _u64p, _dp, _dp, _u32p, _u32p, cint, cint, _u64p, _u64p, _dp,
                              C.c_void_p, _u64p, _dp, C.c_void_p, _u64p, _dp, _u32p, u32, cint, _dp, _dp, _u64p,
                              _dp, _dp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = cint
    L.orc_fnds_run.argtypes = [_dp, u64, u64]
    L.orc_fnds_run.restype = C.c_void_p
    L.orc_fnds_sizes.argtypes = [C.c_void_p, C.POINTER(u64), C.POINTER(u64)]
    L.orc_fnds_sizes.restype = None
    L.orc_fnds_copy.argtypes = [C.c_void_p, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p]
    L.orc_fnds_copy.restype = None
    L.orc_fnds_free.argtypes = [C.c_void_p]
    L.orc_fnds_free.restype = None


class OracleError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _cptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def mt19937_u32(seed: int, n: int, skip: int = 0) -> np.ndarray:
    out = np.empty(n, np.uint32)
    _chk(lib().orc_mt19937_u32(seed, skip, n, out))
    return out


def uniform_real(seed: int, a: float, b: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _chk(lib().orc_uniform_real(seed, a, b, n, out))
    return out


def problem_info(kind: int, dim: int, nobj: int = 1):
    nx, nf = u64(), u64()
    _chk(lib().orc_problem_info(kind, dim, nobj, C.byref(nx), C.byref(nf), None, None))
    lb = np.empty(nx.value)
    ub = np.empty(nx.value)
    _chk(lib().orc_problem_info(kind, dim, nobj, C.byref(nx), C.byref(nf), _cptr(lb), _cptr(ub)))
    return nx.value, nf.value, lb, ub


def fitness_one(kind: int, dim: int, x, centers=None, nobj: int = 1) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    c = None if centers is None else np.ascontiguousarray(centers, np.float64)
    nf = nobj if kind == MO_SPHERE else 1
    f = np.empty(nf)
    _chk(lib().orc_fitness_one(kind, dim, _cptr(c), nobj, x, x.size, f))
    return f


def bfe(kind: int, dim: int, dvs: np.ndarray, centers=None, nobj: int = 1, threads: int = 1) -> np.ndarray:
    dvs = np.ascontiguousarray(dvs, np.float64)
    n, nx = dvs.shape
    c = None if centers is None else np.ascontiguousarray(centers, np.float64)
    nf = nobj if kind == MO_SPHERE else 1
    out = np.empty((n, nf))
    _chk(lib().orc_bfe(kind, dim, _cptr(c), nobj, dvs, n, nx, out, threads))
    return out


def bfe_timed(kind: int, dim: int, dvs: np.ndarray, threads: int, reps: int, centers=None, nobj: int = 1):
    dvs = np.ascontiguousarray(dvs, np.float64)
    n, nx = dvs.shape
    c = None if centers is None else np.ascontiguousarray(centers, np.float64)
    nf = nobj if kind == MO_SPHERE else 1
    out = np.empty((n, nf))
    sec = dbl()
    _chk(lib().orc_bfe_timed(kind, dim, _cptr(c), nobj, dvs, n, nx, out, threads, reps, C.byref(sec)))
    return out, sec.value


def gradient(kind: int, dim: int, dvs: np.ndarray) -> np.ndarray:
    dvs = np.ascontiguousarray(dvs, np.float64)
    n, nx = dvs.shape
    out = np.empty((n, nx))
    _chk(lib().orc_gradient(kind, dim, dvs, n, nx, out))
    return out


def mo_info(kind: int, prob_id: int, dim: int, nobj: int = 2, alpha: int = 100):
    """(nx, nf, lb, ub) of a ZDT (kind ZDT, dim = the constructor's param) or
    DTLZ (dim, nobj = fdim, alpha) problem."""
    nx, nf = u64(), u64()
    _chk(lib().orc_mo_info(kind, prob_id, dim, nobj, alpha, C.byref(nx), C.byref(nf), None, None))
    lb, ub = np.empty(nx.value), np.empty(nx.value)
    _chk(lib().orc_mo_info(kind, prob_id, dim, nobj, alpha, C.byref(nx), C.byref(nf), lb.ctypes.data,
                           ub.ctypes.data))
    return nx.value, nf.value, lb, ub


def mo_fitness(kind: int, prob_id: int, dim: int, dvs, nobj: int = 2, alpha: int = 100) -> np.ndarray:
    """Fitness of every row of dvs for a ZDT or DTLZ problem (n x nf)."""
    dvs = np.ascontiguousarray(dvs, np.float64)
    if dvs.ndim == 1:
        dvs = dvs[None, :]
    n, nx = dvs.shape
    _, nf, _, _ = mo_info(kind, prob_id, dim, nobj, alpha)
    out = np.empty((n, nf))
    _chk(lib().orc_mo_fitness(kind, prob_id, dim, nobj, alpha, dvs, n, nx, out))
    return out


def local_descent(kind: int, dim: int, dvs: np.ndarray, iters: int, step0: float, lb=None, ub=None):
    """The memetic local step (no reference counterpart) on a copy of dvs
    (n x nx); returns (x, f)."""
    x = np.array(dvs, dtype=np.float64, order="C", copy=True)
    n = x.shape[0]
    f = np.empty(max(n, 1))
    lbv = None if lb is None else np.ascontiguousarray(lb, np.float64)
    ubv = None if ub is None else np.ascontiguousarray(ub, np.float64)
    _chk(lib().orc_local_descent(kind, dim, x, n, f, _cptr(lbv), _cptr(ubv), iters, step0))
    return x, f[:n]


def estimate_gradient(kind: int, dim: int, x, dx: float, high: bool, centers=None, nobj: int = 1):
    x = np.ascontiguousarray(x, np.float64)
    c = None if centers is None else np.ascontiguousarray(centers, np.float64)
    nf = nobj if kind == MO_SPHERE else 1
    out = np.empty(nf * x.size)
    _chk(lib().orc_estimate_gradient(kind, dim, _cptr(c), nobj, x, x.size, dx, int(high), out))
    return out


def estimate_gradient_kat(x, dx: float, high: bool) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(12)
    _chk(lib().orc_estimate_gradient_kat(x, dx, int(high), out))
    return out


def pareto_dominance(a, b) -> bool:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if a.size != b.size:
        raise ValueError("Different number of objectives found in input fitnesses: "
                         f"{a.size} and {b.size}. I cannot define dominance")
    r = cint()
    _chk(lib().orc_pareto_dominance(a, b, a.size, C.byref(r)))
    return bool(r.value)


def fnds(objs: np.ndarray):
    """fast_non_dominated_sorting -> (fronts, dom_list, dom_count, rank) in CSR form:
    dict with front_offsets, front_items, dom_offsets, dom_items, dom_count, rank."""
    objs = np.ascontiguousarray(objs, np.float64)
    n, m = objs.shape
    L = lib()
    h = L.orc_fnds_run(objs if objs.size else np.zeros(1), n, m)
    if not h:
        raise OracleError(L.orc_last_error().decode())
    try:
        nfr, ndi = u64(), u64()
        L.orc_fnds_sizes(h, C.byref(nfr), C.byref(ndi))
        r = dict(dom_count=np.empty(n, np.uint64), rank=np.empty(n, np.uint64),
                 front_offsets=np.empty(nfr.value + 1, np.uint64), front_items=np.empty(n, np.uint64),
                 dom_offsets=np.empty(n + 1, np.uint64), dom_items=np.empty(max(ndi.value, 1), np.uint64))
        L.orc_fnds_copy(h, r["dom_count"], r["rank"], r["front_offsets"], r["front_items"], r["dom_offsets"],
                        r["dom_items"])
        r["dom_items"] = r["dom_items"][: ndi.value]
        return r
    finally:
        L.orc_fnds_free(h)


def _pts(objs):
    objs = np.ascontiguousarray(objs, np.float64)
    if objs.ndim != 2:
        objs = objs.reshape(len(objs), -1)
    return objs, objs.shape[0], objs.shape[1], (objs if objs.size else np.zeros(1))


def crowding_distance(objs, stable=False) -> np.ndarray:
    """crowding_distance (multi_objective.cpp:280-318); stable=True breaks ties
    by position (std::stable_sort) instead of std::sort."""
    objs, n, m, buf = _pts(objs)
    out = np.empty(max(n, 1))
    _chk(lib().orc_crowding_distance(buf, n, m, int(stable), out))
    return out[:n]


def select_best_N_mo(objs, N, stable=False) -> np.ndarray:
    """select_best_N_mo (multi_objective.cpp:344-396)."""
    objs, n, m, buf = _pts(objs)
    out = np.empty(max(n, 1), np.uint64)
    ln = u64()
    _chk(lib().orc_select_best_N_mo(buf, n, m, N, int(stable), out, C.byref(ln)))
    return out[: ln.value]


def sort_population_mo(objs, stable=False) -> np.ndarray:
    """sort_population_mo (multi_objective.cpp:425-465)."""
    objs, n, m, buf = _pts(objs)
    out = np.empty(max(n, 1), np.uint64)
    _chk(lib().orc_sort_population_mo(buf, n, m, int(stable), out))
    return out[:n]


def sort_population_mo_timed(objs, reps: int) -> float:
    objs, n, m, buf = _pts(objs)
    sec = dbl()
    _chk(lib().orc_sort_population_mo_timed(buf, n, m, reps, C.byref(sec)))
    return sec.value


def fnds_timed(objs: np.ndarray, reps: int) -> float:
    objs = np.ascontiguousarray(objs, np.float64)
    n, m = objs.shape
    sec = dbl()
    _chk(lib().orc_fnds_timed(objs, n, m, reps, C.byref(sec)))
    return sec.value


def population_init(kind: int, dim: int, pop: int, seed: int, centers=None, nobj: int = 1):
    nx, nf, _, _ = problem_info(kind, dim, nobj)
    c = None if centers is None else np.ascontiguousarray(centers, np.float64)
    X = np.empty((pop, nx))
    F = np.empty((pop, nf))
    _chk(lib().orc_population_init(kind, dim, _cptr(c), nobj, pop, seed, X, F))
    return X, F


def batch_random_dv(kind: int, dim: int, n: int, seed: int) -> np.ndarray:
    nx, _, _, _ = problem_info(kind, dim)
    out = np.empty((n, nx))
    _chk(lib().orc_batch_random_dv(kind, dim, n, seed, out))
    return out


def batch_uniform(lb, ub, n: int, seed: int) -> np.ndarray:
    lb = np.ascontiguousarray(lb, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    out = np.empty((n, lb.size))
    _chk(lib().orc_batch_uniform(lb, ub, lb.size, n, seed, out))
    return out


def pso_gen_evolve(kind: int, dim: int, X: np.ndarray, F: np.ndarray, gen: int, seed: int, omega=0.7298, eta1=2.05,
                   eta2=2.05, max_vel=0.5, variant=5, neighb_type=2, neighb_param=4, memory=False, use_bfe=True,
                   threads=1, trace=False, calls=1):
    X = np.array(X, np.float64, order="C")
    F = np.array(F, np.float64, order="C").reshape(-1)
    pop, nx = X.shape
    tX = np.empty((calls * gen, pop, nx)) if trace else None
    tF = np.empty((calls * gen, pop)) if trace else None
    _chk(lib().orc_pso_gen_evolve(kind, dim, pop, X, F, gen, omega, eta1, eta2, max_vel, variant, neighb_type,
                                  neighb_param, int(memory), seed, int(use_bfe), threads, calls, _cptr(tX),
                                  _cptr(tF)))
    return (X, F, tX, tF) if trace else (X, F)


def pso_gen_timed(kind: int, dim: int, X: np.ndarray, F: np.ndarray, gen: int, seed: int, threads: int, reps: int):
    X = np.ascontiguousarray(X, np.float64)
    F = np.ascontiguousarray(F, np.float64).reshape(-1)
    sec = dbl()
    _chk(lib().orc_pso_gen_timed(kind, dim, X.shape[0], X, F, gen, seed, threads, reps, C.byref(sec)))
    return sec.value


# ---- migration (select_best, fair_replace, topologies, archipelago rounds) ----

def _rate(r):
    """A policy rate as pagmo takes it: an int is an absolute count, a float a
    fraction of the input individuals."""
    if isinstance(r, (float, np.floating)):
        return 1, 0, float(r)
    return 0, int(r), 0.0


def _group(ids, X, F):
    X = np.ascontiguousarray(X, np.float64)
    n = X.shape[0]
    ids = np.arange(n, dtype=np.uint64) if ids is None else np.ascontiguousarray(ids, np.uint64)
    return ids, X, np.ascontiguousarray(F, np.float64).reshape(-1)


def select_best(ids, X, F, rate, stable=False):
    """select_best(rate).select on a single-objective group: (ids, X, F) of
    the selected individuals, best first."""
    ids, X, F = _group(ids, X, F)
    n, nx = X.shape
    fr, cnt, fra = _rate(rate)
    no = u64()
    oi, oX, oF = np.empty(n, np.uint64), np.empty((n, nx)), np.empty(n)
    _chk(lib().orc_select_best(n, nx, ids, X, F, fr, cnt, fra, int(stable), C.byref(no), oi, oX, oF))
    k = no.value
    return oi[:k], oX[:k], oF[:k]


def fair_replace(ids, X, F, mids, mX, mF, rate, stable=False):
    """fair_replace(rate).replace of a single-objective group by migrants."""
    ids, X, F = _group(ids, X, F)
    mids, mX, mF = _group(mids, np.asarray(mX, np.float64).reshape(-1, X.shape[1]), mF)
    n, nx = X.shape
    fr, cnt, fra = _rate(rate)
    no = u64()
    oi, oX, oF = np.empty(n, np.uint64), np.empty((n, nx)), np.empty(n)
    _chk(lib().orc_fair_replace(n, nx, ids, X, F, mX.shape[0], mids, mX, mF, fr, cnt, fra, int(stable), C.byref(no),
                                oi, oX, oF))
    return oi, oX, oF


def topology(which: str, n: int, w: float = 1.0):
    """Inbound connections of ring(n, w) or fully_connected(n, w): a list, per
    island, of (source, weight) in get_connections order."""
    off = np.zeros(n + 1, np.uint64)
    cap = max(n * (n - 1), 1)
    src, wt = np.zeros(cap, np.uint64), np.zeros(cap)
    _chk(lib().orc_topology(0 if which == "ring" else 1, n, w, off, src, wt))
    return [[(int(src[j]), float(wt[j])) for j in range(int(off[i]), int(off[i + 1]))] for i in range(n)]


def archi_migrate(islands, conns, mtype, handling, rounds, stable=True):
    """`rounds` archipelago::evolve(1) rounds with migration.  islands: dicts
    with kind, dim, X (pop x nx), F, seed (pso_gen), gen, sel, rep (rates) and
    migr_seed; conns as topology() returns.  mtype 0 p2p / 1 broadcast;
    handling 0 preserve / 1 evict.  Returns the final (X, F) per island and
    the migrants database as (X, F) per island."""
    n = len(islands)
    kinds = np.array([d["kind"] for d in islands], np.int32)
    dims = np.array([d["dim"] for d in islands], np.uint64)
    pops = np.array([np.asarray(d["X"]).shape[0] for d in islands], np.uint64)
    nxs = [np.asarray(d["X"]).shape[1] for d in islands]
    X0 = np.concatenate([np.ascontiguousarray(d["X"], np.float64).reshape(-1) for d in islands])
    F0 = np.concatenate([np.ascontiguousarray(d["F"], np.float64).reshape(-1) for d in islands])
    seeds = np.array([d["seed"] for d in islands], np.uint32)
    gens = np.array([d["gen"] for d in islands], np.uint32)
    off = np.zeros(n + 1, np.uint64)
    src, wt = [], []
    for k, c in enumerate(conns):
        off[k + 1] = off[k] + len(c)
        src += [s for s, _ in c]
        wt += [w for _, w in c]
    src = np.array(src + [0], np.uint64)
    wt = np.array(wt + [0.0], np.float64)
    sr = [_rate(d["sel"]) for d in islands]
    rr = [_rate(d["rep"]) for d in islands]
    sf = np.array([r[0] for r in sr], np.int32)
    rf = np.array([r[0] for r in rr], np.int32)
    mseeds = np.array([d["migr_seed"] for d in islands], np.uint32)
    Xo, Fo = np.empty_like(X0), np.empty_like(F0)
    dbc = np.zeros(n, np.uint64)
    dbX, dbF = np.zeros_like(X0), np.zeros_like(F0)
    _chk(lib().orc_archi_migrate(n, _cptr(kinds), dims, pops, X0, F0, seeds, gens, mtype, handling, off, src, wt,
                                 _cptr(sf), np.array([r[1] for r in sr], np.uint64), np.array([r[2] for r in sr]),
                                 _cptr(rf), np.array([r[1] for r in rr], np.uint64), np.array([r[2] for r in rr]),
                                 mseeds, rounds, int(stable), Xo, Fo, dbc, dbX, dbF))
    out, db = [], []
    xo = fo = 0
    for k in range(n):
        p, nx = int(pops[k]), nxs[k]
        out.append((Xo[xo:xo + p * nx].reshape(p, nx).copy(), Fo[fo:fo + p].copy()))
        c = int(dbc[k])
        db.append((dbX[xo:xo + c * nx].reshape(c, nx).copy(), dbF[fo:fo + c].copy()))
        xo += p * nx
        fo += p
    return out, db
