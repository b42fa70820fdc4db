"""ctypes binding of libpagmo_hip.so (include/pagmo_hip/pgh.h).

Device entry points take torch tensors that live on the GPU (PyTorch supplies
device memory and the stream; the computation is the library's HIP kernels).
Host entry points take numpy arrays.  Nothing here computes a result itself:
if the library is missing the import of this module fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PGH_LIB names another build of the same sources (the host-sanitized one of
# tests/test_sanitize_cpu.py)
LIB_PATH = os.environ.get("PGH_LIB") or os.path.join(HERE, "libpagmo_hip.so")

OK, INVALID_ARGUMENT, OVERFLOW, UNSUPPORTED, DEVICE_ERROR, NO_DEVICE, OUT_OF_MEMORY = range(7)
ROSENBROCK, RASTRIGIN, ACKLEY, SCHWEFEL, GRIEWANK, LENNARD_JONES, MO_SPHERE, ZDT, DTLZ = range(1, 10)
ARCHI_BATCHED, ARCHI_PER_ISLAND = 0, 1
ARCHI_STEP_ONE, ARCHI_STEP_SPLIT = 0, 1

# every symbol include/pagmo_hip/pgh.h declares
EXPORTS = [
    "pgh_last_error", "pgh_abi_version", "pgh_device_available", "pgh_problem_dims", "pgh_supported",
    "pgh_bfe_eval", "pgh_bfe_eval_host", "pgh_batch_gradient", "pgh_batch_gradient_host",
    "pgh_fnds_create", "pgh_fnds_thread_workspace", "pgh_fnds_destroy", "pgh_fnds_run", "pgh_fnds_run_host", "pgh_fnds_sizes",
    "pgh_fnds_copy", "pgh_fnds_set_peel_blocks", "pgh_fnds_tune", "pgh_lj_set_groups", "pgh_lj_tune",
    "pgh_crowding_distance", "pgh_crowding_distance_host", "pgh_sort_population_mo", "pgh_sort_population_mo_host",
    "pgh_select_best_N_mo", "pgh_select_best_N_mo_host",
    "pgh_local_descent_work_len", "pgh_local_descent", "pgh_local_descent_host", "pgh_local_descent_noalloc",
    "pgh_mt19937_generate", "pgh_batch_random_dv",
    "pgh_archi_create", "pgh_archi_create_seeded", "pgh_archi_destroy", "pgh_archi_evolve", "pgh_archi_get",
    "pgh_archi_set_trace",
    "pgh_archi_get_trace", "pgh_archi_set_migration", "pgh_archi_get_migrants", "pgh_pso_gen_evolve_host",
    "pgh_archi_set_mapping", "pgh_archi_get_mapping", "pgh_archi_tune",
]


class PghError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class pgh_udp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("prob_id", C.c_uint32), ("dim", C.c_uint64), ("nobj", C.c_uint64),
                ("centers", C.c_void_p), ("alpha", C.c_uint64), ("shift", C.c_void_p)]


class pgh_pso_params(C.Structure):
    _fields_ = [("gen", C.c_uint32), ("variant", C.c_uint32), ("omega", C.c_double), ("eta1", C.c_double),
                ("eta2", C.c_double), ("max_vel", C.c_double), ("neighb_type", C.c_uint32),
                ("neighb_param", C.c_uint32), ("memory", C.c_uint32), ("seed", C.c_uint32)]


class pgh_island_spec(C.Structure):
    _fields_ = [("udp", pgh_udp), ("pop", C.c_uint64), ("algo", pgh_pso_params)]


class pgh_archi_mapping(C.Structure):
    _fields_ = [("form", C.c_uint32), ("lj_lds", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("pad", C.c_uint32)]


class pgh_migr_rate(C.Structure):
    _fields_ = [("fractional", C.c_int32), ("pad", C.c_uint32), ("count", C.c_uint64), ("fraction", C.c_double)]


MIGRATION_P2P, MIGRATION_BROADCAST = 0, 1
MIGRANTS_PRESERVE, MIGRANTS_EVICT = 0, 1


def migr_rate(r) -> pgh_migr_rate:
    """A select_best / fair_replace rate as pagmo's constructors take it: an
    int is an absolute count, a float a fraction of the population."""
    if isinstance(r, (float, np.floating)):
        return pgh_migr_rate(1, 0, 0, float(r))
    return pgh_migr_rate(0, 0, int(r), 0.0)


def ring_connections(n: int, w: float = 1.0):
    """topology::get_connections of ring(n, w) for every island: the edges
    ring::push_back adds and removes (src/topologies/ring.cpp:61-100), each
    island's in-edges in the order they were added (base_bgl_topology's
    bidirectional vecS graph, topology.hpp:109-115).  A list of
    [(source, weight), ...] per island."""
    inbound = []
    for size in range(1, n + 1):
        inbound.append([])
        if size == 2:
            edges_add, edges_rm = [(0, 1), (1, 0)], []
        elif size == 3:
            edges_add, edges_rm = [(1, 2), (2, 1), (2, 0), (0, 2)], []
        elif size > 3:
            edges_rm = [(size - 2, 0), (0, size - 2)]
            edges_add = [(size - 2, size - 1), (size - 1, size - 2), (0, size - 1), (size - 1, 0)]
        else:
            edges_add, edges_rm = [], []
        for a, b in edges_rm:
            inbound[b].remove(a)
        for a, b in edges_add:
            inbound[b].append(a)
    return [[(a, w) for a in c] for c in inbound]


def fully_connected_connections(n: int, w: float = 1.0):
    """fully_connected(n, w): every other island in ascending order
    (src/topologies/fully_connected.cpp:86-116)."""
    return [[(j, w) for j in range(n) if j != i] for i in range(n)]


def pso_params(gen=1, omega=0.7298, eta1=2.05, eta2=2.05, max_vel=0.5, variant=5, neighb_type=2, neighb_param=4,
               memory=False, seed=0) -> pgh_pso_params:
    """pso_gen's constructor arguments and defaults (pso_gen.hpp:109-111)."""
    return pgh_pso_params(gen, variant, omega, eta1, eta2, max_vel, neighb_type, neighb_param, int(memory), seed)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m pagmo2_amd.build`")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, u64, u32, ci = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
    udp_p = C.POINTER(pgh_udp)
    sigs = {
        "pgh_abi_version": [],
        "pgh_device_available": [],
        "pgh_problem_dims": [udp_p, C.POINTER(u64), C.POINTER(u64)],
        "pgh_supported": [udp_p],
        "pgh_bfe_eval": [udp_p, vp, u64, u64, vp, u64, vp],
        "pgh_bfe_eval_host": [udp_p, vp, u64, vp, u64],
        "pgh_batch_gradient": [udp_p, vp, u64, u64, vp, vp, vp],
        "pgh_batch_gradient_host": [udp_p, vp, u64, vp, u64, vp],
        "pgh_local_descent": [udp_p, vp, u64, u64, vp, vp, vp, u32, C.c_double, vp, u64, vp],
        "pgh_local_descent_host": [udp_p, vp, u64, vp, u64, vp, vp, u32, C.c_double],
        "pgh_local_descent_noalloc": [udp_p, vp, u64, u64, vp, vp, vp, u32, C.c_double, vp],
        "pgh_fnds_create": [C.POINTER(vp)],
        "pgh_fnds_thread_workspace": [C.POINTER(vp)],
        "pgh_fnds_destroy": [vp],
        "pgh_fnds_run": [vp, vp, u64, u64, vp],
        "pgh_fnds_run_host": [vp, vp, u64, u64],
        "pgh_fnds_sizes": [vp, C.POINTER(u64), C.POINTER(u64)],
        "pgh_fnds_copy": [vp, vp, vp, vp, vp, vp, vp, ci],
        "pgh_fnds_set_peel_blocks": [vp, u32],
        "pgh_fnds_tune": [vp, vp, u64, u64, u32, vp, C.POINTER(u32)],
        "pgh_lj_set_groups": [u64, ci, u32],
        "pgh_lj_tune": [udp_p, vp, u64, vp, vp, u32, vp, C.POINTER(u32)],
        "pgh_crowding_distance": [vp, vp, u64, u64, vp, vp],
        "pgh_crowding_distance_host": [vp, u64, u64, vp],
        "pgh_sort_population_mo": [vp, vp, u64, u64, vp, vp],
        "pgh_sort_population_mo_host": [vp, u64, u64, vp],
        "pgh_select_best_N_mo": [vp, vp, u64, u64, u64, vp, C.POINTER(u64), vp],
        "pgh_select_best_N_mo_host": [vp, u64, u64, u64, vp, C.POINTER(u64)],
        "pgh_mt19937_generate": [u32, u64, u64, vp, vp],
        "pgh_batch_random_dv": [udp_p, u64, u32, vp, vp],
        "pgh_archi_create": [C.POINTER(vp), C.POINTER(pgh_island_spec), u64, C.POINTER(vp), C.POINTER(vp)],
        "pgh_archi_create_seeded": [C.POINTER(vp), C.POINTER(pgh_island_spec), u64, C.POINTER(u32)],
        "pgh_archi_destroy": [vp],
        "pgh_archi_evolve": [vp, ci, vp],
        "pgh_archi_get": [vp, u64, vp, vp],
        "pgh_archi_set_trace": [vp, ci],
        "pgh_archi_get_trace": [vp, u64, vp],
        "pgh_archi_set_mapping": [vp, C.POINTER(pgh_archi_mapping)],
        "pgh_archi_get_mapping": [vp, C.POINTER(pgh_archi_mapping)],
        "pgh_archi_tune": [vp, u32, C.POINTER(pgh_archi_mapping)],
        "pgh_archi_set_migration": [vp, ci, ci, vp, vp, vp, C.POINTER(pgh_migr_rate), C.POINTER(pgh_migr_rate),
                                    C.POINTER(u32)],
        "pgh_archi_get_migrants": [vp, u64, C.POINTER(u64), vp, vp],
        "pgh_pso_gen_evolve_host": [udp_p, C.POINTER(pgh_pso_params), vp, vp, u64, u64],
    }
    for name, args in sigs.items():
        if not hasattr(L, name):
            continue  # reported by missing_exports()
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ci
    L.pgh_last_error.argtypes = []
    L.pgh_last_error.restype = C.c_char_p
    L.pgh_local_descent_work_len.argtypes = [u64, u64]
    L.pgh_local_descent_work_len.restype = u64


def missing_exports():
    L = lib()
    return [n for n in EXPORTS if not hasattr(L, n)]


def check(status: int) -> None:
    if status != OK:
        raise PghError(status, lib().pgh_last_error().decode())


def make_udp(kind: int, dim: int, centers=None, nobj: int = 1, prob_id: int = 0, alpha: int = 0,
             shift=None) -> pgh_udp:
    """centers: a device tensor (device entry points) or a numpy array (host).
    ZDT: dim = the constructor's param, prob_id 1..6; DTLZ: dim, nobj = fdim,
    prob_id 1..7, alpha.  shift: a translation vector (numpy, host memory;
    the caller keeps it alive during the call)."""
    ptr = None
    if centers is not None:
        ptr = centers.data_ptr() if hasattr(centers, "data_ptr") else centers.ctypes.data
    sp = None if shift is None else shift.ctypes.data
    return pgh_udp(kind, prob_id, dim, nobj, ptr, alpha, sp)


def dims(kind: int, dim: int, nobj: int = 1, prob_id: int = 0, alpha: int = 0):
    nx, nf = C.c_uint64(), C.c_uint64()
    u = make_udp(kind, dim, None, nobj, prob_id, alpha)
    if kind == MO_SPHERE:
        u.centers = 1  # only the shape is resolved here
    check(lib().pgh_problem_dims(C.byref(u), C.byref(nx), C.byref(nf)))
    return nx.value, nf.value


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def _dptr(t):
    return None if t is None else t.data_ptr()


def _hptr(a):
    return None if a is None else a.ctypes.data


# ---- device entry points (torch tensors on the GPU) -----------------------

def bfe_eval(kind, dim, dvs, fvs, centers=None, nobj=1, stream=None, prob_id=0, alpha=0):
    u = make_udp(kind, dim, centers, nobj, prob_id, alpha)
    n, nx = dvs.shape
    check(lib().pgh_bfe_eval(C.byref(u), dvs.data_ptr(), n, nx, fvs.data_ptr(), fvs.shape[1], _stream(stream)))


def batch_gradient(kind, dim, dvs, grad, fvs=None, stream=None):
    u = make_udp(kind, dim)
    n, nx = dvs.shape
    check(lib().pgh_batch_gradient(C.byref(u), dvs.data_ptr(), n, nx, grad.data_ptr(), _dptr(fvs), _stream(stream)))


def lj_set_groups(atoms, grad, groups):
    check(lib().pgh_lj_set_groups(atoms, int(bool(grad)), groups))


def lj_tune(atoms, dvs, fvs=None, grad=None, reps=10, stream=None):
    """Times Lennard-Jones individuals-per-workgroup values on dvs; keeps and
    returns the fastest (energy kernel when grad is None)."""
    u = make_udp(LENNARD_JONES, atoms)
    out = C.c_uint32()
    check(lib().pgh_lj_tune(C.byref(u), dvs.data_ptr(), dvs.shape[0], _dptr(fvs), _dptr(grad), reps, _stream(stream),
                            C.byref(out)))
    return out.value


def local_descent_work_len(n, nx):
    return lib().pgh_local_descent_work_len(n, nx)


def local_descent(kind, dim, dvs, fvs, iters, step0, lb=None, ub=None, work=None, stream=None):
    """Memetic local step on device tensors: dvs (n x nx) in place, fvs (n)
    out; lb/ub device tensors of nx or None; work is allocated when None."""
    u = make_udp(kind, dim)
    n, nx = dvs.shape
    need = local_descent_work_len(n, nx)
    if work is None:
        import torch
        work = torch.empty(max(need, 1), dtype=torch.float64, device=dvs.device)
    check(lib().pgh_local_descent(C.byref(u), dvs.data_ptr(), n, nx, fvs.data_ptr(), _dptr(lb), _dptr(ub), iters,
                                  step0, work.data_ptr(), work.numel(), _stream(stream)))


# doubles of pgh_local_descent_noalloc's device array (PGH_LOCAL_DESCENT_NOALLOC_MAX_WORK)
LOCAL_DESCENT_NOALLOC_MAX_WORK = 1 << 23


def local_descent_noalloc(kind, dim, dvs, fvs, iters, step0, lb=None, ub=None, stream=None):
    """local_descent with the work buffer in the library's own device array
    (nothing allocated; calls overlapping on one GPU race)."""
    u = make_udp(kind, dim)
    n, nx = dvs.shape
    check(lib().pgh_local_descent_noalloc(C.byref(u), dvs.data_ptr(), n, nx, fvs.data_ptr(), _dptr(lb), _dptr(ub),
                                          iters, step0, _stream(stream)))


def mt19937_generate(seed, n, out, skip=0, stream=None):
    check(lib().pgh_mt19937_generate(seed, skip, n, out.data_ptr(), _stream(stream)))


def batch_random_dv(kind, dim, n, seed, out, stream=None):
    u = make_udp(kind, dim)
    if kind == MO_SPHERE:
        u.centers = 1
        u.nobj = 1
    check(lib().pgh_batch_random_dv(C.byref(u), n, seed, out.data_ptr(), _stream(stream)))


class Fnds:
    """fast_non_dominated_sorting workspace (device-resident results)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().pgh_fnds_create(C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pgh_fnds_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process end reclaims device memory; no HIP
        # call is made from the finalizer then
        if not _finalizing():
            self.close()

    def run(self, objs, stream=None):
        n, m = objs.shape
        check(lib().pgh_fnds_run(self.h, objs.data_ptr(), n, m, _stream(stream)))

    def set_peel_blocks(self, blocks):
        check(lib().pgh_fnds_set_peel_blocks(self.h, blocks))

    def tune(self, objs, reps=5, stream=None):
        """Times the peel grids on objs, keeps the fastest; returns it (0 when
        there is nothing to tune).  The workspace then holds objs' result."""
        n, m = objs.shape
        out = C.c_uint32()
        check(lib().pgh_fnds_tune(self.h, objs.data_ptr(), n, m, reps, _stream(stream), C.byref(out)))
        return out.value

    def run_host(self, objs: np.ndarray):
        objs = np.ascontiguousarray(objs, np.float64)
        n, m = objs.shape
        check(lib().pgh_fnds_run_host(self.h, objs.ctypes.data if objs.size else None, n, m))

    def sizes(self):
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().pgh_fnds_sizes(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def to_host(self, n):
        n_fronts, n_dom = self.sizes()
        r = dict(dom_count=np.empty(n, np.uint64), rank=np.empty(n, np.uint64),
                 front_offsets=np.empty(n_fronts + 1, np.uint64), front_items=np.empty(n, np.uint64),
                 dom_offsets=np.empty(n + 1, np.uint64), dom_items=np.empty(max(n_dom, 1), np.uint64))
        check(lib().pgh_fnds_copy(self.h, *(_hptr(r[k]) for k in ("dom_count", "rank", "front_offsets",
                                                                   "front_items", "dom_offsets", "dom_items")), 1))
        r["dom_items"] = r["dom_items"][:n_dom]
        return r


    # crowding distance and the orderings built on fnds (these replace the
    # workspace's fnds results and synchronise)
    def crowding_distance(self, points, out, stream=None):
        n, m = points.shape
        check(lib().pgh_crowding_distance(self.h, points.data_ptr(), n, m, out.data_ptr(), _stream(stream)))

    def sort_population_mo(self, objs, out, stream=None):
        n, m = objs.shape
        check(lib().pgh_sort_population_mo(self.h, objs.data_ptr(), n, m, out.data_ptr(), _stream(stream)))

    def select_best_N_mo(self, objs, N, out, stream=None):
        n, m = objs.shape
        ln = C.c_uint64()
        check(lib().pgh_select_best_N_mo(self.h, objs.data_ptr(), n, m, N, out.data_ptr(), C.byref(ln),
                                         _stream(stream)))
        return ln.value


# ---- host entry points (numpy) ------------------------------------------------

def bfe_eval_host(kind, dim, dvs: np.ndarray, centers=None, nobj=1, prob_id=0, alpha=0, shift=None) -> np.ndarray:
    dvs = np.ascontiguousarray(dvs, np.float64)
    c = None if centers is None else np.ascontiguousarray(centers, np.float64)
    t = None if shift is None else np.ascontiguousarray(shift, np.float64)
    u = make_udp(kind, dim, c, nobj, prob_id, alpha, t)
    nx, nf = dims(kind, dim, nobj, prob_id, alpha)
    out = np.empty((dvs.size // nx, nf))
    check(lib().pgh_bfe_eval_host(C.byref(u), dvs.ctypes.data, dvs.size, out.ctypes.data, out.size))
    return out


def batch_gradient_host(kind, dim, dvs: np.ndarray, with_fitness=False, shift=None):
    dvs = np.ascontiguousarray(dvs, np.float64)
    t = None if shift is None else np.ascontiguousarray(shift, np.float64)
    u = make_udp(kind, dim, shift=t)
    g = np.empty_like(dvs)
    nx = dims(kind, dim)[0]
    f = np.empty(dvs.size // nx) if with_fitness else None
    check(lib().pgh_batch_gradient_host(C.byref(u), dvs.ctypes.data, dvs.size, g.ctypes.data, g.size, _hptr(f)))
    return (g, f) if with_fitness else g


def local_descent_host(kind, dim, dvs: np.ndarray, iters, step0, lb=None, ub=None):
    """Memetic local step on host arrays; returns (x, f) (copies)."""
    x = np.array(dvs, np.float64, order="C", copy=True)
    u = make_udp(kind, dim)
    n = x.shape[0] if x.ndim == 2 else 0
    f = np.empty(max(n, 1))
    lbv = None if lb is None else np.ascontiguousarray(lb, np.float64)
    ubv = None if ub is None else np.ascontiguousarray(ub, np.float64)
    check(lib().pgh_local_descent_host(C.byref(u), x.ctypes.data, x.size, f.ctypes.data, n, _hptr(lbv), _hptr(ubv),
                                       iters, step0))
    return x, f[:n]


def pso_gen_evolve_host(kind, dim, X, F, prm: pgh_pso_params):
    X = np.array(X, np.float64, order="C")
    F = np.array(F, np.float64, order="C").reshape(-1)
    u = make_udp(kind, dim)
    check(lib().pgh_pso_gen_evolve_host(C.byref(u), C.byref(prm), X.ctypes.data, F.ctypes.data, X.shape[0],
                                        X.shape[1]))
    return X, F


class Archipelago:
    """Independent pso_gen islands evolved on the device."""

    def __init__(self, islands, trace=False):
        """islands: list of dicts {kind, dim, X0 (pop x nx), F0 (pop), algo (pgh_pso_params)}."""
        n = len(islands)
        specs = (pgh_island_spec * n)()
        self._keep = []
        xs = (C.c_void_p * n)()
        fs = (C.c_void_p * n)()
        self.shapes = []
        for k, isl in enumerate(islands):
            X0 = np.ascontiguousarray(isl["X0"], np.float64)
            F0 = np.ascontiguousarray(isl["F0"], np.float64).reshape(-1)
            self._keep += [X0, F0]
            specs[k].udp = make_udp(isl["kind"], isl["dim"])
            specs[k].pop = X0.shape[0]
            specs[k].algo = isl["algo"]
            xs[k] = X0.ctypes.data
            fs[k] = F0.ctypes.data
            self.shapes.append(X0.shape)
        h = C.c_void_p()
        check(lib().pgh_archi_create(C.byref(h), specs, n, xs, fs))
        self.h = h
        self.gens = [isl["algo"].gen for isl in islands]
        if trace:
            check(lib().pgh_archi_set_trace(self.h, 1))

    @classmethod
    def seeded(cls, islands, trace=False):
        """islands: list of dicts {kind, dim, pop, seed, algo}; each initial
        population is drawn (std::mt19937(seed)) and evaluated on the device."""
        self = cls.__new__(cls)
        n = len(islands)
        specs = (pgh_island_spec * n)()
        seeds = (C.c_uint32 * n)()
        self._keep = []
        self.shapes = []
        for k, isl in enumerate(islands):
            specs[k].udp = make_udp(isl["kind"], isl["dim"])
            specs[k].pop = isl["pop"]
            specs[k].algo = isl["algo"]
            seeds[k] = isl["seed"]
            self.shapes.append((isl["pop"], dims(isl["kind"], isl["dim"])[0]))
        h = C.c_void_p()
        check(lib().pgh_archi_create_seeded(C.byref(h), specs, n, seeds))
        self.h = h
        self.gens = [isl["algo"].gen for isl in islands]
        if trace:
            check(lib().pgh_archi_set_trace(self.h, 1))
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().pgh_archi_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # as for Fnds: nothing is released while the interpreter finalizes
        if not _finalizing():
            self.close()

    def evolve(self, mode=ARCHI_BATCHED, stream=None):
        check(lib().pgh_archi_evolve(self.h, mode, _stream(stream) if stream is not None else None))

    def set_mapping(self, form=ARCHI_STEP_ONE, lj_lds=0, blocks_per_cu=0):
        m = pgh_archi_mapping(form, lj_lds, blocks_per_cu, 0)
        check(lib().pgh_archi_set_mapping(self.h, C.byref(m)))

    def mapping(self):
        m = pgh_archi_mapping()
        check(lib().pgh_archi_get_mapping(self.h, C.byref(m)))
        return dict(form=m.form, lj_lds=m.lj_lds, blocks_per_cu=m.blocks_per_cu)

    def tune(self, reps=1):
        """Times the candidate launch mappings from the current state (which it
        restores) and keeps the fastest; returns it."""
        m = pgh_archi_mapping()
        check(lib().pgh_archi_tune(self.h, reps, C.byref(m)))
        return dict(form=m.form, lj_lds=m.lj_lds, blocks_per_cu=m.blocks_per_cu)

    def get(self, k):
        pop, nx = self.shapes[k]
        X = np.empty((pop, nx))
        F = np.empty(pop)
        check(lib().pgh_archi_get(self.h, k, X.ctypes.data, F.ctypes.data))
        return X, F

    def trace(self, k):
        pop, _ = self.shapes[k]
        T = np.empty((self.gens[k], pop))
        check(lib().pgh_archi_get_trace(self.h, k, T.ctypes.data))
        return T

    def set_migration(self, conns, select_rate=1, replace_rate=1, seeds=None, mtype=MIGRATION_P2P,
                      handling=MIGRANTS_PRESERVE):
        """conns: per island, [(source, weight), ...] in get_connections
        order (ring_connections, fully_connected_connections), or None to turn
        migration off.  Rates: one for all islands or one per island (int:
        absolute, float: fraction); seeds: the migration engines' seeds."""
        n = len(self.shapes)
        if conns is None:
            check(lib().pgh_archi_set_migration(self.h, mtype, handling, None, None, None, None, None, None))
            return
        off = np.zeros(n + 1, np.uint64)
        src, wt = [], []
        for k, c in enumerate(conns):
            off[k + 1] = off[k] + len(c)
            src += [a for a, _ in c]
            wt += [w for _, w in c]
        src = np.array(src + [0], np.uint64)
        wt = np.array(wt + [0.0], np.float64)
        per = lambda r: list(r) if isinstance(r, (list, tuple)) else [r] * n
        sel = (pgh_migr_rate * n)(*[migr_rate(r) for r in per(select_rate)])
        rep = (pgh_migr_rate * n)(*[migr_rate(r) for r in per(replace_rate)])
        sd = (C.c_uint32 * n)(*(seeds if seeds is not None else range(n)))
        self._migr_keep = [off, src, wt]
        check(lib().pgh_archi_set_migration(self.h, mtype, handling, off.ctypes.data, src.ctypes.data,
                                            wt.ctypes.data, sel, rep, sd))

    def migrants(self, k):
        """Island k's entry of the migrants database: (X, F), best first."""
        _, nx = self.shapes[k]
        cnt = C.c_uint64()
        check(lib().pgh_archi_get_migrants(self.h, k, C.byref(cnt), None, None))
        X = np.empty((cnt.value, nx))
        F = np.empty(cnt.value)
        check(lib().pgh_archi_get_migrants(self.h, k, C.byref(cnt), X.ctypes.data, F.ctypes.data))
        return X, F


def _mat(objs):
    objs = np.ascontiguousarray(objs, np.float64)
    if objs.ndim != 2:
        objs = objs.reshape(len(objs), -1)
    return objs, objs.shape[0], objs.shape[1]


def crowding_distance_host(points) -> np.ndarray:
    pts, n, m = _mat(points)
    out = np.empty(max(n, 1))
    check(lib().pgh_crowding_distance_host(_hptr(pts) if pts.size else None, n, m, _hptr(out)))
    return out[:n]


def sort_population_mo_host(objs) -> np.ndarray:
    pts, n, m = _mat(objs)
    out = np.empty(max(n, 1), np.uint64)
    check(lib().pgh_sort_population_mo_host(_hptr(pts) if pts.size else None, n, m, _hptr(out)))
    return out[:n]


def select_best_N_mo_host(objs, N) -> np.ndarray:
    pts, n, m = _mat(objs)
    out = np.empty(max(n, 1), np.uint64)
    ln = C.c_uint64()
    check(lib().pgh_select_best_N_mo_host(_hptr(pts) if pts.size else None, n, m, N, _hptr(out), C.byref(ln)))
    return out[: ln.value]
