"""Benchmark of the MI355X batch fitness evaluator (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extra] [--no-cpu]

Headline workload: separable bulk evaluation -- Rastrigin and Ackley,
n = 1000, population 65536, FP64, decision vectors drawn on the device from
std::mt19937(42) in pagmo's batch_random_decision_vector order (Rastrigin in
[-5.12, 5.12], Ackley in [-15, 30]).  One step = one bfe call per function on
the HBM-resident matrices.  value = fitness evaluations per second.

roofline: the dominant kernel's algorithmic bytes (8*n + 8 per evaluation)
over its average duration, timed with HIP events on the launch stream.
Defaults: 50 warm-up and 200 timed steps (about 45 ms of GPU time).  Launched
back to back, these kernels' durations vary by up to 10% over tens of launches
(the device's power management: serialised launches and a plain streaming read
of the same matrix show no such variation), so the timed region is long enough
to average over it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fitness+grad evals/s on 1 MI355X vs per-island launches; % HBM/FP64 peak"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
POP, NX = 65536, 1000
SEP_LANES = {"rastrigin": 64, "ackley": 32}  # lanes per individual the library uses at n = 1000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensure_built():
    import pagmo2_amd.build as B
    if not os.path.exists(B.LIB):
        B.build()


def cores():
    """Host threads for the CPU baselines: the cores this process may run on,
    at most 16 (a one-GPU box's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return dict(nproc=os.cpu_count(), affinity=affinity, model=model)


class EventTimer:
    """HIP events on the stream the kernels are launched on."""

    def __init__(self, torch, n):
        self.ev = [torch.cuda.Event(enable_timing=True) for _ in range(n)]

    def mark(self, i):
        self.ev[i].record()

    def ms(self, i, j):
        return self.ev[i].elapsed_time(self.ev[j])


def bench_separable(P, torch, steps, warmup):
    dev = "cuda:0"
    xr = torch.empty((POP, NX), dtype=torch.float64, device=dev)
    xa = torch.empty((POP, NX), dtype=torch.float64, device=dev)
    P.batch_random_dv(P.RASTRIGIN, NX, POP, 42, xr)
    P.batch_random_dv(P.ACKLEY, NX, POP, 42, xa)
    fr = torch.empty((POP, 1), dtype=torch.float64, device=dev)
    fa = torch.empty((POP, 1), dtype=torch.float64, device=dev)
    for _ in range(warmup):
        P.bfe_eval(P.RASTRIGIN, NX, xr, fr)
        P.bfe_eval(P.ACKLEY, NX, xa, fa)
    torch.cuda.synchronize()
    # per-call GPU time (one k_separable launch per call): HIP events on the
    # launch stream around the K rastrigin calls and the K ackley calls of the
    # timed region, issued back to back per function (a step is still one call
    # of each); three events in all, so they add nothing to the calls
    tm = EventTimer(torch, 3)
    t0 = time.perf_counter()
    tm.mark(0)
    for _ in range(steps):
        P.bfe_eval(P.RASTRIGIN, NX, xr, fr)
    tm.mark(1)
    for _ in range(steps):
        P.bfe_eval(P.ACKLEY, NX, xa, fa)
    tm.mark(2)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ras_ms = tm.ms(0, 1) / steps
    ack_ms = tm.ms(1, 2) / steps
    ras_graph = graph_ms(torch, lambda: P.bfe_eval(P.RASTRIGIN, NX, xr, fr), steps)
    ack_graph = graph_ms(torch, lambda: P.bfe_eval(P.ACKLEY, NX, xa, fa), steps)
    return dict(wall=wall, ras_ms=ras_ms, ack_ms=ack_ms, ras_graph=ras_graph, ack_graph=ack_graph, xr=xr, xa=xa,
                fr=fr, fa=fa)


def cpu_baseline(O, sep):
    """thread_bfe restatement on this host, on the first 8192 rows of each
    matrix (a bounded sample of the same workload)."""
    rows = 8192
    n = cores()
    xr = sep["xr"][:rows].cpu().numpy()
    xa = sep["xa"][:rows].cpu().numpy()
    _, tr = O.bfe_timed(O.RASTRIGIN, NX, xr, n, 5)
    _, ta = O.bfe_timed(O.ACKLEY, NX, xa, n, 5)
    # the same on one thread, on a quarter of the rows
    r1 = rows // 4
    _, tr1 = O.bfe_timed(O.RASTRIGIN, NX, xr[:r1], 1, 3)
    _, ta1 = O.bfe_timed(O.ACKLEY, NX, xa[:r1], 1, 3)
    return dict(value=2 * rows / (tr + ta), unit="evals/s", cores=n, kind="port",
                value_1thread=2 * r1 / (tr1 + ta1), host=host_info(),
                sample=f"oracle thread_bfe restatement, rastrigin + ackley, {rows} x {NX} rows each of the same "
                       f"seed-42 matrices, median of 5 calls, {n} threads (the host's cores available to this "
                       f"process, at most 16: one GPU's share of its box); value_1thread: {r1} rows each, "
                       f"one thread, median of 3")


def uniform_on_device(P, torch, count, seed, a, b):
    """std::uniform_real_distribution<double>(a, b) over std::mt19937(seed),
    from words generated by the library (libstdc++'s canonical mapping)."""
    w = torch.empty(2 * count, dtype=torch.int32, device="cuda:0")
    P.mt19937_generate(seed, 2 * count, w)
    u = w.view(count, 2).to(torch.int64) & 0xFFFFFFFF
    c = (u[:, 0].double() + u[:, 1].double() * 4294967296.0) * 2.0 ** -64
    c = torch.where(c >= 1.0, torch.full_like(c, 1.0 - 2.0 ** -53), c)
    return c * (b - a) + a


def graph_ms(torch, fn, reps):
    """Average GPU time of fn() (a few launches on the current stream), from
    `reps` back-to-back replays of a captured graph between two events: for
    kernels of a few tens of microseconds the host could not keep the GPU
    busy with one ctypes call per launch."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    g.replay()
    torch.cuda.synchronize()
    tm = EventTimer(torch, 2)
    tm.mark(0)
    g.replay()
    tm.mark(1)
    torch.cuda.synchronize()
    return tm.ms(0, 1) / reps


def bench_lj(P, torch, O=None, reps=20):
    atoms, pop = 48, 16384
    nx = 3 * atoms - 6
    x = uniform_on_device(P, torch, pop * nx, 42, -3.0, 3.0).view(pop, nx)
    f = torch.empty((pop, 1), dtype=torch.float64, device="cuda:0")
    g = torch.empty((pop, nx), dtype=torch.float64, device="cuda:0")
    # default mapping, then the autotuner's choice of individuals per
    # workgroup (results are identical for every choice)
    P.lj_set_groups(atoms, 0, 0)
    P.lj_set_groups(atoms, 1, 0)
    default_fit_ms = graph_ms(torch, lambda: P.bfe_eval(P.LENNARD_JONES, atoms, x, f), reps)
    default_grad_ms = graph_ms(torch, lambda: P.batch_gradient(P.LENNARD_JONES, atoms, x, g, f), reps)
    fit_groups = P.lj_tune(atoms, x, f, None, reps=reps)
    grad_groups = P.lj_tune(atoms, x, f, g, reps=reps)
    fit_ms = graph_ms(torch, lambda: P.bfe_eval(P.LENNARD_JONES, atoms, x, f), reps)
    grad_ms = graph_ms(torch, lambda: P.batch_gradient(P.LENNARD_JONES, atoms, x, g, f), reps)
    P.lj_set_groups(atoms, 0, 0)
    P.lj_set_groups(atoms, 1, 0)
    pairs = atoms * (atoms - 1) // 2 * pop
    cpu = {}
    if O is not None:
        # the oracle's thread_bfe restatement (the reference's path) on 2048
        # rows, and its analytic gradient (no reference counterpart) on 512
        xs = x[:2048].cpu().numpy()
        _, t = O.bfe_timed(P.LENNARD_JONES, atoms, xs, cores(), 3)
        cpu["fitness_evals_per_s"] = 2048 / t
        cpu["fitness_cores"] = cores()
        t0 = time.perf_counter()
        O.gradient(P.LENNARD_JONES, atoms, xs[:512])
        cpu["gradient_evals_per_s_1thread"] = 512 / (time.perf_counter() - t0)
    return dict(workload="lennard_jones 48 atoms, pop 16384, FP64", cpu_baseline=cpu,
                tuned={"fitness_groups": fit_groups, "fitness_grad_groups": grad_groups,
                       "default_fitness_ms": default_fit_ms, "default_fitness_grad_ms": default_grad_ms},
                fitness_ms=fit_ms, fitness_grad_ms=grad_ms,
                fitness_evals_per_s=pop / (fit_ms * 1e-3), fitness_grad_evals_per_s=pop / (grad_ms * 1e-3),
                fitness_fp64_frac=14 * pairs / (fit_ms * 1e-3) / (FP64_PEAK_TFLOPS * 1e12),
                fitness_grad_fp64_frac=28 * pairs / (grad_ms * 1e-3) / (FP64_PEAK_TFLOPS * 1e12))


def bench_fnds(P, torch, O, with_cpu):
    rng = np.random.default_rng(42)
    n = 16384
    objs = rng.uniform(0, 1, (n, 3))
    dup = n // 10
    objs[n - dup:] = objs[rng.integers(0, n - dup, dup)]
    c = rng.uniform(0, 1, (2, 30))
    x = rng.uniform(0, 1, (8192, 30))
    xd = torch.from_numpy(x).to("cuda:0")
    f2 = torch.empty((8192, 2), dtype=torch.float64, device="cuda:0")
    P.bfe_eval(P.MO_SPHERE, 30, xd, f2, centers=torch.from_numpy(c).to("cuda:0"), nobj=2)
    out = {}
    for name, o in (("8192x2", f2), ("16384x3", torch.from_numpy(objs).to("cuda:0"))):
        ws = P.Fnds()
        reps = 5

        def timed():
            for _ in range(2):
                ws.run(o)
                ws.sizes()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                ws.run(o)
                ws.sizes()  # synchronises and writes the domination lists
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps * 1e3

        default_ms = timed()
        blocks = ws.tune(o, reps=3)  # the peel's workgroup count; same results for any
        out[name] = dict(ms=timed(), peel_workgroups=blocks, default_ms=default_ms)
        if with_cpu:  # cpu baseline: the oracle's single-threaded restatement, as the reference runs it
            out[name]["cpu_baseline_ms_1thread"] = O.fnds_timed(o.cpu().numpy(), 1) * 1e3
        ws.close()
    return out


def bench_mo_problems(P, torch, O, with_cpu, n=1 << 22, reps=20):
    """ZDT and DTLZ batch evaluation (pop 2^22, inputs uniform in the bounds,
    resident in HBM; every matrix exceeds the 256 MB last-level cache): GPU
    time per call from graph replays, algorithmic bytes 8 (nx + nf) per
    evaluation over 8 TB/s; the oracle's single-threaded restatement on 65536
    rows beside it."""
    out = {}
    for name, kind, pid, dim, nobj in (("zdt1_30", P.ZDT, 1, 30, 2), ("zdt4_30", P.ZDT, 4, 30, 2),
                                       ("dtlz2_12x3", P.DTLZ, 2, 12, 3)):
        nx, nf = P.dims(kind, dim, nobj=nobj, prob_id=pid, alpha=100)
        lb, ub = (np.zeros(nx), np.ones(nx)) if pid != 4 else (np.r_[0.0, np.full(nx - 1, -5.0)],
                                                                 np.r_[1.0, np.full(nx - 1, 5.0)])
        u = uniform_on_device(P, torch, n * nx, 7, 0.0, 1.0).view(n, nx)
        x = (u * torch.from_numpy(ub - lb).to("cuda:0") + torch.from_numpy(lb).to("cuda:0")).contiguous()
        f = torch.empty((n, nf), dtype=torch.float64, device="cuda:0")
        ms = graph_ms(torch, lambda: P.bfe_eval(kind, dim, x, f, nobj=nobj, prob_id=pid, alpha=100), reps)
        bytes_per = 8 * (nx + nf)
        r = dict(ms=ms, evals_per_s=n / (ms * 1e-3), GBps=n * bytes_per / (ms * 1e-3) / 1e9,
                 hbm_frac=n * bytes_per / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
        if with_cpu:
            xs = x[:65536].cpu().numpy()
            t0 = time.perf_counter()
            O.mo_fitness(kind, pid, dim, xs, nobj, 100)
            r["cpu_baseline_evals_per_s_1thread"] = 65536 / (time.perf_counter() - t0)
        out[name] = r
        del x, u, f
    return out


def bench_plain_method(P, O, reps=5):
    """The CPU-runnable configuration: pso_gen (variant 5, lbest, 100
    generations, seed 42) on rosenbrock n = 10 with 64 individuals seeded by
    population{rosenbrock{10}, 64, 42}; the device run against the oracle's
    single-threaded pso_gen restatement, trajectories compared bit for bit."""
    if O is None:
        return {}
    X0, F0 = O.population_init(P.ROSENBROCK, 10, 64, 42)
    prm = P.pso_params(gen=100, seed=42)
    P.pso_gen_evolve_host(P.ROSENBROCK, 10, X0, F0, prm)  # warm-up
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        X, F = P.pso_gen_evolve_host(P.ROSENBROCK, 10, X0, F0, prm)
        t.append(time.perf_counter() - t0)
    Xo, Fo = O.pso_gen_evolve(P.ROSENBROCK, 10, X0, F0, 100, 42)
    return dict(workload="pso_gen 100 generations, rosenbrock n=10, pop 64, seed 42", ms=sorted(t)[len(t) // 2] * 1e3,
                cpu_baseline_ms_1thread=O.pso_gen_timed(P.ROSENBROCK, 10, X0, F0, 100, 42, 1, reps) * 1e3,
                identical_to_oracle=bool(np.array_equal(X, Xo) and np.array_equal(F, Fo)))


def bench_gradients(P, torch, O, reps=20):
    """Batched gradients of rosenbrock and rastrigin (pop 65536, n = 1000):
    algorithmic bytes 16 n + 8 per evaluation (read x, write the gradient and
    the fitness) over 8 TB/s; the oracle's gradients on 256 rows beside."""
    out = {}
    for name, kind in (("rosenbrock", P.ROSENBROCK), ("rastrigin", P.RASTRIGIN)):
        x = torch.empty((POP, NX), dtype=torch.float64, device="cuda:0")
        P.batch_random_dv(kind, NX, POP, 7, x)
        g = torch.empty_like(x)
        f = torch.empty((POP, 1), dtype=torch.float64, device="cuda:0")
        ms = graph_ms(torch, lambda: P.batch_gradient(kind, NX, x, g, f), reps)
        bytes_per = 16 * NX + 8
        r = dict(ms=ms, GBps=POP * bytes_per / (ms * 1e-3) / 1e9,
                 hbm_frac=POP * bytes_per / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
        if O is not None:
            xs = x[:256].cpu().numpy()
            t0 = time.perf_counter()
            O.gradient(kind, NX, xs)
            r["cpu_baseline_evals_per_s_1thread"] = 256 / (time.perf_counter() - t0)
        out[name] = r
        del x, g, f
    return out


def bench_seeding(P, torch, O, reps=5):
    """batch_random_decision_vector of rastrigin, 65536 x 1000 on the device
    (std::mt19937 words in order, bit-exact), against the oracle on 8192 rows."""
    x = torch.empty((POP, NX), dtype=torch.float64, device="cuda:0")
    P.batch_random_dv(P.RASTRIGIN, NX, POP, 42, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        P.batch_random_dv(P.RASTRIGIN, NX, POP, 42, x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    r = dict(workload="batch_random_decision_vector, rastrigin 65536 x 1000", ms=ms,
             draws_per_s=POP * NX / (ms * 1e-3))
    if O is not None:
        t0 = time.perf_counter()
        ref = O.batch_random_dv(P.RASTRIGIN, NX, 8192, 42)
        r["cpu_baseline_draws_per_s_1thread"] = 8192 * NX / (time.perf_counter() - t0)
        r["first_8192_rows_identical"] = bool(np.array_equal(ref, x[:8192].cpu().numpy()))
    del x
    return r


def bench_host_path(P, torch, reps=3):
    """The drop-in path as pagmo drives it: the dvs matrix in pageable host
    memory, one pgh_bfe_eval_host call (copy in, kernels, copy out)."""
    xd = torch.empty((POP, NX), dtype=torch.float64, device="cuda:0")
    P.batch_random_dv(P.RASTRIGIN, NX, POP, 42, xd)
    x = xd.cpu().numpy()
    del xd
    P.bfe_eval_host(P.RASTRIGIN, NX, x)  # warm-up: staging buffers
    t0 = time.perf_counter()
    for _ in range(reps):
        P.bfe_eval_host(P.RASTRIGIN, NX, x)
    dt = (time.perf_counter() - t0) / reps
    return dict(workload="rastrigin n=1000, pop 65536, dvs in pageable host memory (pgh_bfe_eval_host)",
                ms=dt * 1e3, evals_per_s=POP / dt, host_to_device_GBps=POP * NX * 8 / dt / 1e9)


def bench_mo_sort(P, torch, O, with_cpu):
    """sort_population_mo (fnds + crowding of every front + the final sort) on
    the fnds workloads."""
    rng = np.random.default_rng(42)
    n = 16384
    objs = rng.uniform(0, 1, (n, 3))
    dup = n // 10
    objs[n - dup:] = objs[rng.integers(0, n - dup, dup)]
    c = rng.uniform(0, 1, (2, 30))
    x = rng.uniform(0, 1, (8192, 30))
    f2 = torch.empty((8192, 2), dtype=torch.float64, device="cuda:0")
    P.bfe_eval(P.MO_SPHERE, 30, torch.from_numpy(x).to("cuda:0"), f2, centers=torch.from_numpy(c).to("cuda:0"),
               nobj=2)
    out = {}
    for name, o in (("8192x2", f2), ("16384x3", torch.from_numpy(objs).to("cuda:0"))):
        ws = P.Fnds()
        ws.tune(o, reps=3)  # the fnds peel's workgroup count
        idx = torch.empty(o.shape[0], dtype=torch.int64, device="cuda:0")
        ws.sort_population_mo(o, idx)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            ws.sort_population_mo(o, idx)  # synchronises
        out[name] = dict(ms=(time.perf_counter() - t0) / reps * 1e3)
        if with_cpu:
            out[name]["cpu_baseline_ms_1thread"] = O.sort_population_mo_timed(o.cpu().numpy(), 1) * 1e3
        ws.close()
    return out


def bench_archipelago(P, torch, gens, O=None):
    from pagmo2_amd.workloads import archipelago_islands
    isl = archipelago_islands(gens=gens)
    res = {}
    finals = {}
    for mode, name in ((P.ARCHI_BATCHED, "batched"), (P.ARCHI_PER_ISLAND, "per_island")):
        t0 = time.perf_counter()
        a = P.Archipelago.seeded(isl)  # synchronous
        res[name + "_create_s"] = time.perf_counter() - t0
        if mode == P.ARCHI_BATCHED:
            # the launch mapping from the tuner (one timed evolve call per
            # candidate, from the initial state it then restores)
            t0 = time.perf_counter()
            res["mapping"] = a.tune(reps=1)
            res["tune_s"] = time.perf_counter() - t0
        torch.cuda.synchronize()
        # three evolve calls of `gens` generations each; the first two capture
        # the batched mode's graphs (the first call starts from X, later ones
        # from lbX), the third is timed
        t0 = time.perf_counter()
        a.evolve(mode)
        a.get(0)  # synchronises
        res[name + "_first_call_s"] = time.perf_counter() - t0
        a.evolve(mode)
        a.get(0)
        t0 = time.perf_counter()
        a.evolve(mode)
        a.get(0)
        res[name + "_s"] = time.perf_counter() - t0
        finals[name] = [a.get(k) for k in range(len(isl))]
        a.close()
    same = all(np.array_equal(finals["batched"][k][0], finals["per_island"][k][0])
               and np.array_equal(finals["batched"][k][1], finals["per_island"][k][1]) for k in range(len(isl)))
    if O is not None:
        # one generation of every island with the oracle's single-threaded
        # pso_gen (the reference runs one island per thread): the summed time
        # per generation, on the islands' initial populations
        t = 0.0
        for k, d in enumerate(isl):
            X0 = O.batch_random_dv(d["kind"], d["dim"], d["pop"], d["seed"])
            F0 = O.bfe(d["kind"], d["dim"], X0, threads=cores())
            t += O.pso_gen_timed(d["kind"], d["dim"], X0, F0, 1, d["seed"], 1, 1)
        res["cpu_baseline_s_per_generation_1thread_summed"] = t
    res.update(workload=f"64 uneven islands (pop logU[32,4096], dim logU[10,500], rosenbrock/rastrigin/"
                        f"lennard_jones round robin), pso_gen {gens} generations per evolve call (the third call "
                        f"timed); initial populations drawn and evaluated on the device",
               speedup=res["per_island_s"] / res["batched_s"], identical=bool(same),
               batched_us_per_generation=res["batched_s"] / gens * 1e6,
               per_island_us_per_generation=res["per_island_s"] / gens * 1e6)
    return res


def bench_migration(P, torch, O=None, n_islands=16, pop=1024, dim=100, gens=10, rounds=20):
    """A migrating archipelago: n_islands rastrigin(dim) islands of pop
    individuals in a ring (weight 1, p2p, preserve, select_best(4),
    fair_replace(4)); a round is [immigration, pso_gen(gens), emigration],
    archipelago::evolve(1).  Device seconds per round with and without
    migration on the same islands, one round against the oracle's round, and
    the oracle's round on one thread."""
    isl = [dict(kind=P.RASTRIGIN, dim=dim, pop=pop, seed=42 + k, algo=P.pso_params(gen=gens, seed=42 + k))
           for k in range(n_islands)]
    conns = P.ring_connections(n_islands)
    res = dict(workload=f"{n_islands} rastrigin({dim}) islands of {pop}, ring p2p preserve, select_best(4) / "
                        f"fair_replace(4), {gens} pso_gen generations per round")
    for name, on in (("with_migration", True), ("without_migration", False)):
        a = P.Archipelago.seeded(isl)
        X0 = [a.get(k) for k in range(n_islands)]
        if on:
            a.set_migration(conns, 4, 4, seeds=list(range(n_islands)))
        a.evolve()
        first = [a.get(k) for k in range(n_islands)]
        a.evolve()  # the steady-state graph
        a.get(0)
        t0 = time.perf_counter()
        for _ in range(rounds):
            a.evolve()
        a.get(0)
        res[name + "_s_per_round"] = (time.perf_counter() - t0) / rounds
        a.close()
        if on and O is not None:
            oi = [dict(kind=P.RASTRIGIN, dim=dim, X=X0[k][0], F=X0[k][1], seed=42 + k, gen=gens, sel=4, rep=4,
                       migr_seed=k) for k in range(n_islands)]
            t0 = time.perf_counter()
            ref, _ = O.archi_migrate(oi, conns, 0, 0, rounds=1)
            res["cpu_baseline_s_per_round_1thread"] = time.perf_counter() - t0
            res["first_round_matches_oracle"] = bool(all(
                np.array_equal(first[k][0], ref[k][0]) and np.allclose(first[k][1], ref[k][1], rtol=1e-12, atol=0)
                for k in range(n_islands)))
    res["migration_s_per_round"] = res["with_migration_s_per_round"] - res["without_migration_s_per_round"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--extras", default="plain_method,lennard_jones,gradients,seeding,fnds,mo_sort,mo_problems,"
                                        "archipelago,migration,host_path",
                    help="comma-separated extra sections to run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--archi-gens", type=int, default=200)
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 256 MiB device copy (the reference kernel for HBM byte counters)")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the extra sections (to profile them on their own); the line has no headline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the islands and individuals of this "
              f"path shard across GPUs without exchange)", file=sys.stderr)
        return 2
    ensure_built()
    import torch

    import pagmo2_amd as P
    O = None
    if not args.no_cpu:
        import oracle as O  # the CPU baselines only

    if args.copy_calib:
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.skip_headline:
        line = {"metric": METRIC, "skipped_headline": True}
    else:
        log("separable bulk: generating inputs and timing")
        sep = bench_separable(P, torch, args.steps, args.warmup)
        ms_per_step = sep["wall"] / args.steps * 1e3
        value = 2 * POP / (ms_per_step * 1e-3)
        bytes_per_launch = POP * (8 * NX + 8)
        dom_name, dom_ms = ("rastrigin", sep["ras_ms"]) if sep["ras_ms"] >= sep["ack_ms"] else ("ackley", sep["ack_ms"])
        achieved = bytes_per_launch / (dom_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=None, kernel=f"k_separable<{dom_name}, {SEP_LANES[dom_name]} lanes, aligned> (one bfe call)", kernel_ms=dom_ms,
                    other={"rastrigin_ms": sep["ras_ms"], "ackley_ms": sep["ack_ms"],
                           "rastrigin_graph_replay_ms": sep["ras_graph"], "ackley_graph_replay_ms": sep["ack_graph"],
                           "rastrigin_GBps": bytes_per_launch / (sep["ras_ms"] * 1e-3) / 1e9,
                           "ackley_GBps": bytes_per_launch / (sep["ack_ms"] * 1e-3) / 1e9})
        traffic_file = os.path.join(ROOT, "profiles", "separable_hbm_traffic.json")
        if os.path.exists(traffic_file):  # from rocprofv3 FETCH_SIZE / WRITE_SIZE runs, copy-kernel calibrated
            with open(traffic_file) as f:
                pmc = json.load(f)
            roof["traffic"] = pmc.get(dom_name, {}).get("traffic_bytes")
        line = {
            "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic: std::mt19937(42) draws on the device in batch_random_decision_vector order",
            "config": {"workload": "separable bulk evaluation: rastrigin + ackley, n=1000, pop=65536, one bfe call "
                                   "each per step", "n": NX, "pop": POP, "functions": ["rastrigin", "ackley"]},
            "roofline": roof,
        }
        if not args.no_cpu:
            log("cpu baseline")
            line["cpu_baseline"] = cpu_baseline(O, sep)
        del sep
        torch.cuda.empty_cache()
    if not args.no_extra:
        extra = {}
        for name, fn in (("plain_method", lambda: bench_plain_method(P, O)),
                         ("lennard_jones", lambda: bench_lj(P, torch, O)),
                         ("gradients", lambda: bench_gradients(P, torch, O)),
                         ("seeding", lambda: bench_seeding(P, torch, O)),
                         ("fnds", lambda: bench_fnds(P, torch, O, not args.no_cpu)),
                         ("mo_sort", lambda: bench_mo_sort(P, torch, O, not args.no_cpu)),
                         ("mo_problems", lambda: bench_mo_problems(P, torch, O, not args.no_cpu)),
                         ("archipelago", lambda: bench_archipelago(P, torch, args.archi_gens, O)),
                         ("migration", lambda: bench_migration(P, torch, O)),
                         ("host_path", lambda: bench_host_path(P, torch))):
            if name not in args.extras.split(","):
                continue
            log(f"extra: {name}")
            try:
                extra[name] = fn()
            except Exception as e:  # an extra section never hides the headline line
                extra[name] = {"error": repr(e)}
        line["extra"] = extra
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    # a zero status ends the interpreter normally rather than through SystemExit
    rc = main()
    if rc:
        sys.exit(rc)
