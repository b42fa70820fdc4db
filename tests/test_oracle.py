"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a swapped index, a wrong loop bound, a different draw order) fails it: the
reference's own known answers, closed forms at best_known(), invariants and
brute force on small inputs.  CPU only.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O

KIND = {"rosenbrock": O.ROSENBROCK, "rastrigin": O.RASTRIGIN, "ackley": O.ACKLEY, "schwefel": O.SCHWEFEL,
        "griewank": O.GRIEWANK, "lennard_jones": O.LENNARD_JONES}


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# --- reference known answers -------------------------------------------------

def test_fitness_kats(kats):
    for k in kats["fitness"]:
        f = O.fitness_one(KIND[k["kind"]], k["dim"], k["x"])
        if k.get("exact"):
            assert f[0] == k["f"][0], k["src"]
        else:
            assert rel(f[0], k["f"][0]) <= k["rel_tol"], (k["src"], f[0])


def test_gradient_kats(kats):
    for k in kats["gradient"]:
        g = O.gradient(KIND[k["kind"]], k["dim"], np.array([k["x"]]))[0]
        for gi, ei in zip(g, k["g"]):
            if "abs_tol" in k:
                assert abs(gi - ei) < k["abs_tol"], k["src"]
            else:
                assert rel(gi, ei) <= k["rel_tol"], k["src"]


def test_estimate_gradient_h_kat(kats):
    k = kats["estimate_gradient_h"]
    gh = O.estimate_gradient_kat(k["x"], k["dx"], True)
    for tol in k["rel_tol"]:
        for a, b in zip(gh, k["g"]):
            if b == 0:
                assert a == 0
            else:
                assert rel(a, b) <= tol
    # The low-order estimate of the same test (computed but unchecked there)
    # must at least agree to the finite-difference accuracy.
    g = O.estimate_gradient_kat(k["x"], 1e-8, False)
    np.testing.assert_allclose(g, k["g"], rtol=1e-6, atol=1e-9)


def test_pareto_dominance_kats(kats):
    for k in kats["pareto_dominance"]:
        assert O.pareto_dominance(k["a"], k["b"]) == k["r"], k["src"]
    with pytest.raises(ValueError):
        O.pareto_dominance([1, 2], [3, 4, 5])


def _fnds_lists(r):
    fo, fi = r["front_offsets"], r["front_items"]
    fronts = [list(map(int, fi[fo[i]:fo[i + 1]])) for i in range(len(fo) - 1)]
    do, di = r["dom_offsets"], r["dom_items"]
    dl = [list(map(int, di[do[i]:do[i + 1]])) for i in range(len(do) - 1)]
    return fronts, dl


def test_fnds_kats(kats):
    for k in kats["fnds"]:
        pts = np.array(k["points"], np.float64).reshape(len(k["points"]), -1)
        r = O.fnds(pts)
        fronts, dl = _fnds_lists(r)
        assert fronts == k["fronts"], k["src"]
        assert dl == k["dom_list"], k["src"]
        assert list(r["dom_count"]) == k["dom_count"], k["src"]
        assert list(r["rank"]) == k["rank"], k["src"]
    for k in kats["fnds_throws"]:
        pts = np.array(k["points"], np.float64).reshape(len(k["points"]), -1 if k["points"] else 0)
        with pytest.raises(O.OracleError, match="At least two points"):
            O.fnds(pts)


def test_mt19937_kat(kats):
    k = kats["mt19937"]
    assert int(O.mt19937_u32(k["seed"], 1, skip=k["index"] - 1)[0]) == k["value"]


# --- closed forms ------------------------------------------------------------

def test_best_known_closed_forms():
    assert O.fitness_one(O.ROSENBROCK, 10, np.ones(10))[0] == 0.0       # rosenbrock.cpp:98-101
    assert O.fitness_one(O.RASTRIGIN, 7, np.zeros(7))[0] == 0.0         # rastrigin.cpp:154-157
    assert abs(O.fitness_one(O.ACKLEY, 7, np.zeros(7))[0]) < 1e-15      # ackley.cpp:96-99
    assert O.fitness_one(O.GRIEWANK, 7, np.zeros(7))[0] == 0.0          # griewank.cpp:95-98
    assert abs(O.fitness_one(O.SCHWEFEL, 7, np.full(7, 420.9687))[0]) < 1e-3  # schwefel.cpp:88-91


def test_fitness_identities_numpy():
    """Independent float64 numpy evaluations of the textbook formulas."""
    rng = np.random.default_rng(1)
    x = rng.uniform(-5, 5, size=(20, 13))
    for row in x:
        ros = O.fitness_one(O.ROSENBROCK, 13, row)[0]
        ref = np.sum(100 * (row[:-1] ** 2 - row[1:]) ** 2 + (row[:-1] - 1) ** 2)
        assert rel(ros, ref) < 1e-13
        ras = O.fitness_one(O.RASTRIGIN, 13, row)[0]
        assert rel(ras, 10 * 13 + np.sum(row ** 2 - 10 * np.cos(2 * np.pi * row))) < 1e-12
        ack = O.fitness_one(O.ACKLEY, 13, row)[0]
        ref = (-20 * np.exp(-0.2 * np.sqrt(np.mean(row ** 2))) - np.exp(np.mean(np.cos(2 * np.pi * row)))
               + 20 + np.e)
        assert rel(ack, ref) < 1e-12
        sch = O.fitness_one(O.SCHWEFEL, 13, row)[0]
        assert rel(sch, 418.9828872724338 * 13 - np.sum(row * np.sin(np.sqrt(np.abs(row))))) < 1e-12
        gri = O.fitness_one(O.GRIEWANK, 13, row)[0]
        ref = np.sum(row ** 2) / 4000 - np.prod(np.cos(row / np.sqrt(np.arange(1, 14)))) + 1
        assert rel(gri, ref) < 1e-12


def _lj_positions(x):
    atoms = (len(x) + 6) // 3
    p = np.zeros((atoms, 3))
    p[1, 2] = x[0]
    p[2, 1:] = x[1:3]
    p[3:] = np.asarray(x[3:]).reshape(-1, 3)
    return p


def test_lennard_jones_brute_force():
    rng = np.random.default_rng(2)
    for atoms in (3, 4, 7, 13):
        x = rng.uniform(-3, 3, 3 * atoms - 6)
        p = _lj_positions(x)
        e = 0.0
        for i, j in itertools.combinations(range(atoms), 2):
            r2 = np.sum((p[i] - p[j]) ** 2)
            e += r2 ** -6 - r2 ** -3
        assert rel(O.fitness_one(O.LENNARD_JONES, atoms, x)[0], 4 * e) < 1e-12


def test_lennard_jones_coincident_atoms():
    # lennard_jones.cpp:81-83,89: dist == 0 -> DBL_MAX, final *4 -> +inf.
    x = np.array([0.0, 0.0, 1.0])  # atom 1 at the origin, on top of atom 0
    assert O.fitness_one(O.LENNARD_JONES, 3, x)[0] == math.inf


def lj_lattice(atoms, rng, spacing=1.1, jitter=0.1):
    """Decision vector of a jittered cubic lattice (pairs stay ~1 apart), in the
    encoding of lennard_jones.cpp:132-151 (atom 0 at the origin, atom 1 on z,
    atom 2 in the yz plane)."""
    grid = np.array(list(itertools.product(range(4), repeat=3)), np.float64)[:atoms] * spacing
    p = grid + rng.uniform(-jitter, jitter, grid.shape)
    p -= p[0]
    p[1, :2] = 0.0
    p[2, 0] = 0.0
    return np.concatenate([[p[1, 2]], p[2, 1:], p[3:].reshape(-1)])


def test_lennard_jones_gradient_fd():
    rng = np.random.default_rng(3)
    for atoms in (3, 5, 9, 20):
        x = lj_lattice(atoms, rng)
        g = O.gradient(O.LENNARD_JONES, atoms, x[None, :])[0]
        h = 1e-6
        fd = np.empty_like(x)
        for k in range(x.size):
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            fd[k] = (O.fitness_one(O.LENNARD_JONES, atoms, xp)[0] - O.fitness_one(O.LENNARD_JONES, atoms, xm)[0]) / (2 * h)
        assert np.linalg.norm(fd - g) / np.linalg.norm(g) < 1e-7
        gh = O.estimate_gradient(O.LENNARD_JONES, atoms, x, 1e-3, True)
        assert np.linalg.norm(gh - g) / np.linalg.norm(g) < 1e-7


def test_rosenbrock_rastrigin_gradient_vs_estimate():
    rng = np.random.default_rng(4)
    for kind in (O.ROSENBROCK, O.RASTRIGIN):
        x = rng.uniform(-2, 2, 9)
        g = O.gradient(kind, 9, x[None, :])[0]
        gh = O.estimate_gradient(kind, 9, x, 1e-2, True)
        assert np.linalg.norm(gh - g) / np.linalg.norm(g) < 1e-6


# --- fnds invariants -------------------------------------------------------

def _brute(points):
    n = len(points)
    dom = [[O.pareto_dominance(points[i], points[j]) for j in range(n)] for i in range(n)]
    count = [sum(dom[j][i] for j in range(n)) for i in range(n)]
    rank = [None] * n
    for _ in range(n):  # longest chain by relaxation
        for i in range(n):
            rs = [rank[j] for j in range(n) if dom[j][i]]
            if all(r is not None for r in rs):
                rank[i] = 1 + max(rs) if rs else 0
    lists = [[j for j in range(n) if dom[i][j]] for i in range(n)]
    return lists, count, rank


@pytest.mark.parametrize("m,dup", [(2, False), (3, True), (1, True), (4, False)])
def test_fnds_brute_force(m, dup):
    rng = np.random.default_rng(10 + m)
    pts = rng.integers(0, 6, size=(40, m)).astype(np.float64)  # many ties
    if dup:
        pts[30:] = pts[:10]
    r = O.fnds(pts)
    fronts, dl = _fnds_lists(r)
    lists, count, rank = _brute(pts)
    assert dl == lists
    assert list(map(int, r["dom_count"])) == count
    assert list(map(int, r["rank"])) == rank
    assert sorted(itertools.chain(*fronts)) == list(range(len(pts)))
    for f, items in enumerate(fronts):
        assert all(rank[i] == f for i in items)
    assert fronts[0] == sorted(fronts[0])


def test_fnds_nan_ordering():
    # custom_comparisons.hpp:54-100: NaN sorts after +inf, NaN == NaN.
    nan, inf = math.nan, math.inf
    assert O.pareto_dominance([inf, 0], [nan, 0])
    assert not O.pareto_dominance([nan, 0], [inf, 0])
    assert not O.pareto_dominance([nan, 1], [nan, 1])
    assert O.pareto_dominance([nan, 0], [nan, 1])


# --- bfe restatement ---------------------------------------------------------

@pytest.mark.parametrize("kind,dim", [(O.ROSENBROCK, 7), (O.RASTRIGIN, 9), (O.ACKLEY, 5), (O.SCHWEFEL, 6),
                                      (O.GRIEWANK, 8), (O.LENNARD_JONES, 6)])
def test_thread_bfe_equals_serial(kind, dim):
    # tests/thread_bfe.cpp:66-96: exact equality with per-dv fitness.
    nx, _, lb, ub = O.problem_info(kind, dim)
    dvs = O.batch_uniform(lb, ub, 101, 7)
    f1 = O.bfe(kind, dim, dvs, threads=1)
    f4 = O.bfe(kind, dim, dvs, threads=4)
    assert np.array_equal(f1, f4)
    for i in range(0, 101, 17):
        assert f1[i, 0] == O.fitness_one(kind, dim, dvs[i])[0]


def test_bfe_input_message(kats):
    msg = kats["bfe_messages"]["input_not_multiple"].format(n=7, nx=3)
    lib = O.lib()
    dvs = np.zeros(7)
    out = np.zeros(7)
    rc = lib.orc_bfe(O.ROSENBROCK, 3, None, 1, dvs, 1, 7, out, 1)
    assert rc != 0
    assert lib.orc_last_error().decode() == msg


def test_population_matches_batch_random(kats):
    # tests/population.cpp:399-451: serial ctor and bfe ctor draw the same dvs.
    X, F = O.population_init(O.ROSENBROCK, 10, 64, 42)
    B = O.batch_random_dv(O.ROSENBROCK, 10, 64, 42)
    assert np.array_equal(X, B)
    assert X.min() >= -5 and X.max() < 10
    assert np.array_equal(F[:, 0], O.bfe(O.ROSENBROCK, 10, X)[:, 0])


# --- RNG mapping ---------------------------------------------------------------

def canonical(u0, u1):
    """libstdc++ generate_canonical<double,53> over two mt19937 outputs
    (bits/random.tcc, GCC 11): (u0 + u1*2^32) / 2^64 rounded, clamped below 1."""
    r = (float(u0) + float(u1) * 4294967296.0) / 18446744073709551616.0
    return math.nextafter(1.0, 0.0) if r >= 1.0 else r


def test_uniform_real_mapping():
    u = O.mt19937_u32(42, 2000)
    d = O.uniform_real(42, -5.0, 10.0, 1000)
    for i in range(1000):
        assert d[i] == canonical(u[2 * i], u[2 * i + 1]) * 15.0 + -5.0


# --- pso_gen -----------------------------------------------------------------

def test_pso_gen_bfe_equals_serial():
    # tests/pso_gen.cpp:302-318: evolve with and without a bfe give equal results.
    X, F = O.population_init(O.ROSENBROCK, 10, 64, 42)
    a = O.pso_gen_evolve(O.ROSENBROCK, 10, X, F, 30, 42, use_bfe=True, threads=3)
    b = O.pso_gen_evolve(O.ROSENBROCK, 10, X, F, 30, 42, use_bfe=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("topo", [1, 2, 3, 4])
def test_pso_gen_determinism(variant, topo):
    # tests/pso_gen.cpp:154-240: the same seed gives the same result.
    X, F = O.population_init(O.RASTRIGIN, 6, 24, 3)
    a = O.pso_gen_evolve(O.RASTRIGIN, 6, X, F, 10, 5, variant=variant, neighb_type=topo)
    b = O.pso_gen_evolve(O.RASTRIGIN, 6, X, F, 10, 5, variant=variant, neighb_type=topo)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the memory only ever improves (pso_gen.cpp:443-449)
    assert np.all(a[1] <= F[:, 0])


def _pso_one_gen_python(X, F, seed, lb, ub, fit):
    """Independent single-generation pso_gen (variant 5, lbest radius 2)
    written from pso_gen.cpp:177-200,203-218,232-363,443-449."""
    pop, nx = X.shape
    u = O.mt19937_u32(seed, 4 * pop * nx + 2 * pop * nx)
    k = 0

    def draw(a, b):
        nonlocal k
        c = canonical(u[k], u[k + 1])
        k += 2
        return c * (b - a) + a

    maxv = [(ub[j] - lb[j]) * 0.5 for j in range(nx)]
    minv = [-1.0 * m for m in maxv]
    V = [[draw(minv[j], maxv[j]) for j in range(nx)] for _ in range(pop)]
    neigh = [[(p - 2) % pop, (p - 1) % pop, (p + 1) % pop, (p + 2) % pop] for p in range(pop)]
    Xn = [list(r) for r in X]
    for p in range(pop):
        b = neigh[p][0]
        for q in neigh[p][1:]:
            if F[q] <= F[b]:
                b = q
        for d in range(nx):
            r1 = draw(0.0, 1.0)
            r2 = draw(0.0, 1.0)
            V[p][d] = 0.7298 * (V[p][d] + 2.05 * r1 * (X[p][d] - X[p][d]) + 2.05 * r2 * (X[b][d] - X[p][d]))
    for p in range(pop):
        for d in range(nx):
            v = min(max(V[p][d], minv[d]), maxv[d]) if (V[p][d] > maxv[d] or V[p][d] < minv[d]) else V[p][d]
            nxv = Xn[p][d] + v
            if nxv < lb[d]:
                nxv = lb[d]
            elif nxv > ub[d]:
                nxv = ub[d]
            Xn[p][d] = nxv
    Xn = np.array(Xn)
    fn = np.array([fit(r) for r in Xn])
    better = fn <= F
    return np.where(better[:, None], Xn, X), np.where(better, fn, F)


def test_pso_gen_single_generation_independent():
    X, F = O.population_init(O.ROSENBROCK, 5, 12, 9)
    nx, _, lb, ub = O.problem_info(O.ROSENBROCK, 5)
    Xo, Fo = O.pso_gen_evolve(O.ROSENBROCK, 5, X, F, 1, 77)
    Xp, Fp = _pso_one_gen_python(X, F[:, 0], 77, lb, ub, lambda r: O.fitness_one(O.ROSENBROCK, 5, r)[0])
    assert np.array_equal(Xo, Xp)
    assert np.array_equal(Fo, Fp)


# Pure-Python fitness in the reference's operation order (rosenbrock.cpp:59-66,
# rastrigin.cpp:62-72); math.cos is the same glibc routine the reference calls.
def _py_rosenbrock(x):
    f = 0.0
    for i in range(len(x) - 1):
        t = x[i] * x[i] - x[i + 1]
        f = f + ((100.0 * t) * t + (x[i] - 1.0) * (x[i] - 1.0))
    return f


def _py_rastrigin(x):
    omega = 2.0 * math.pi
    f = 0.0
    for v in x:
        f += v * v - 10.0 * math.cos(omega * v)
    return f + 10.0 * len(x)


def _le_f(a, b):
    # less_than_f(a, b) || equal_to_f(a, b), custom_comparisons.hpp:54-100
    if math.isnan(a):
        return math.isnan(b)
    if math.isnan(b):
        return True
    return a <= b


class _PsoPython:
    """Independent pso_gen (variant 5, lbest) over several generations and
    evolve calls, written from pso_gen.cpp:120-530 and :593-699.

    The random words come from numpy's MT19937 (RandomState with a 32-bit
    seed seeds like std::mt19937; the standard's 10000th-output known answer
    pins it below), mapped by the libstdc++ canonical rule `canonical`.  The
    `mistake` argument plants one plausible restatement error so that the
    tests can show each of them changes the trajectory.
    """

    def __init__(self, seed, lb, ub, fit, radius=2, omega=0.7298, eta1=2.05, eta2=2.05, max_vel=0.5, memory=False,
                 mistake=None, variant=5, topo=2):
        self.words = np.random.RandomState(seed)
        self.lb, self.ub, self.fit = list(lb), list(ub), fit
        self.radius, self.omega, self.eta1, self.eta2, self.max_vel = radius, omega, eta1, eta2, max_vel
        self.memory, self.mistake = memory, mistake
        self.variant, self.topo = variant, topo  # pso_gen.cpp:241-327; 1 gbest, 2 lbest, 3 von Neumann
        self.V = None
        self.cognitive_nonzero = 0  # visits where lbX[p][d] - X[p][d] != 0

    def neighbours(self, pop):
        if self.topo == 3:  # pso_gen.cpp:719-746, offsets (-1,0), (1,0), (0,-1), (0,1)
            rows = int(math.sqrt(pop))
            while pop % rows:
                rows -= 1
            cols = pop // rows
            if self.mistake == "von_transposed":
                rows, cols = cols, rows
            return [[((p // cols + dy) % rows) * cols + (p % cols + dx) % cols
                     for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1))] for p in range(pop)]
        if self.topo == 1:  # pso_gen.cpp:644-662: FIPS sees every particle
            return [list(range(pop)) for _ in range(pop)]
        R = self.radius  # pso_gen.cpp:679-699
        return [[(p - j) % pop for j in range(R, 0, -1)] + [(p + j) % pop for j in range(1, R + 1)]
                for p in range(pop)]

    def draw(self, a, b):
        u0, u1 = (int(w) for w in self.words.randint(0, 2 ** 32, size=2, dtype=np.uint64))
        return canonical(u0, u1) * (b - a) + a

    def evolve(self, X, F, gens):
        pop, nx = len(X), len(X[0])
        lb, ub = self.lb, self.ub
        maxv = [(ub[j] - lb[j]) * self.max_vel for j in range(nx)]
        minv = [-1.0 * m for m in maxv]
        X = [list(map(float, r)) for r in X]
        lbX = [list(r) for r in X]
        lbfit = [float(f) for f in F]
        if self.V is None or len(self.V) != pop or not self.memory or self.mistake == "velocity_reset":
            self.V = [[self.draw(minv[j], maxv[j]) for j in range(nx)] for _ in range(pop)]
        V = self.V
        neigh = self.neighbours(pop)
        if self.topo == 1:
            # pop.best_idx(): the first index of the minimum by operator<
            g = 0
            for i in range(1, pop):
                if lbfit[i] < lbfit[g]:
                    g = i
            gbX, gbfit = list(X[g]), lbfit[g]
        acc = self.eta1 + self.eta2
        w, e1, e2 = self.omega, self.eta1, self.eta2
        for _ in range(gens):
            for p in range(pop):
                if self.topo == 1:
                    best = gbX
                elif self.variant != 6:
                    b = neigh[p][0]
                    for q in neigh[p][1:]:
                        if _le_f(lbfit[q], lbfit[b]) if self.mistake != "first_wins" else lbfit[q] < lbfit[b]:
                            b = q
                    best = X[b] if self.mistake == "best_from_x" else lbX[b]
                mine = X[p] if self.mistake == "cognitive_x" else lbX[p]
                v = self.variant
                if v in (3, 4) and self.mistake != "v4_per_element":  # one draw (4) or two (3) per particle
                    r1 = self.draw(0.0, 1.0)
                    r2 = self.draw(0.0, 1.0) if v == 3 else r1
                for d in range(nx):
                    self.cognitive_nonzero += lbX[p][d] != X[p][d]
                    if v == 6:
                        sf = 0.0
                        for n in neigh[p]:
                            sf += self.draw(0.0, 1.0) * acc * (lbX[n][d] - X[p][d])
                        V[p][d] = w * (V[p][d] + (sf if self.mistake == "fips_no_mean" else sf / float(len(neigh[p]))))
                        continue
                    if v == 4 and self.mistake == "v4_per_element":
                        r1 = r2 = self.draw(0.0, 1.0)
                    if v in (1, 5):
                        r1 = self.draw(0.0, 1.0)
                        r2 = self.draw(0.0, 1.0)
                    elif v == 2:
                        r1 = r2 = self.draw(0.0, 1.0)
                    if v == 5:
                        V[p][d] = w * (V[p][d] + e1 * r1 * (mine[d] - X[p][d]) + e2 * r2 * (best[d] - X[p][d]))
                    else:
                        V[p][d] = w * V[p][d] + e1 * r1 * (mine[d] - X[p][d]) + e2 * r2 * (best[d] - X[p][d])
            for p in range(pop):
                for d in range(nx):
                    if V[p][d] > maxv[d]:
                        V[p][d] = maxv[d]
                    elif V[p][d] < minv[d]:
                        V[p][d] = minv[d]
                    nxv = X[p][d] + V[p][d]
                    if nxv < lb[d]:
                        nxv = lb[d]
                        V[p][d] = 0.0
                    elif nxv > ub[d]:
                        nxv = ub[d]
                        V[p][d] = 0.0
                    X[p][d] = nxv
            fit = [self.fit(r) for r in X]
            for p in range(pop):
                if _le_f(fit[p], lbfit[p]):
                    lbfit[p] = fit[p]
                    lbX[p] = list(X[p])
                    if self.topo == 1 and _le_f(fit[p], gbfit) and self.mistake != "gbest_frozen":
                        gbX, gbfit = list(X[p]), fit[p]
        # pop.set_xf(i, lbX[i], lbfit[i]), pso_gen.cpp:526-528
        return np.array(lbX), np.array(lbfit)


def test_numpy_mt19937_known_answer():
    # the standard's known answer for std::mt19937 (10000th output, default
    # seed 5489): pins the independent word source of _PsoPython
    w = np.random.RandomState(5489).randint(0, 2 ** 32, size=10000, dtype=np.uint64)
    assert int(w[-1]) == 4123659995


def _rastrigin_pop_with_mirrors(pop, dim, seed):
    # rastrigin is even in every coordinate: row 2k+1 = -row 2k has a
    # bit-identical fitness at a different position, so the later-neighbour
    # tie rule (pso_gen.cpp:615-617) decides which position is followed
    X, _ = O.population_init(O.RASTRIGIN, dim, pop, seed)
    X[1::2] = -X[0::2][: pop // 2]
    F = np.array([_py_rastrigin(r) for r in X])
    assert np.all(F[1::2] == F[0::2][: pop // 2])
    return X, F


@pytest.mark.parametrize("kind,dim,pop,gens,calls", [
    (O.ROSENBROCK, 5, 12, 4, 1),
    (O.ROSENBROCK, 3, 9, 3, 2),
    (O.RASTRIGIN, 4, 10, 5, 1),
])
def test_pso_gen_generations_independent(kind, dim, pop, gens, calls):
    """Several generations (and evolve calls) of the oracle's pso_gen against
    the independent Python loop: from generation 2 on lbX != X for every
    particle that did not improve, so the cognitive term eta1 r1 (lbX - X),
    the best neighbour's memory lbX[best], the velocity carried from one
    generation to the next and the `<=` memory rule all enter."""
    fit = _py_rosenbrock if kind == O.ROSENBROCK else _py_rastrigin
    if kind == O.RASTRIGIN:
        X, F = _rastrigin_pop_with_mirrors(pop, dim, 5)
    else:
        X, F = O.population_init(kind, dim, pop, 9)
        F = F[:, 0]
        assert np.array_equal(F, [fit(r) for r in X])
    _, _, lb, ub = O.problem_info(kind, dim)
    seed = 77
    Xo, Fo = O.pso_gen_evolve(kind, dim, X, F, gens, seed, calls=calls, memory=calls > 1)
    py = _PsoPython(seed, lb, ub, fit, memory=calls > 1)
    Xp, Fp = X, F
    for _ in range(calls):
        Xp, Fp = py.evolve(Xp, Fp, gens)
    assert np.array_equal(Xo, Xp)
    assert np.array_equal(Fo, Fp)
    # the memory differs from the position in many visits after generation
    # 1, so the cognitive term is not identically zero in this run
    assert py.cognitive_nonzero > pop * dim


@pytest.mark.parametrize("variant,topo", [(v, t) for v in range(1, 7) for t in (1, 2, 3)])
def test_pso_gen_variants_and_topologies_independent(variant, topo):
    """Every variant (pso_gen.cpp:241-327) with the gbest, lbest and von
    Neumann topologies, over two evolve calls: the oracle's pso_gen against the
    independent Python loop, bit for bit."""
    kind, dim, pop, gens = O.ROSENBROCK, 3, 12, 3
    X, F = O.population_init(kind, dim, pop, 21)
    F = F[:, 0]
    _, _, lb, ub = O.problem_info(kind, dim)
    Xo, Fo = O.pso_gen_evolve(kind, dim, X, F, gens, 55, variant=variant, neighb_type=topo, calls=2)
    py = _PsoPython(55, lb, ub, _py_rosenbrock, variant=variant, topo=topo)
    Xp, Fp = X, F
    for _ in range(2):
        Xp, Fp = py.evolve(Xp, Fp, gens)
    assert np.array_equal(Xo, Xp)
    assert np.array_equal(Fo, Fp)


@pytest.mark.parametrize("mistake,variant,topo", [("v4_per_element", 4, 2), ("von_transposed", 5, 3),
                                                  ("fips_no_mean", 6, 2), ("gbest_frozen", 1, 1)])
def test_pso_gen_variant_loop_detects_mistakes(mistake, variant, topo):
    """Planted errors in the variant and topology rules of the Python loop
    change the result, so the comparison above would catch each of them."""
    kind, dim, pop, gens = O.ROSENBROCK, 3, 12, 3
    X, F = O.population_init(kind, dim, pop, 21)
    F = F[:, 0]
    _, _, lb, ub = O.problem_info(kind, dim)
    Xo, Fo = O.pso_gen_evolve(kind, dim, X, F, gens, 55, variant=variant, neighb_type=topo, calls=2)
    py = _PsoPython(55, lb, ub, _py_rosenbrock, variant=variant, topo=topo, mistake=mistake)
    Xp, Fp = X, F
    for _ in range(2):
        Xp, Fp = py.evolve(Xp, Fp, gens)
    assert not (np.array_equal(Xo, Xp) and np.array_equal(Fo, Fp))


@pytest.mark.parametrize("mistake", ["cognitive_x", "best_from_x", "velocity_reset", "first_wins"])
def test_pso_gen_independent_loop_detects_mistakes(mistake):
    """Each planted restatement error changes the result of the run above, so
    the independent comparison would catch it (the mistakes are planted in the
    Python loop, never in the product)."""
    if mistake == "first_wins":
        kind, dim, pop, gens, calls = O.RASTRIGIN, 4, 10, 5, 1
        X, F = _rastrigin_pop_with_mirrors(pop, dim, 5)
        fit = _py_rastrigin
    else:
        kind, dim, pop, gens, calls = O.ROSENBROCK, 3, 9, 3, 2
        X, F = O.population_init(kind, dim, pop, 9)
        F = F[:, 0]
        fit = _py_rosenbrock
    _, _, lb, ub = O.problem_info(kind, dim)
    Xo, Fo = O.pso_gen_evolve(kind, dim, X, F, gens, 77, calls=calls, memory=calls > 1)
    py = _PsoPython(77, lb, ub, fit, memory=calls > 1, mistake=mistake)
    Xp, Fp = X, F
    for _ in range(calls):
        Xp, Fp = py.evolve(Xp, Fp, gens)
    assert not (np.array_equal(Xo, Xp) and np.array_equal(Fo, Fp))


# ---- the two-objective sphere of the multi-objective ranking workload --------
# f_k(x) = ||x - c_k||^2 has no reference counterpart; pinned by its closed
# forms and by a sequential sum written here.

def test_mo_sphere_closed_forms():
    rng = np.random.default_rng(3)
    n, nobj = 30, 2
    C = rng.uniform(0, 1, (nobj, n))
    for k in range(nobj):
        f = O.fitness_one(O.MO_SPHERE, n, C[k], centers=C, nobj=nobj)
        assert f[k] == 0.0
        other = 1 - k
        assert f[other] == pytest.approx(float(np.sum((C[k] - C[other]) ** 2)), rel=1e-15)
        for i in (0, 7, n - 1):
            e = C[k].copy()
            e[i] += 1.0
            # (c + 1 - c)^2 = 1 exactly when c + 1 rounds back to c + 1 - c = 1
            d = e[i] - C[k][i]
            assert O.fitness_one(O.MO_SPHERE, n, e, centers=C, nobj=nobj)[k] == d * d


def test_mo_sphere_sequential_sum():
    rng = np.random.default_rng(4)
    n, nobj, pop = 30, 3, 64
    C = rng.uniform(0, 1, (nobj, n))
    X = rng.uniform(-1, 2, (pop, n))
    got = O.bfe(O.MO_SPHERE, n, X, centers=C, nobj=nobj)
    want = np.empty((pop, nobj))
    for p in range(pop):
        for k in range(nobj):
            acc = 0.0
            for i in range(n):
                d = float(X[p, i]) - float(C[k, i])
                acc += d * d
            want[p, k] = acc
    assert np.array_equal(got, want)


# ---- crowding distance and multi-objective sorting (multi_objective.cpp:280-465)

def _f(v):
    return float("inf") if v == "inf" else float(v)


def _mat(points):
    return np.array(points, np.float64).reshape(len(points), -1 if points else 0)


@pytest.mark.parametrize("stable", [False, True])
def test_crowding_distance_kats(kats, stable):
    for k in kats["crowding_distance"]:
        r = O.crowding_distance(_mat(k["points"]), stable=stable)
        assert list(r) == [_f(v) for v in k["result"]], k["src"]
    for k in kats["crowding_distance_throws"]:
        with pytest.raises(O.OracleError, match=k["match"]):
            O.crowding_distance(_mat(k["points"]), stable=stable)


@pytest.mark.parametrize("stable", [False, True])
def test_sort_population_mo_kats(kats, stable):
    for k in kats["sort_population_mo"]:
        assert list(O.sort_population_mo(_mat(k["points"]), stable=stable)) == k["result"], k["src"]


@pytest.mark.parametrize("stable", [False, True])
def test_select_best_N_mo_kats(kats, stable):
    for k in kats["select_best_N_mo"]:
        assert list(O.select_best_N_mo(_mat(k["points"]), k["N"], stable=stable)) == k["result"], k["src"]
    for k in kats["select_best_N_mo_permutation"]:
        pts = _mat(k["points"])
        full = list(O.sort_population_mo(pts, stable=stable))
        for N in range(1, len(pts) + 3):
            sel = list(O.select_best_N_mo(pts, N, stable=stable))
            assert sorted(sel) == sorted(full[: len(sel)]), (k["src"], N)


def _crowding_brute(pts):
    """Crowding distance from its definition, with the neighbours in each
    objective found by value (distinct values only)."""
    n, m = pts.shape
    out = np.zeros(n)
    for i in range(m):
        order = np.argsort(pts[:, i], kind="stable")
        df = pts[order[-1], i] - pts[order[0], i]
        for j in range(n):
            pos = int(np.where(order == j)[0][0])
            if pos in (0, n - 1):
                out[j] = np.inf
            elif np.isfinite(out[j]):
                out[j] += (pts[order[pos + 1], i] - pts[order[pos - 1], i]) / df
    return out


@pytest.mark.parametrize("n,m", [(2, 2), (5, 3), (40, 2), (120, 4)])
def test_crowding_distance_brute_force(n, m):
    # distinct values: the sort order is unique and the definition applies
    pts = np.random.default_rng(n * m).permutation(n * m * 3)[: n * m].reshape(n, m) / 7.0
    np.testing.assert_array_equal(O.crowding_distance(pts), _crowding_brute(pts))
    np.testing.assert_array_equal(O.crowding_distance(pts, stable=True), _crowding_brute(pts))


def test_sort_population_mo_properties():
    # ordered by (rank asc, crowding desc); the stable variant breaks the
    # remaining ties by index; both are permutations of the population
    rng = np.random.default_rng(4)
    pts = rng.integers(0, 8, size=(300, 2)).astype(np.float64)
    r = O.fnds(pts)
    rank = r["rank"].astype(np.int64)
    crowd = np.zeros(len(pts))
    fo, fi = r["front_offsets"], r["front_items"]
    for f in range(len(fo) - 1):
        front = fi[fo[f]:fo[f + 1]].astype(np.int64)
        if len(front) > 1:
            crowd[front] = O.crowding_distance(pts[front], stable=True)
    for stable in (False, True):
        order = O.sort_population_mo(pts, stable=stable).astype(np.int64)
        assert sorted(order) == list(range(len(pts)))
        key = [(rank[i], -crowd[i]) for i in order]
        if stable:
            assert key == sorted(key)
            key_i = [(rank[i], -crowd[i], i) for i in order]
            assert key_i == sorted(key_i)


# ---- the memetic local step (no reference counterpart) ----------------------

def test_local_descent_hand_traces():
    # rosenbrock at the origin: g = (-2, 0), m = 2.  alpha 0.5: y = (0.5, 0),
    # f = 100 * 0.25^2 + 0.25 = 6.5 > 1, rejected; alpha 0.25: y = (0.25, 0),
    # f = 100 * 0.0625^2 + 0.75^2 = 0.953125 < 1, accepted.  All exact.
    x, f = O.local_descent(O.ROSENBROCK, 2, np.zeros((1, 2)), 2, 0.5)
    assert x.tolist() == [[0.25, 0.0]] and f.tolist() == [0.953125]
    # one more round from (0.25, 0) with alpha 0.5: g = (4.75, -12.5), m = 12.5,
    # y = (0.06, 0.5), f(y) = 25.52... > 0.953125: rejected, x stays
    x3, f3 = O.local_descent(O.ROSENBROCK, 2, np.zeros((1, 2)), 3, 0.5)
    s = 0.5 / 12.5
    y = np.array([0.25 - s * 4.75, 0.0 - s * -12.5])
    fy = 100 * (y[0] ** 2 - y[1]) ** 2 + (y[0] - 1) ** 2
    assert fy > 0.953125 and x3.tolist() == [[0.25, 0.0]] and f3.tolist() == [0.953125]
    # a bound clamps the trial: ub_0 = 0.125 -> (0.125, 0), f = 0.7900390625
    x, f = O.local_descent(O.ROSENBROCK, 2, np.zeros((1, 2)), 1, 0.5, [-5.0, -5.0], [0.125, 10.0])
    assert x.tolist() == [[0.125, 0.0]] and f.tolist() == [0.7900390625]
    # zero rounds: the fitness only; the minimum does not move (g = 0)
    x, f = O.local_descent(O.ROSENBROCK, 4, np.ones((1, 4)), 10, 0.5)
    assert x.tolist() == [[1.0] * 4] and f.tolist() == [0.0]


def test_local_descent_rastrigin_reaches_the_local_minimum():
    # from 0.1 the steps of a one-dimensional rastrigin run into its minimum 0
    x, f = O.local_descent(O.RASTRIGIN, 1, np.array([[0.1]]), 40, 0.05)
    assert abs(x[0, 0]) < 1e-6 and f[0] < 1e-10


@pytest.mark.parametrize("kind,dim", [(O.ROSENBROCK, 7), (O.RASTRIGIN, 9), (O.LENNARD_JONES, 6)])
def test_local_descent_invariants(kind, dim):
    nx, _, lb, ub = O.problem_info(kind, dim)
    lo, hi = (lb, ub) if kind != O.LENNARD_JONES else (np.full(nx, 0.3), np.full(nx, 2.5))
    x0 = O.batch_uniform(lo, hi, 40, 2)
    f0 = O.bfe(kind, dim, x0)[:, 0]
    x, f = O.local_descent(kind, dim, x0, 15, 0.05, lb, ub)
    assert np.all(f <= f0) and np.mean(f < f0) > 0.5
    assert np.array_equal(f, O.bfe(kind, dim, x)[:, 0])
    assert np.all((x >= lb) & (x <= ub))
    assert np.array_equal(np.any(x != x0, axis=1), f < f0)



# ---- ZDT and DTLZ (tests/zdt.cpp, tests/dtlz.cpp) ---------------------------

def kat_x(spec):
    if "fill" in spec:
        n, v = spec["fill"]
        return np.full(n, v)
    n, v, at, w = spec["fill2"]
    x = np.full(n, v)
    x[at:] = w
    return x


def test_zdt_kats(kats):
    tol = kats["zdt_fitness"]["rel_tol"]
    for c in kats["zdt_fitness"]["cases"]:
        f = O.mo_fitness(O.ZDT, c["prob_id"], c["param"], kat_x(c["x"]))[0]
        for got, want in zip(f, c["f"]):
            assert abs(got - want) <= tol * abs(want), c


def test_dtlz_kats(kats):
    tol = kats["dtlz_fitness"]["rel_tol"]
    for c in kats["dtlz_fitness"]["cases"]:
        f = O.mo_fitness(O.DTLZ, c["prob_id"], c["dim"], np.array(c["x"]), c["fdim"], c["alpha"])[0]
        for got, want in zip(f, c["f"]):
            assert abs(got - want) <= tol * abs(want), c


def test_zdt_dtlz_shapes_and_bounds():
    # zdt.cpp:108-132 and tests/zdt.cpp:62; dtlz.cpp:116-119 and tests/dtlz.cpp:184-190
    assert O.mo_info(O.ZDT, 5, 11)[0] == (11 - 1) * 5 + 30
    nx, nf, lb, ub = O.mo_info(O.ZDT, 4, 10)
    assert (nx, nf) == (10, 2) and lb[0] == 0.0 and ub[0] == 1.0 and np.all(lb[1:] == -5.0) and np.all(ub[1:] == 5.0)
    for pid in range(1, 8):
        nx, nf, lb, ub = O.mo_info(O.DTLZ, pid, 4, 3)
        assert (nx, nf) == (4, 3) and np.all(lb == 0.0) and np.all(ub == 1.0)


def test_dtlz_closed_forms():
    # the Pareto-optimal sets: x_M = 0.5 (DTLZ1-5) gives g = 0, where DTLZ1's
    # objectives sum to 0.5 and DTLZ2's squares sum to 1 (unit sphere)
    rng = np.random.default_rng(4)
    for _ in range(20):
        x = np.concatenate([rng.uniform(0, 1, 2), np.full(5, 0.5)])
        f1 = O.mo_fitness(O.DTLZ, 1, 7, x, 3)[0]
        assert abs(f1.sum() - 0.5) < 1e-15
        f2 = O.mo_fitness(O.DTLZ, 2, 7, x, 3)[0]
        assert abs(np.sum(f2 * f2) - 1.0) < 1e-15
