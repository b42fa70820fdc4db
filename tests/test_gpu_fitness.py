"""Parity of the HIP batch fitness / gradient kernels with the oracle.

Bars: rosenbrock and the multi-objective sphere are bit-exact (only + - *
in the reference's order); the transcendental problems are within 1e-12
relative; Lennard-Jones within 1e-12 of the sum of the absolute pair terms
(its energy cancels); gradients as stated per test.
"""
import math

import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def gpu_bfe(kind, dim, dvs, centers=None, nobj=1):
    d = dev(dvs)
    nf = nobj if kind == P.MO_SPHERE else 1
    out = torch.empty((dvs.shape[0], nf), dtype=torch.float64, device=DEV)
    c = None if centers is None else dev(centers)
    P.bfe_eval(kind, dim, d, out, centers=c, nobj=nobj)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def inputs(kind, dim, n, seed, lo=None, hi=None):
    nx, _, lb, ub = O.problem_info(kind, dim)
    if lo is not None:
        lb, ub = np.full(nx, lo), np.full(nx, hi)
    return O.batch_uniform(lb, ub, n, seed)


def lj_abs_scale(x):
    atoms = (x.size + 6) // 3
    p = np.zeros((atoms, 3))
    p[1, 2] = x[0]
    p[2, 1:] = x[1:3]
    p[3:] = x[3:].reshape(-1, 3)
    d = p[:, None, :] - p[None, :, :]
    r2 = np.sum(d * d, axis=-1)[np.triu_indices(atoms, 1)]
    s = r2 ** -3
    return 4 * np.sum(np.abs(s * s) + np.abs(s))


@pytest.mark.parametrize("n", [1, 3, 64, 257])
@pytest.mark.parametrize("dim", [2, 3, 10, 37, 63, 64, 65, 500, 1000])
def test_rosenbrock_bit_exact(n, dim):
    x = inputs(P.ROSENBROCK, dim, n, 100 + dim)
    assert np.array_equal(gpu_bfe(P.ROSENBROCK, dim, x), O.bfe(P.ROSENBROCK, dim, x, threads=8))


@pytest.mark.parametrize("kind", [P.RASTRIGIN, P.ACKLEY, P.SCHWEFEL, P.GRIEWANK])
@pytest.mark.parametrize("dim", [1, 2, 7, 31, 32, 33, 255, 256, 1000, 1001])
@pytest.mark.parametrize("n", [1, 5, 130])
def test_separable_rel_1e12(kind, dim, n):
    x = inputs(kind, dim, n, 7 * dim + n)
    g = gpu_bfe(kind, dim, x)
    r = O.bfe(kind, dim, x, threads=8)
    np.testing.assert_allclose(g, r, rtol=1e-12, atol=0)


def test_every_dimension_to_300():
    """Every dimension from 1 (2 for rosenbrock) to 300 and a few beyond, for
    the five problems of the row kernels (lanes per individual, staged and
    streaming forms, 16-byte pairs for even rosenbrock rows), batches of 3 and
    67 rows: rosenbrock bit for bit, the others within 1e-12 relative."""
    dims = list(range(1, 301)) + [511, 512, 513, 1023, 1024, 1025]
    for kind in (P.ROSENBROCK, P.RASTRIGIN, P.ACKLEY, P.SCHWEFEL, P.GRIEWANK):
        for dim in dims:
            if kind == P.ROSENBROCK and dim < 2:
                continue
            for n in (3, 67):
                x = inputs(kind, dim, n, 31 * dim + n + kind)
                g = gpu_bfe(kind, dim, x)
                r = O.bfe(kind, dim, x, threads=8)
                if kind == P.ROSENBROCK:
                    assert np.array_equal(g, r), (dim, n)
                else:
                    np.testing.assert_allclose(g, r, rtol=1e-12, atol=0, err_msg=f"{kind} {dim} {n}")


@pytest.mark.parametrize("kind", [P.RASTRIGIN, P.ACKLEY])
def test_separable_out_of_range_arguments(kind):
    """Arguments beyond the streaming kernel's cosine range (|x| >= 2^10), inf
    and NaN are legal inputs (bounds are never checked, problem.cpp:948-949);
    those rows are recomputed with the reference's expression and an accurate
    cosine."""
    x = inputs(kind, 300, 14, 9)
    x[1, 17] = 3.0e5
    x[2, 0] = -7.5e9
    x[3, 299] = 1.0e300
    x[4, 5] = np.inf
    x[5, 6] = np.nan
    x[6, 7] = 166886.0  # just below 2^20 / (2 pi): the accurate cosine's own fast reduction
    x[7, 8] = 1024.5    # just past the streaming range
    x[8, :] = 1023.75   # inside it, but the lanes' sums of squares exceed 2^20
    x[9, 9] = -np.inf
    x[10, :] = 400.3    # streaming range, identical elements: their errors add up
    x[11, :] = -400.3
    g = gpu_bfe(kind, 300, x)
    r = O.bfe(kind, 300, x, threads=4)
    np.testing.assert_allclose(g, r, rtol=1e-12, atol=0, equal_nan=True)


def test_separable_unaligned_rows_same_order():
    # an odd row offset selects the unaligned load path; the reduction order
    # (and so every bit of the result) must not change
    x = inputs(P.RASTRIGIN, 1000, 33, 5)
    d = dev(np.concatenate([np.zeros(1), x.reshape(-1)]))[1:].view(33, 1000)
    out = torch.empty((33, 1), dtype=torch.float64, device=DEV)
    P.bfe_eval(P.RASTRIGIN, 1000, d, out)
    assert np.array_equal(out.cpu().numpy(), gpu_bfe(P.RASTRIGIN, 1000, x))


# 30-32, 45-48, 60-64, 90-96 and 120-128 atoms: the energy on pair-once
# blocks; the other sizes on the cyclic schedule
@pytest.mark.parametrize("atoms,n", [(3, 5), (4, 33), (5, 17), (9, 100), (13, 7), (30, 40), (32, 33), (40, 20),
                                     (45, 30), (48, 300), (60, 9), (64, 20), (90, 7), (96, 20), (100, 20),
                                     (120, 5), (128, 9), (169, 9)])
def test_lennard_jones(atoms, n):
    x = inputs(P.LENNARD_JONES, atoms, n, atoms, lo=-3.0, hi=3.0)
    g = gpu_bfe(P.LENNARD_JONES, atoms, x)[:, 0]
    r = O.bfe(P.LENNARD_JONES, atoms, x, threads=8)[:, 0]
    for i in range(n):
        assert abs(g[i] - r[i]) <= 1e-12 * lj_abs_scale(x[i]), (i, g[i], r[i])


def test_lennard_jones_reference_kats(kats):
    for k in kats["fitness"]:
        if k["kind"] == "lennard_jones":
            g = gpu_bfe(P.LENNARD_JONES, 3, np.array([k["x"]]))[0, 0]
            assert abs(g - k["f"][0]) <= 1e-13 * abs(k["f"][0])


def test_lennard_jones_coincident_atoms():
    # lennard_jones.cpp:81-83,89: dist == 0 overwrites with DBL_MAX -> +inf
    x = inputs(P.LENNARD_JONES, 6, 4, 1, lo=-3.0, hi=3.0)
    x[0, 3:6] = 0.0               # atom 3 on atom 0
    x[1, 0] = 0.0                 # atom 1 on atom 0
    x[2, 3:6] = x[2, 6:9]         # atoms 3 and 4 coincide
    g = gpu_bfe(P.LENNARD_JONES, 6, x)[:, 0]
    r = O.bfe(P.LENNARD_JONES, 6, x)[:, 0]
    assert g[0] == math.inf and r[0] == math.inf
    assert g[1] == math.inf and r[1] == math.inf
    assert g[2] == math.inf and r[2] == math.inf
    assert np.isfinite(g[3])


@pytest.mark.parametrize("kind", [P.RASTRIGIN, P.ACKLEY, P.ROSENBROCK])
def test_reference_kats_on_device(kats, kind):
    name = {P.RASTRIGIN: "rastrigin", P.ACKLEY: "ackley", P.ROSENBROCK: "rosenbrock"}[kind]
    for k in kats["fitness"]:
        if k["kind"] != name:
            continue
        x = np.array([k["x"]])
        dim = x.shape[1]
        g = gpu_bfe(kind, dim, x)[0, 0]
        if k.get("exact"):
            assert g == k["f"][0], k["src"]
        else:
            assert abs(g - k["f"][0]) <= 1e-12 * abs(k["f"][0]), k["src"]


def test_mo_sphere_bit_exact():
    rng = np.random.default_rng(0)
    c = rng.uniform(0, 1, (2, 30))
    x = rng.uniform(0, 1, (1001, 30))
    g = gpu_bfe(P.MO_SPHERE, 30, x, centers=c, nobj=2)
    r = O.bfe(P.MO_SPHERE, 30, x, centers=c, nobj=2)
    assert np.array_equal(g, r)


def test_host_entry_point_matches_device():
    x = inputs(P.ACKLEY, 100, 50, 3)
    h = P.bfe_eval_host(P.ACKLEY, 100, x)
    assert np.array_equal(h, gpu_bfe(P.ACKLEY, 100, x))


@pytest.mark.parametrize("rows", [2048, 5243, 9000])
def test_large_host_batch_goes_through_pinned_chunks(rows):
    """Batches of 16 MiB and more are uploaded through two pinned 16 MiB
    chunks filled by host threads (one chunk exactly, a ragged last chunk,
    and more than two chunks so that both halves are reused); the result is
    the device-resident evaluation's, and a second call reuses the buffers."""
    x = np.random.default_rng(rows).uniform(-5.12, 5.12, (rows, 1024))
    want = gpu_bfe(P.RASTRIGIN, 1024, x)
    assert np.array_equal(P.bfe_eval_host(P.RASTRIGIN, 1024, x), want)
    y = x[::-1].copy()
    assert np.array_equal(P.bfe_eval_host(P.RASTRIGIN, 1024, y), want[::-1])


def test_large_host_gradient_batch_goes_through_pinned_chunks():
    x = np.random.default_rng(4).uniform(-2.0, 2.0, (5000, 1000))
    d = dev(x)
    g = torch.empty_like(d)
    P.batch_gradient(P.ROSENBROCK, 1000, d, g)
    torch.cuda.synchronize()
    assert np.array_equal(P.batch_gradient_host(P.ROSENBROCK, 1000, x), g.cpu().numpy())


# ---- gradients ---------------------------------------------------------------

@pytest.mark.parametrize("dim,n", [(2, 3), (10, 64), (501, 17), (64, 257), (1000, 300), (1001, 513)])
def test_rosenbrock_gradient_bit_exact(dim, n):
    x = inputs(P.ROSENBROCK, dim, n, dim)
    g = torch.empty((n, dim), dtype=torch.float64, device=DEV)
    f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    P.batch_gradient(P.ROSENBROCK, dim, dev(x), g, f)
    torch.cuda.synchronize()
    assert np.array_equal(g.cpu().numpy(), O.gradient(P.ROSENBROCK, dim, x))
    assert np.array_equal(f.cpu().numpy(), O.bfe(P.ROSENBROCK, dim, x))


def test_gradients_every_dimension_to_300():
    """Fitness + gradient of rosenbrock (bit for bit) and rastrigin (within
    1e-13 (|2x| + 20 pi) per component, the fitness equal to a bfe call's) at
    every dimension from 2 to 300 and around 512 and 1024, batches of 3 and
    67 rows: the fused, paired, staged and element-wise kernels' thresholds."""
    for dim in list(range(2, 301)) + [511, 512, 513, 1023, 1024, 1025]:
        for n in (3, 67):
            for kind in (P.ROSENBROCK, P.RASTRIGIN):
                x = inputs(kind, dim, n, 17 * dim + n + kind)
                g = torch.empty((n, dim), dtype=torch.float64, device=DEV)
                f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
                P.batch_gradient(kind, dim, dev(x), g, f)
                torch.cuda.synchronize()
                got, fit, ref = g.cpu().numpy(), f.cpu().numpy(), O.gradient(kind, dim, x)
                if kind == P.ROSENBROCK:
                    assert np.array_equal(got, ref), (dim, n)
                    assert np.array_equal(fit, O.bfe(kind, dim, x)), (dim, n)
                else:
                    assert np.all(np.abs(got - ref) <= 1e-13 * (np.abs(2 * x) + 20 * np.pi)), (dim, n)
                    assert np.array_equal(fit, gpu_bfe(kind, dim, x)), (dim, n)


@pytest.mark.parametrize("dim,n", [(300, 40), (1000, 257), (301, 33), (12, 1000), (2, 5)])
def test_rastrigin_gradient(dim, n):
    # even dimensions take the fused fitness + gradient pass, odd ones the
    # element-wise gradient kernel
    x = inputs(P.RASTRIGIN, dim, n, 2 + dim)
    if n > 200:
        # rows that leave the streaming routines' range (|x| >= 2^10, inf,
        # NaN) are recomputed with the general routines
        x[3, dim // 2] = 3000.5
        x[7, 0] = -1e9
        x[11, dim - 1] = np.inf
        x[13, 1] = np.nan
    g = torch.empty((n, dim), dtype=torch.float64, device=DEV)
    f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    P.batch_gradient(P.RASTRIGIN, dim, dev(x), g, f)
    torch.cuda.synchronize()
    ref = O.gradient(P.RASTRIGIN, dim, x)
    got = g.cpu().numpy()
    fin = np.isfinite(x)
    scale = np.abs(2 * x) + 10 * 2 * np.pi
    assert np.all(np.abs(got[fin] - ref[fin]) <= 1e-13 * scale[fin])
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    # the fitness equals a bfe call's bit for bit
    assert np.array_equal(f.cpu().numpy(), gpu_bfe(P.RASTRIGIN, dim, x), equal_nan=True)


# 22 to 128 atoms take the pair-once block kernel (16-lane rows up to 64
# atoms, 32-lane rows beyond; multiples of the row without padded slots, the
# others with); the rest visit each pair from both atoms
@pytest.mark.parametrize("atoms,n", [(3, 4), (7, 20), (16, 9), (17, 40), (21, 9), (22, 30), (24, 30), (31, 20),
                                     (32, 64), (33, 50), (47, 20), (48, 64), (49, 9), (63, 10), (64, 30), (65, 6),
                                     (80, 9), (96, 7), (97, 5), (120, 5), (128, 6), (129, 3), (150, 2)])
def test_lennard_jones_gradient(atoms, n):
    x = inputs(P.LENNARD_JONES, atoms, n, 11 + atoms, lo=-3.0, hi=3.0)
    nx = x.shape[1]
    g = torch.empty((n, nx), dtype=torch.float64, device=DEV)
    f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    P.batch_gradient(P.LENNARD_JONES, atoms, dev(x), g, f)
    torch.cuda.synchronize()
    g, f = g.cpu().numpy(), f.cpu().numpy()[:, 0]
    ref = O.gradient(P.LENNARD_JONES, atoms, x)
    rf = O.bfe(P.LENNARD_JONES, atoms, x)[:, 0]
    for i in range(n):
        assert np.linalg.norm(g[i] - ref[i]) <= 1e-12 * np.linalg.norm(ref[i])
        assert abs(f[i] - rf[i]) <= 1e-12 * lj_abs_scale(x[i])


def test_lennard_jones_gradient_special_distances():
    """Rows that leave the streaming path: coincident atoms (energy +inf, the
    forces between them NaN, like the oracle's 1/r^2), and an atom so far away
    that r^6 overflows (its terms and forces vanish)."""
    atoms, n = 6, 4
    x = inputs(P.LENNARD_JONES, atoms, n, 5, lo=-3.0, hi=3.0)
    x[0, 3:6] = 0.0           # atom 3 on atom 0
    x[1, 6:9] = x[1, 9:12]    # atoms 4 and 5 coincide
    x[2, 9:12] = 1.0e120      # atom 5 far away: r^2 ~ 1e240, r^6 overflows
    g = torch.empty((n, x.shape[1]), dtype=torch.float64, device=DEV)
    f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    P.batch_gradient(P.LENNARD_JONES, atoms, dev(x), g, f)
    torch.cuda.synchronize()
    g, f = g.cpu().numpy(), f.cpu().numpy()[:, 0]
    ref = O.gradient(P.LENNARD_JONES, atoms, x)
    rf = O.bfe(P.LENNARD_JONES, atoms, x)[:, 0]
    assert f[0] == math.inf and f[1] == math.inf
    for i in range(n):
        assert np.array_equal(np.isnan(g[i]), np.isnan(ref[i])), i
        ok = ~np.isnan(ref[i])
        assert np.linalg.norm(g[i][ok] - ref[i][ok]) <= 1e-12 * max(np.linalg.norm(ref[i][ok]), 1e-300), i
    for i in (2, 3):
        assert abs(f[i] - rf[i]) <= 1e-12 * lj_abs_scale(x[i]), (i, f[i], rf[i])


@pytest.mark.parametrize("atoms", [40, 48, 80])
def test_lennard_jones_gradient_special_distances_blocks(atoms):
    """The same special rows at sizes of the pair-once kernel, plus an atom at
    1e25 (inside the range of its batched reciprocals) and one at 1e40 (beyond
    2^100: the careful path)."""
    n = 6
    x = inputs(P.LENNARD_JONES, atoms, n, 7 + atoms, lo=-3.0, hi=3.0)

    def at(a):  # flat elements of atom a >= 3
        return slice(3 * a - 6, 3 * a - 3)

    x[0, at(3)] = 0.0                 # atom 3 on atom 0
    x[1, at(20)] = x[1, at(21)]       # atoms 20 and 21 coincide
    x[2, at(atoms - 1)] = 1.0e120     # the last atom far away
    x[3, at(10)] = 1.0e25
    x[4, at(17)] = 1.0e40
    g = torch.empty((n, x.shape[1]), dtype=torch.float64, device=DEV)
    f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    P.batch_gradient(P.LENNARD_JONES, atoms, dev(x), g, f)
    torch.cuda.synchronize()
    g, f = g.cpu().numpy(), f.cpu().numpy()[:, 0]
    ref = O.gradient(P.LENNARD_JONES, atoms, x)
    rf = O.bfe(P.LENNARD_JONES, atoms, x)[:, 0]
    assert f[0] == math.inf and f[1] == math.inf
    for i in range(n):
        assert np.array_equal(np.isnan(g[i]), np.isnan(ref[i])), i
        ok = ~np.isnan(ref[i])
        assert np.linalg.norm(g[i][ok] - ref[i][ok]) <= 1e-12 * max(np.linalg.norm(ref[i][ok]), 1e-300), i
    for i in range(2, n):
        assert abs(f[i] - rf[i]) <= 1e-12 * lj_abs_scale(x[i]), (i, f[i], rf[i])
    # the energy kernel on the same rows (its reciprocals are batched too)
    fb = gpu_bfe(P.LENNARD_JONES, atoms, x)[:, 0]
    assert fb[0] == math.inf and fb[1] == math.inf
    for i in range(2, n):
        assert abs(fb[i] - rf[i]) <= 1e-12 * lj_abs_scale(x[i]), (i, fb[i], rf[i])


def test_lennard_jones_every_size_to_140():
    """Every cluster size from 3 to 140 atoms goes through the lane and block
    rules that pick its kernels (cyclic schedule, pair-once blocks of 16 or 32
    lanes, with and without padded slots): energy and gradient of two random
    rows against the oracle, and a third row with two coincident atoms that
    takes the careful paths."""
    for atoms in range(3, 141):
        x = inputs(P.LENNARD_JONES, atoms, 3, 1000 + atoms, lo=-3.0, hi=3.0)
        if atoms >= 5:
            x[2, 3:6] = x[2, 6:9]  # atoms 3 and 4 coincide
        fb = gpu_bfe(P.LENNARD_JONES, atoms, x)[:, 0]
        g = torch.empty((3, x.shape[1]), dtype=torch.float64, device=DEV)
        f = torch.empty((3, 1), dtype=torch.float64, device=DEV)
        P.batch_gradient(P.LENNARD_JONES, atoms, dev(x), g, f)
        torch.cuda.synchronize()
        g, f = g.cpu().numpy(), f.cpu().numpy()[:, 0]
        rf = O.bfe(P.LENNARD_JONES, atoms, x)[:, 0]
        rg = O.gradient(P.LENNARD_JONES, atoms, x)
        for i in range(3):
            if math.isfinite(rf[i]):
                sc = lj_abs_scale(x[i])
                assert abs(fb[i] - rf[i]) <= 1e-12 * sc, (atoms, i)
                assert abs(f[i] - rf[i]) <= 1e-12 * sc, (atoms, i)
                assert np.linalg.norm(g[i] - rg[i]) <= 1e-12 * np.linalg.norm(rg[i]), (atoms, i)
            else:
                assert fb[i] == rf[i] and f[i] == rf[i], (atoms, i, fb[i], f[i], rf[i])
                assert np.array_equal(np.isnan(g[i]), np.isnan(rg[i])), (atoms, i)


def test_lennard_jones_gradient_finite_differences():
    # central differences, h = 1e-6, and estimate_gradient_h
    # (gradients_and_hessians.hpp:199-244) at dx = 1e-4; <= 1e-7 relative
    atoms, n = 48, 4
    x = inputs(P.LENNARD_JONES, atoms, n, 99, lo=-3.0, hi=3.0)
    g = P.batch_gradient_host(P.LENNARD_JONES, atoms, x)
    for i in range(n):
        fd = np.empty(x.shape[1])
        for k in range(x.shape[1]):
            xp, xm = x[i].copy(), x[i].copy()
            xp[k] += 1e-6
            xm[k] -= 1e-6
            fd[k] = (O.fitness_one(P.LENNARD_JONES, atoms, xp)[0] - O.fitness_one(P.LENNARD_JONES, atoms, xm)[0]) / 2e-6
        assert np.linalg.norm(fd - g[i]) <= 1e-7 * np.linalg.norm(g[i])
        gh = O.estimate_gradient(P.LENNARD_JONES, atoms, x[i], 1e-4, True)
        assert np.linalg.norm(gh - g[i]) <= 1e-7 * np.linalg.norm(g[i])


# ---- RNG ----------------------------------------------------------------------

# beyond 2^21 words the stream is split across workgroups by jump-ahead
@pytest.mark.parametrize("seed,skip,n", [(5489, 0, 10000), (42, 0, 1), (42, 1000, 5000), (7, 623, 700),
                                         (5489, 0, (1 << 21) + 1), (42, 0, 5_000_000), (9, 0, (1 << 24) + 12345)])
def test_mt19937_bit_exact(seed, skip, n):
    out = torch.empty(n, dtype=torch.int32, device=DEV)
    P.mt19937_generate(seed, n, out, skip=skip)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), O.mt19937_u32(seed, n, skip=skip))


@pytest.mark.parametrize("kind,dim,n", [(P.ROSENBROCK, 10, 64), (P.RASTRIGIN, 1000, 33), (P.LENNARD_JONES, 48, 20),
                                        (P.RASTRIGIN, 1000, 2000), (P.ROSENBROCK, 10, 150_001)])
def test_batch_random_dv_bit_exact(kind, dim, n):
    nx, _ = P.dims(kind, dim)
    out = torch.empty((n, nx), dtype=torch.float64, device=DEV)
    P.batch_random_dv(kind, dim, n, 42, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), O.batch_random_dv(kind, dim, n, 42))


@pytest.mark.parametrize("kind,dim", [(P.ROSENBROCK, 10), (P.RASTRIGIN, 1000), (P.LENNARD_JONES, 13)])
def test_translated_problem_host_path(kind, dim):
    # translate (translate.cpp:101-105, 245-249): fitness and gradient of the
    # inner problem at x - t, with t subtracted on the device
    x = inputs(kind, dim, 300, 5, lo=-3.0 if kind == P.LENNARD_JONES else None,
               hi=3.0 if kind == P.LENNARD_JONES else None)
    t = np.random.default_rng(dim).uniform(-0.5, 0.5, x.shape[1])
    f = P.bfe_eval_host(kind, dim, x, shift=t)[:, 0]
    ref = O.bfe(kind, dim, x - t, threads=8)[:, 0]
    g = P.batch_gradient_host(kind, dim, x, shift=t)
    gref = O.gradient(kind, dim, x - t)
    if kind == P.ROSENBROCK:
        assert np.array_equal(f, ref) and np.array_equal(g, gref)
    elif kind == P.RASTRIGIN:
        assert np.all(np.abs(f - ref) <= 1e-12 * np.abs(ref))
        assert np.allclose(g, gref, rtol=1e-12, atol=1e-12)
    else:
        for i in range(x.shape[0]):
            assert abs(f[i] - ref[i]) <= 1e-12 * lj_abs_scale(x[i] - t)
            assert np.linalg.norm(g[i] - gref[i]) <= 1e-12 * np.linalg.norm(gref[i])
