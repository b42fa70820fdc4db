"""crowding_distance, sort_population_mo and select_best_N_mo on the device
(multi_objective.cpp:280-465) against the oracle.

The reference sorts with std::sort, which leaves the order of equal keys open;
the device keeps equal keys in input order.  So: against the reference's known
answers and the oracle's std::sort form whenever no order-sensitive ties
occur, and bit for bit against the oracle's stable form (std::stable_sort,
the device's rule) always."""
import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"
INF = float("inf")


def _f(v):
    return INF if v == "inf" else float(v)


def _mat(points):
    return np.array(points, np.float64).reshape(len(points), -1 if points else 0)


def dev_crowding(pts):
    ws = P.Fnds()
    try:
        d = torch.from_numpy(np.ascontiguousarray(pts)).to(DEV)
        out = torch.empty(len(pts), dtype=torch.float64, device=DEV)
        ws.crowding_distance(d, out)
        return out.cpu().numpy()
    finally:
        ws.close()


def dev_sort(pts):
    ws = P.Fnds()
    try:
        d = torch.from_numpy(np.ascontiguousarray(pts)).to(DEV)
        out = torch.empty(max(len(pts), 1), dtype=torch.int64, device=DEV)
        ws.sort_population_mo(d, out)
        return out.cpu().numpy()[: len(pts)].astype(np.uint64)
    finally:
        ws.close()


def dev_select(pts, N):
    ws = P.Fnds()
    try:
        d = torch.from_numpy(np.ascontiguousarray(pts)).to(DEV)
        out = torch.empty(max(len(pts), 1), dtype=torch.int64, device=DEV)
        k = ws.select_best_N_mo(d, N, out)
        return out.cpu().numpy()[:k].astype(np.uint64)
    finally:
        ws.close()


def test_crowding_distance_kats(kats):
    for k in kats["crowding_distance"]:
        pts = _mat(k["points"])
        want = [_f(v) for v in k["result"]]
        assert list(dev_crowding(pts)) == want, k["src"]
        assert list(P.crowding_distance_host(pts)) == want, k["src"]


def test_sort_population_mo_kats(kats):
    for k in kats["sort_population_mo"]:
        pts = _mat(k["points"])
        if len(pts) >= 2:
            assert list(map(int, dev_sort(pts))) == k["result"], k["src"]
        assert list(map(int, P.sort_population_mo_host(pts))) == k["result"], k["src"]


def test_select_best_N_mo_kats(kats):
    for k in kats["select_best_N_mo"]:
        assert list(map(int, P.select_best_N_mo_host(_mat(k["points"]), k["N"]))) == k["result"], k["src"]
    for k in kats["select_best_N_mo_permutation"]:
        pts = _mat(k["points"])
        full = list(map(int, P.sort_population_mo_host(pts)))
        for N in range(1, len(pts) + 3):
            sel = list(map(int, P.select_best_N_mo_host(pts, N)))
            assert sorted(sel) == sorted(full[: len(sel)]), (k["src"], N)


@pytest.mark.parametrize("n,m", [(2, 2), (3, 3), (17, 2), (300, 2), (1000, 3), (4097, 5)])
def test_crowding_distance_distinct(n, m):
    pts = np.random.default_rng(n + m).random((n, m))
    got = dev_crowding(pts)
    assert np.array_equal(got, O.crowding_distance(pts))          # std::sort form: no ties
    assert np.array_equal(got, O.crowding_distance(pts, stable=True))


@pytest.mark.parametrize("n,m", [(40, 2), (500, 3), (2000, 2)])
def test_crowding_distance_ties(n, m):
    pts = np.random.default_rng(7 * n).integers(0, 6, size=(n, m)).astype(np.float64)
    assert np.array_equal(dev_crowding(pts), O.crowding_distance(pts, stable=True))


def test_crowding_distance_specials():
    rng = np.random.default_rng(2)
    pts = rng.random((64, 3))
    pts[3, 0] = np.nan
    pts[10, 1] = np.inf
    pts[11, 1] = -np.inf
    pts[20, 2] = -0.0
    pts[21, 2] = 0.0
    np.testing.assert_array_equal(dev_crowding(pts), O.crowding_distance(pts, stable=True))


def _key(pts, order):
    """(rank, -crowding) along an order, from the oracle's fnds and stable
    per-front crowding."""
    r = O.fnds(pts)
    crowd = np.zeros(len(pts))
    fo, fi = r["front_offsets"], r["front_items"]
    for f in range(len(fo) - 1):
        front = fi[fo[f]:fo[f + 1]].astype(np.int64)
        if len(front) > 1:
            crowd[front] = O.crowding_distance(pts[front], stable=True)
    return [(int(r["rank"][i]), -crowd[i]) for i in order]


@pytest.mark.parametrize("n,m", [(2, 2), (16, 2), (17, 3), (300, 2), (1200, 3)])
def test_sort_population_mo_continuous(n, m):
    pts = np.random.default_rng(100 + n).random((n, m))
    got = dev_sort(pts)
    assert np.array_equal(got, O.sort_population_mo(pts, stable=True))
    # the std::sort form agrees up to the order of equal (rank, crowding) pairs
    ref = O.sort_population_mo(pts)
    assert _key(pts, got) == _key(pts, ref)
    if n <= 16:  # libstdc++ sorts 16 elements or fewer by insertion: stable
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("n,m", [(50, 2), (400, 3)])
def test_sort_population_mo_ties(n, m):
    pts = np.random.default_rng(n).integers(0, 5, size=(n, m)).astype(np.float64)
    assert np.array_equal(dev_sort(pts), O.sort_population_mo(pts, stable=True))


@pytest.mark.parametrize("n,m", [(30, 2), (500, 3), (2048, 2)])
def test_select_best_N_mo(n, m):
    rng = np.random.default_rng(n + 1)
    pts = rng.random((n, m))
    tied = rng.integers(0, 6, size=(n, m)).astype(np.float64)
    for N in sorted({1, 2, n // 7, n // 3, n // 2, n - 1, n, n + 5}):
        assert np.array_equal(dev_select(pts, N), O.select_best_N_mo(pts, N, stable=True)), N
        assert np.array_equal(dev_select(tied, N), O.select_best_N_mo(tied, N, stable=True)), N
        assert np.array_equal(P.select_best_N_mo_host(pts, N), O.select_best_N_mo(pts, N, stable=True)), N


def test_sort_population_mo_ranking_workload():
    """The two sorting workloads of the benchmark: 8192 x 2 (distances to two
    centres, n = 30) and 16384 x 3 with 10% duplicated rows."""
    rng = np.random.default_rng(42)
    c = rng.uniform(0, 1, (2, 30))
    x = rng.uniform(0, 1, (8192, 30))
    f2 = O.bfe(P.MO_SPHERE, 30, x, centers=c, nobj=2)
    assert np.array_equal(dev_sort(f2), O.sort_population_mo(f2, stable=True))
    n = 16384
    objs = rng.uniform(0, 1, (n, 3))
    dup = n // 10
    objs[n - dup:] = objs[rng.integers(0, n - dup, dup)]
    assert np.array_equal(dev_sort(objs), O.sort_population_mo(objs, stable=True))
    for N in (100, n // 2):
        assert np.array_equal(dev_select(objs, N), O.select_best_N_mo(objs, N, stable=True))


def test_single_objective_fronts_of_one():
    # m = 1 with distinct values: every front has one element, so no crowding
    # is needed and no error is raised (multi_objective.cpp:201-203 example)
    pts = np.arange(11, dtype=np.float64).reshape(11, 1)
    assert list(map(int, dev_sort(pts))) == list(range(11))
    with pytest.raises(P.PghError, match="at least two objectives: 1 detected"):
        dev_sort(np.zeros((3, 1)))
