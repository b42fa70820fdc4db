"""fast_non_dominated_sorting on the device: fronts (with their order),
domination lists, domination counts and ranks bit-exact against the oracle
(multi_objective.cpp:200-257)."""
import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"

KEYS = ("dom_count", "rank", "front_offsets", "front_items", "dom_offsets", "dom_items")


def gpu_fnds(objs):
    objs = np.ascontiguousarray(objs, np.float64)
    n = objs.shape[0]
    d = torch.from_numpy(objs if objs.size else np.zeros((n, 0))).to(DEV)
    ws = P.Fnds()
    try:
        ws.run(d)
        return ws.to_host(n)
    finally:
        ws.close()


def assert_same(objs):
    g = gpu_fnds(objs)
    r = O.fnds(objs)
    for k in KEYS:
        assert np.array_equal(g[k], r[k]), k


def test_reference_kats(kats):
    for k in kats["fnds"]:
        pts = np.array(k["points"], np.float64).reshape(len(k["points"]), -1)
        g = gpu_fnds(pts)
        fo, fi = g["front_offsets"], g["front_items"]
        fronts = [list(map(int, fi[fo[i]:fo[i + 1]])) for i in range(len(fo) - 1)]
        assert fronts == k["fronts"], k["src"]
        do, di = g["dom_offsets"], g["dom_items"]
        assert [list(map(int, di[do[i]:do[i + 1]])) for i in range(len(do) - 1)] == k["dom_list"], k["src"]
        assert list(map(int, g["dom_count"])) == k["dom_count"]
        assert list(map(int, g["rank"])) == k["rank"]


def test_too_few_points():
    with pytest.raises(P.PghError, match="At least two points are needed for fast_non_dominated_sorting: 1 detected."):
        gpu_fnds(np.zeros((1, 3)))


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 300, 1025])
def test_random_with_ties(m, n):
    rng = np.random.default_rng(1000 * m + n)
    objs = rng.integers(0, 5, size=(n, m)).astype(np.float64)
    assert_same(objs)


def test_every_size_to_300():
    """Every population size from 2 to 300 and a few around the peel's
    thresholds (the bit matrix's 32-bit words, the column partition), with
    ties (integer objectives) and without: all six outputs bit for bit."""
    sizes = list(range(2, 301)) + [511, 512, 513, 1023, 1024, 1025, 4095, 4096, 4097]
    for n in sizes:
        m = 2 + n % 3
        rng = np.random.default_rng(n)
        assert_same(rng.integers(0, 6, size=(n, m)).astype(np.float64))
        if n % 7 == 0 or n > 300:
            assert_same(rng.uniform(0, 1, (n, m)))


@pytest.mark.parametrize("n,m", [(500, 2), (777, 3)])
def test_continuous(n, m):
    rng = np.random.default_rng(n)
    assert_same(rng.uniform(0, 1, (n, m)))


def test_specials():
    rng = np.random.default_rng(3)
    vals = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, -1.0, 1e-300, -1e-300, 5e-324])
    objs = rng.choice(vals, size=(400, 3))
    assert_same(objs)


def test_anticorrelated_single_front():
    x = np.linspace(0, 1, 3000)
    assert_same(np.stack([x, 1 - x], axis=1))


def test_chain():
    # every point dominates the next: n fronts of one element
    x = np.arange(700, dtype=np.float64)
    assert_same(np.stack([x, x, x], axis=1))


def test_ranking_workload_mo_sphere():
    """8192 x 2 objectives of f_k(x) = ||x - c_k||^2 (n = 30) computed on the
    device, then sorted on the device."""
    rng = np.random.default_rng(42)
    c = rng.uniform(0, 1, (2, 30))
    x = rng.uniform(0, 1, (8192, 30))
    xd = torch.from_numpy(x).to(DEV)
    f = torch.empty((8192, 2), dtype=torch.float64, device=DEV)
    P.bfe_eval(P.MO_SPHERE, 30, xd, f, centers=torch.from_numpy(c).to(DEV), nobj=2)
    ws = P.Fnds()
    ws.run(f)
    g = ws.to_host(8192)
    ws.close()
    objs = O.bfe(P.MO_SPHERE, 30, x, centers=c, nobj=2)
    assert np.array_equal(f.cpu().numpy(), objs)
    r = O.fnds(objs)
    for k in KEYS:
        assert np.array_equal(g[k], r[k]), k


def test_ranking_workload_duplicates():
    """16384 x 3 uniform objectives, the last 10% exact copies of earlier rows."""
    rng = np.random.default_rng(42)
    n = 16384
    objs = rng.uniform(0, 1, (n, 3))
    dup = n // 10
    objs[n - dup:] = objs[rng.integers(0, n - dup, dup)]
    assert_same(objs)


def layered(n, layers, delta, seed):
    """n points on `layers` parallel anti-diagonals x + y = c_L (c_L = L * delta),
    rows shuffled: a point dominates the points of later layers whose x lies
    within (L' - L) * delta above its own, so fronts are wide and each element
    has few dominators at varied positions."""
    rng = np.random.default_rng(seed)
    lay = np.arange(n) % layers
    x = rng.uniform(0, 1, n)
    objs = np.stack([x, lay * delta - x], axis=1)
    return objs[rng.permutation(n)]


def test_cooperative_peel_at_its_size_limit():
    # n = 32768 is the largest size the grid-wide peel takes (16-bit positions
    # and indices in its sort keys); one level completes thousands of elements
    assert_same(layered(32768, 8, 0.002, 5))


def test_cooperative_peel_one_dominator_of_all():
    # F1 = every other point, all completed by the one member of F0
    x = np.linspace(0, 1, 20000)
    objs = np.stack([x, 1 - x], axis=1)
    objs[7777] = (-1.0, -1.0)
    assert_same(objs)


def test_single_workgroup_peel_beyond_cooperative_limit():
    assert_same(layered(40000, 20, 0.0005, 6))
