"""The autotuner's parameters change no result: every candidate of the front
peel's workgroup count and of the Lennard-Jones individuals per workgroup gives
bit-identical outputs, the tuners pick one of their candidates, and an
archipelago built after a change of the Lennard-Jones mapping evolves exactly
as one built with the default."""
import ctypes

import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"

KEYS = ("dom_count", "rank", "front_offsets", "front_items", "dom_offsets", "dom_items")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def lj_inputs(atoms, n, seed):
    nx = 3 * atoms - 6
    return O.batch_uniform(np.full(nx, -3.0), np.full(nx, 3.0), n, seed)


def lj_run(atoms, x, grad):
    n, nx = x.shape
    f = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    if grad:
        g = torch.empty((n, nx), dtype=torch.float64, device=DEV)
        P.batch_gradient(P.LENNARD_JONES, atoms, x, g, f)
        torch.cuda.synchronize()
        return f.cpu().numpy(), g.cpu().numpy()
    P.bfe_eval(P.LENNARD_JONES, atoms, x, f)
    torch.cuda.synchronize()
    return f.cpu().numpy(), None


@pytest.fixture
def lj_defaults():
    yield
    for atoms in (5, 13, 20, 37, 48):
        for grad in (0, 1):
            P.lj_set_groups(atoms, grad, 0)


# 37 and 48 atoms: the gradient's pair-once kernel, with and without padded
# block slots
@pytest.mark.parametrize("atoms,n", [(5, 301), (13, 1000), (37, 515), (48, 2049)])
@pytest.mark.parametrize("grad", [False, True])
def test_lj_groups_change_no_result(atoms, n, grad, lj_defaults):
    x = dev(lj_inputs(atoms, n, 3 + atoms))
    f0, g0 = lj_run(atoms, x, grad)
    tried = 0
    for groups in (1, 2, 3, 4, 8, 16, 64, 256):
        try:
            P.lj_set_groups(atoms, grad, groups)
        except P.PghError:
            continue  # beyond the LDS limit for this size
        f, g = lj_run(atoms, x, grad)
        assert np.array_equal(f, f0), groups
        if grad:
            assert np.array_equal(g, g0), groups
        tried += 1
    assert tried >= 3
    P.lj_set_groups(atoms, grad, 0)
    fo = torch.empty((n, 1), dtype=torch.float64, device=DEV)
    go = torch.empty((n, x.shape[1]), dtype=torch.float64, device=DEV) if grad else None
    chosen = P.lj_tune(atoms, x, fo, go, reps=3)
    torch.cuda.synchronize()
    assert chosen >= 1
    assert np.array_equal(fo.cpu().numpy(), f0)
    if grad:
        assert np.array_equal(go.cpu().numpy(), g0)
    f, g = lj_run(atoms, x, grad)  # the tuned mapping stays in use
    assert np.array_equal(f, f0)


def test_lj_set_groups_limits(lj_defaults):
    # 48 atoms take 16 lanes each (pair-once blocks): 16 individuals fill a
    # 256-lane workgroup; 40 atoms take the cyclic schedule's 32 lanes: 8
    with pytest.raises(P.PghError, match="at most 16 individuals per workgroup for 48 atoms"):
        P.lj_set_groups(48, 0, 17)
    with pytest.raises(P.PghError, match="at most 8 individuals per workgroup for 40 atoms"):
        P.lj_set_groups(40, 0, 9)
    with pytest.raises(P.PghError, match="atoms out of range"):
        P.lj_set_groups(2, 0, 1)
    x = dev(lj_inputs(13, 4, 1))
    u = P.make_udp(P.RASTRIGIN, 13)
    out = ctypes.c_uint32()
    with pytest.raises(P.PghError, match="needs a lennard_jones problem"):
        P.check(P.lib().pgh_lj_tune(ctypes.byref(u), x.data_ptr(), 4, x.data_ptr(), None, 1, None, ctypes.byref(out)))


def fnds_outputs(ws, n):
    ws.sizes()
    return ws.to_host(n)


@pytest.mark.parametrize("n,m,dup", [(2000, 2, 0.0), (8192, 2, 0.0), (16384, 3, 0.1), (32768, 2, 0.05)])
def test_peel_grid_changes_no_result(n, m, dup):
    rng = np.random.default_rng(n + m)
    objs = rng.uniform(0, 1, (n, m))
    k = int(n * dup)
    if k:
        objs[n - k:] = objs[rng.integers(0, n - k, k)]
    d = dev(objs)
    ref = O.fnds(objs)
    ws = P.Fnds()
    try:
        for blocks in (0, 1, 3, 8, 64, 256, 1024, 4096):
            ws.set_peel_blocks(blocks)
            ws.run(d)
            got = fnds_outputs(ws, n)
            for key in KEYS:
                assert np.array_equal(got[key], ref[key]), (blocks, key)
        ws.set_peel_blocks(0)
        chosen = ws.tune(d, reps=2)
        # the peel's range: at most 16 words (1024 points) per workgroup, at
        # most one workgroup per word and 256 in all
        words = (n + 63) // 64
        assert -(-words // 16) <= chosen <= min(words, 256)
        got = fnds_outputs(ws, n)
        for key in KEYS:
            assert np.array_equal(got[key], ref[key]), key
        with pytest.raises(P.PghError, match="at most 4096 peel workgroups"):
            ws.set_peel_blocks(4097)
    finally:
        ws.close()


def test_peel_tune_beyond_grid_limit_is_a_plain_run():
    n = 33000
    objs = np.random.default_rng(5).uniform(0, 1, (n, 2))
    ws = P.Fnds()
    try:
        assert ws.tune(dev(objs), reps=1) == 0
        got = fnds_outputs(ws, n)
        ref = O.fnds(objs)
        for key in ("rank", "front_offsets", "front_items"):
            assert np.array_equal(got[key], ref[key]), key
    finally:
        ws.close()


def test_archipelago_with_a_changed_lj_mapping(lj_defaults):
    rng = np.random.default_rng(8)
    islands = []
    for k, atoms in enumerate((5, 13, 20)):
        pop = int(rng.integers(20, 90))
        X0, F0 = O.population_init(P.LENNARD_JONES, atoms, pop, 70 + k)
        islands.append(dict(kind=P.LENNARD_JONES, dim=atoms, X0=X0, F0=F0, algo=P.pso_params(gen=6, seed=70 + k)))
    base = P.Archipelago(islands)
    base.evolve()
    want = [base.get(k) for k in range(len(islands))]
    base.close()
    for atoms in (5, 13, 20):
        P.lj_set_groups(atoms, 0, 2)
    for mode in (P.ARCHI_BATCHED, P.ARCHI_PER_ISLAND):
        a = P.Archipelago(islands)
        a.evolve(mode)
        for k in range(len(islands)):
            X, F = a.get(k)
            assert np.array_equal(X, want[k][0]) and np.array_equal(F, want[k][1]), (mode, k)
        a.close()
